// Device implementations of the UMANS cost functions, optimisation methods, contact
// forces and integration for one agent (one lane). Each function follows the
// reference file cited beside it; numerics follow device_math.h.
#pragma once
#include "device_math.h"
#include "step_types.h"

namespace umans {

// --------------------------------------------------------------------- neighbours
// One agent's neighbour list: entries (sorted index | segment << 28) in LDS with a
// fixed stride, plus the cell-sorted agent arrays they index. A neighbour is seen
// as the reference's PhantomAgent (worldBase.h:40-79): its position shifted by minus
// the query displacement, its float squared distance to the agent's undisplaced
// position, its live velocity, radius and preferred velocity.
struct Nb {
	V2 pos, vel;
	float radius, dsq;
	int j;
};

struct NbList {
	const unsigned* list;
	int stride;
	int count;
	const float2* pos;
	const float2* vel;
	const float4* prm;
	const float2* pref;
	float W, H;
	V2 self;
	// neighbouring obstacle edges (NeighborList::second): (ax, ay, bx, by), already moved
	// back by their query displacement; only the generic class collects them
	const float4* obs;
	int n_obs;

	UD unsigned raw(int e) const { return list[e * stride]; }
	UD Nb get(int e) const {
		const unsigned u = raw(e);
		const int j = (int)(u & kIndexMask);
		const unsigned seg = u >> kSegShift;
		const float2 pj = pos[j];
		// positionOffset = -displacement (worldToric.cpp:36,47), added in float; selects,
		// not branches (-0 adds nothing, whatever the sign of the coordinate)
		const float ox = seg == 1 ? -W : (seg == 2 ? W : -0.0f);
		const float oy = seg == 3 ? -H : (seg == 4 ? H : -0.0f);
		Nb n;
		n.pos = V2{pj.x + ox, pj.y + oy};
		n.vel = v2(vel[j]);
		n.radius = prm[j].x;
		n.dsq = sqr_mag(self - n.pos);       // distanceSquared, vector2D.h:157-160
		n.j = j;
		return n;
	}
	UD V2 pref_of(int j) const { return v2(pref[j]); }
};

// The agent being updated.
struct Self {
	V2 pos, vel, pref;
	float radius, pref_speed, max_speed, max_accel;
	float dt, relaxation;
};

// ------------------------------------------------------------- shared numerics
// CostFunction::SolveQuadraticEquation (costFunction.cpp:173-191), incl. the D == 0
// branch that evaluates (-b / 2) * a.
UD int solve_quadratic(float a, float b, float c, float& r1, float& r2) {
	const float D = b * b - 4.0f * a * c;
	if (D < 0.0f) return 0;
	if (D == 0.0f) { r1 = -b / 2.0f * a; return 1; }
	const float m = -b / (2.0f * a);
	const float s = sqrtf(D) / (2.0f * a);
	r1 = m + s;
	r2 = m - s;
	return 2;
}
// CostFunction::ComputeTimeToCollision (costFunction.cpp:193-227)
UD float time_to_collision(V2 p1, V2 v1, float r1, V2 p2, V2 v2_, float r2) {
	const V2 P = p1 - p2;
	const float R = r1 + r2;
	const float RSq = R * R;
	if (sqr_mag(P) <= RSq) return 0.0f;
	const V2 V = v1 - v2_;
	const float a = dot(V, V);
	const float b = 2.0f * dot(P, V);
	const float c = dot(P, P) - RSq;
	float t1 = kMaxFloat, t2 = kMaxFloat;
	solve_quadratic(a, b, c, t1, t2);
	if (t1 < 0.0f) t1 = kMaxFloat;
	if (t2 < 0.0f) t2 = kMaxFloat;
	return smin(t1, t2);
}
// ------------------------------------------------------------ obstacle edges
// nearestPointOnLine (vector2D.h:274-296); on_seg reports 0 <= k <= 1 of the unclamped k
UD V2 nearest_point(V2 pt, V2 la, V2 lb, bool on_segment, bool* on_seg = nullptr) {
	if (la.x == lb.x && la.y == lb.y) return la;
	const V2 line = la - lb;
	float k = dot(pt - lb, line) / sqr_mag(line);
	if (on_segment) {
		if (k > 1.0f) k = 1.0f;
		else if (k < 0.0f) k = 0.0f;
	}
	if (on_seg) *on_seg = (k >= 0.0f && k <= 1.0f);
	return (la * k) + (lb * (1.0f - k));
}
// distanceToLineSquared (vector2D.h:303-306)
UD float dist_line_sq(V2 pt, V2 la, V2 lb, bool on_segment) {
	return sqr_mag(pt - nearest_point(pt, la, lb, on_segment));
}
UD V2 seg_a(float4 e) { return V2{e.x, e.y}; }
UD V2 seg_b(float4 e) { return V2{e.z, e.w}; }

// CostFunction::ComputeTimeToCollision_LineSegmentInterior (costFunction.cpp:293-325)
UD float ttc_segment_interior(V2 pos, V2 vel, float radius, V2 sa, V2 sb) {
	if (dist_line_sq(pos, sa, sb, true) <= radius * radius) return 0.0f;
	V2 X;
	if (!line_intersection(pos, pos + vel, sa, sb, X)) return kMaxFloat;
	if (dot(vel, X - pos) < 0.0f) return kMaxFloat;
	bool on = false;
	nearest_point(X, sa, sb, false, &on);
	if (!on) return kMaxFloat;
	const float speed = mag(vel);
	const float ttc = mag(X - pos) / speed;
	const float line_angle = angle_between(vel, sb - sa);
	const float sub = radius / sinf(line_angle);
	const float tsub = sub / speed;
	return smax(0.0f, ttc - tsub);
}
// CostFunction::ComputeTimeToCollision_LineSegment (costFunction.cpp:270-291)
UD float ttc_segment(V2 pos, V2 vel, float radius, V2 sa, V2 sb) {
	const float t1 = time_to_collision(pos, vel, radius, sa, V2{0.0f, 0.0f}, 0.0f);
	if (t1 == 0.0f) return 0.0f;
	if (sa.x == sb.x && sa.y == sb.y) return t1;
	const float t2 = time_to_collision(pos, vel, radius, sb, V2{0.0f, 0.0f}, 0.0f);
	if (t2 == 0.0f) return 0.0f;
	return smin(smin(t1, t2), ttc_segment_interior(pos, vel, radius, sa, sb));
}

// CostFunction::ComputeTimeToFirstCollision (costFunction.cpp:229-268)
UD float time_to_first_collision(const NbList& L, V2 pos, V2 vel, float radius, float maxDist, bool ignoreCurrent) {
	float minTTC = kMaxFloat;
	const float maxDistSq = maxDist * maxDist;
	for (int e = 0; e < L.count; ++e) {
		const Nb n = L.get(e);
		if (n.dsq > maxDistSq) continue;
		const float ttc = time_to_collision(pos, vel, radius, n.pos, n.vel, n.radius);
		if (ignoreCurrent && ttc == 0.0f) continue;
		if (ttc < minTTC) minTTC = ttc;
	}
	for (int e = 0; e < L.n_obs; ++e) {
		const float4 sg = L.obs[e];
		if (dist_line_sq(pos, seg_a(sg), seg_b(sg), true) > maxDistSq) continue;
		const float ttc = ttc_segment(pos, vel, radius, seg_a(sg), seg_b(sg));
		if (ignoreCurrent && ttc == 0.0f) continue;
		if (ttc < minTTC) minTTC = ttc;
	}
	return minTTC;
}

// --------------------------------------------------------------- force models
// PowerLaw::ComputeAgentInteractionForce (PowerLaw.cpp:28-66)
UD V2 power_law_pair(const float* prm, const Self& a, const Nb& o) {
	const float k = prm[1], tau0 = prm[2];
	const float R = o.radius + a.radius;
	const V2 x = a.pos - o.pos;
	const V2 v = a.vel - o.vel;
	const float aa = dot(v, v);
	const float b = -dot(x, v);
	const float c = dot(x, x) - R * R;
	if (c <= 0.0f) return V2{0.0f, 0.0f};
	const float d = b * b - aa * c;
	if (d <= 0.0f) return V2{0.0f, 0.0f};
	const float sqrtD = sqrtf(d);
	const float tau = (b - sqrtD) / aa;
	if (tau < 0.001f) return V2{0.0f, 0.0f};
	if (tau > tau0) return V2{0.0f, 0.0f};
	const float comp1 = -k * expf(-tau / tau0) * (2.0f / tau + 1.0f / tau0) / (aa * tau * tau);
	const V2 comp2 = v - (aa * x + b * v) / sqrtD;
	return comp1 * comp2;
}

// angle(a, b) < (float)(100 deg) as the reference evaluates it (vector2D.h:177-190, a
// double acos cast to float), decided by comparing the cosine with view_cos, the
// largest double whose float-rounded acos is still >= 100 degrees (computed on the host
// with the same libm): no transcendental on the device, and the reference's rounding.
UD bool in_view(V2 a, V2 b, double view_cos) {
	const double lengths = (double)(mag(a) * mag(b));
	if (lengths == 0.0) return true;
	const double frac = (double)dot(a, b) / lengths;
	if (frac < -1.0 || frac > 1.0) return true;
	return frac > view_cos;
}

// SocialForcesAvoidance::ComputeAgentInteractionForce (SocialForcesAvoidance.cpp:26-110)
UD V2 social_force_pair(const float* prm, const Self& a, const Nb& o, V2 o_pref, double view_cos) {
	const float sdt = prm[1], V0 = prm[2], sigma = prm[3];
	const V2 R = a.pos - o.pos;
	const float magR = mag(R);
	if (magR <= a.radius + o.radius) return V2{0.0f, 0.0f};
	const float vb = mag(o.vel);
	const V2 Eb = normalized(o_pref);
	const V2 V = (vb * sdt) * Eb;
	const float magV = mag(V);
	const V2 R2 = R - V;
	const float magR2 = mag(R2);
	const double s1 = (double)(magR + magR2), s2 = (double)(vb * sdt);
	const float bSq2 = (float)(s1 * s1 - s2 * s2);        // pow(float, 2) in double
	if (bSq2 <= 0.0f) return V2{0.0f, 0.0f};
	const float b = 0.5f * sqrtf(bSq2);
	if (b < 0.001f) return V2{0.0f, 0.0f};
	const float d_magR2 = 1.0f / (2.0f * magR2) * (2.0f * magR - 2.0f * magV * cos_angle(R, V));
	const float d_bSq2 = 2.0f * (magR + magR2) * (1.0f + d_magR2);
	const float d_b = (1.0f / (8.0f * b)) * d_bSq2;
	const float d_V = -V0 / sigma * expf(-b / sigma) * d_b;
	const V2 F = (-R) * d_V;
	const float scale = in_view(a.vel, -R, view_cos) ? 1.0f : 0.5f;
	return scale * F;
}

// SocialForcesAvoidance::ComputeObstacleInteractionForce (SocialForcesAvoidance.cpp:112-140).
// U0 keeps its default 10: parseParameters reads the "U0" attribute into V0
// (SocialForcesAvoidance.cpp:142-150); R is parameter slot 4.
UD V2 social_force_obstacle(const float* prm, const Self& a, V2 sa, V2 sb) {
	const float U0 = 10.0f, R = prm[4];
	const V2 diff = a.pos - nearest_point(a.pos, sa, sb, true);
	const float dist = mag(diff);
	if (dist < 0.001f) return V2{0.0f, 0.0f};
	const float d_U = -U0 / R * expf(-dist / R);
	return (-diff) * d_U;
}

// ObjectInteractionForces::ComputeForce (ObjectInteractionForces.cpp:26-48) for
// PowerLaw / SocialForcesAvoidance, GoalReachingForce::ComputeForce
// (GoalReachingForce.cpp:26-29). PowerLaw's obstacle force is zero (PowerLaw.cpp:68-72).
UD V2 force_of(const DevPolicy& P, int c, const Self& a, const NbList& L) {
	const int kind = P.kind[c];
	if (kind == KIND_GOAL_REACHING)
		return (a.pref - a.vel) / smax(a.relaxation, a.dt);
	const float rangeSq = P.range[c] * P.range[c];
	V2 F{0.0f, 0.0f};
	for (int e = 0; e < L.count; ++e) {
		const Nb n = L.get(e);
		if (n.dsq <= rangeSq) {
			const V2 f = (kind == KIND_POWER_LAW) ? power_law_pair(P.params[c], a, n)
			                                      : social_force_pair(P.params[c], a, n, L.pref_of(n.j), P.view_cos);
			F = F + f;
		}
	}
	V2 Fo{0.0f, 0.0f};
	for (int e = 0; e < L.n_obs; ++e) {
		const float4 sg = L.obs[e];
		if (dist_line_sq(a.pos, seg_a(sg), seg_b(sg), true) <= rangeSq)
			Fo = Fo + (kind == KIND_SOCIAL_FORCES ? social_force_obstacle(P.params[c], a, seg_a(sg), seg_b(sg))
			                                      : V2{0.0f, 0.0f});
	}
	return F + Fo;
}
// ForceBasedFunction::ComputeTargetVelocity (ForceBasedFunction.cpp:49-52)
UD V2 target_velocity(const DevPolicy& P, int c, const Self& a, const NbList& L) {
	const float mass = P.params[c][0];
	return a.vel + force_of(P, c, a, L) / mass * a.dt;
}

UD bool is_force_based(int kind) {
	return kind == KIND_GOAL_REACHING || kind == KIND_POWER_LAW || kind == KIND_SOCIAL_FORCES;
}

// ----------------------------------------------------------------------- ORCA
struct Line { V2 point, dir; };
// A lane's half-planes: contiguous in private memory (stride 1) or interleaved with the
// other lanes of the block in LDS (stride = block size).
struct Lines {
	Line* p;
	int s;
	UD Line& operator[](int i) const { return p[i * s]; }
};
constexpr float kOrcaEps = 0.00001f;   // ORCASolver.h:62

// The half-plane of one neighbour (ORCASolver.cpp:104-168), including the collision
// branch that divides by the squared length (ORCASolver.cpp:161-165).
UD Line orca_line(const Self& a, const Nb& o, float tau) {
	const V2 relPos = o.pos - a.pos;
	const V2 relVel = a.vel - o.vel;
	const float distSq = sqr_mag(relPos);
	const float cr = a.radius + o.radius;
	const float crSq = cr * cr;
	Line line;
	V2 u;
	if (distSq > crSq) {
		const V2 w = relVel - (relPos / tau);
		const float wLenSq = sqr_mag(w);
		const float dp1 = dot(w, relPos);
		if (dp1 < 0.0f && dp1 * dp1 > crSq * wLenSq) {
			const float wLen = sqrtf(wLenSq);
			const V2 unitW = w / wLen;
			line.dir = V2{unitW.y, -unitW.x};
			u = unitW * (cr / tau - wLen);
		} else {
			const float leg = sqrtf(distSq - crSq);
			if (det(relPos, w) > 0.0f)
				line.dir = V2{relPos.x * leg - relPos.y * cr, relPos.x * cr + relPos.y * leg} / distSq;
			else
				line.dir = V2{relPos.x * leg + relPos.y * cr, -relPos.x * cr + relPos.y * leg} * -1.0f / distSq;
			const float dp2 = dot(relVel, line.dir);
			u = (line.dir * dp2) - relVel;
		}
	} else {
		const V2 w = relVel - (relPos / a.dt);
		const float wLen = sqr_mag(w);
		const V2 unitW = w / wLen;
		line.dir = V2{unitW.y, -unitW.x};
		u = unitW * (cr / a.dt - wLen);
	}
	line.point = a.vel + (u * 0.5f);
	return line;
}
// Solver::createAgentOrcaLines (ORCASolver.cpp:96-171): the neighbours within range, in
// list order.
UD int orca_lines(const float* prm, float range, const Self& a, const NbList& L, Lines lines) {
	const float tau = prm[0];
	const float maxDistSq = range * range;
	int n = 0;
	for (int e = 0; e < L.count; ++e) {
		const Nb o = L.get(e);
		if (o.dsq > maxDistSq) continue;
		lines[n++] = orca_line(a, o, tau);
	}
	return n;
}
// linearProgram1 (ORCASolver.cpp:173-244)
UD bool lp1(Lines L, int no, float radius, V2 opt, bool dirOpt, V2& res) {
	const float dp = dot(L[no].point, L[no].dir);
	const float disc = dp * dp + radius * radius - sqr_mag(L[no].point);
	if (disc < 0.0f) return false;
	const float sd = sqrtf(disc);
	float tl = -dp - sd, tr = -dp + sd;
	for (int i = 0; i < no; ++i) {
		const float den = det(L[no].dir, L[i].dir);
		const float num = det(L[i].dir, L[no].point - L[i].point);
		if (fabsf(den) <= kOrcaEps) {
			if (num < 0.0f) return false;
			continue;
		}
		const float t = num / den;
		if (den >= 0.0f) tr = smin(tr, t);
		else tl = smax(tl, t);
		if (tl > tr) return false;
	}
	if (dirOpt) {
		res = dot(opt, L[no].dir) > 0.0f ? L[no].point + (L[no].dir * tr) : L[no].point + (L[no].dir * tl);
	} else {
		const float t = dot(L[no].dir, opt - L[no].point);
		if (t < tl) res = L[no].point + (L[no].dir * tl);
		else if (t > tr) res = L[no].point + (L[no].dir * tr);
		else res = L[no].point + (L[no].dir * t);
	}
	return true;
}
// linearProgram2 (ORCASolver.cpp:246-277)
UD int lp2(Lines L, int n, float radius, V2 opt, bool dirOpt, V2& res) {
	if (dirOpt) res = opt * radius;
	else if (sqr_mag(opt) > radius * radius) res = normalized(opt) * radius;
	else res = opt;
	for (int i = 0; i < n; ++i) {
		if (det(L[i].dir, L[i].point - res) > 0.0f) {
			const V2 tmp = res;
			if (!lp1(L, i, radius, opt, dirOpt, res)) { res = tmp; return i; }
		}
	}
	return n;
}
// linearProgram3 (ORCASolver.cpp:279-326) without obstacle lines
UD void lp3(Lines L, int n, int begin, float radius, V2& res, Lines proj) {
	float distance = 0.0f;
	for (int i = begin; i < n; ++i) {
		if (det(L[i].dir, L[i].point - res) > distance) {
			int np = 0;
			for (int j = 0; j < i; ++j) {
				Line line;
				const float dt = det(L[i].dir, L[j].dir);
				if (fabsf(dt) <= kOrcaEps) {
					if (dot(L[i].dir, L[j].dir) > 0.0f) continue;
					line.point = (L[i].point + L[j].point) * 0.5f;
				} else {
					line.point = L[i].point + L[i].dir * (det(L[j].dir, L[i].point - L[j].point) / dt);
				}
				line.dir = normalized(L[j].dir - L[i].dir);
				proj[np++] = line;
			}
			const V2 tmp = res;
			if (lp2(proj, np, radius, V2{-L[i].dir.y, L[i].dir.x}, true, res) < np) res = tmp;
			distance = det(L[i].dir, L[i].point - res);
		}
	}
}
// Solver::solveOrcaProgram (ORCASolver.cpp:61-94)
UD V2 orca_solve(Lines L, int n, const Self& a, Lines proj) {
	V2 res;
	const int fail = lp2(L, n, a.max_speed, a.pref, false, res);
	if (fail != n) lp3(L, n, fail, a.max_speed, res, proj);
	return res;
}
// The same solver with the half-planes in registers (lists of at most 16 entries): slot
// e holds neighbour e's half-plane when bit e of `valid` is set, so every index is a
// compile-time constant once the loops are unrolled, and the lines never touch memory.
// The valid slots in slot order are the lines of orca_lines in list order, so each loop
// below visits the lines the loop it mirrors visits, in the same order.
template <int K>
struct RegLines {
	Line l[K];
	unsigned valid;
	UD bool has(int e) const { return (valid >> e) & 1u; }
};
template <int K>
UD void orca_lines_reg(const float* prm, float range, const Self& a, const NbList& L, RegLines<K>& R) {
	const float tau = prm[0];
	const float maxDistSq = range * range;
	R.valid = 0u;
#pragma unroll
	for (int e = 0; e < K; ++e) {
		R.l[e] = Line{V2{0.0f, 0.0f}, V2{0.0f, 0.0f}};
		if (e < L.count) {
			const Nb o = L.get(e);
			if (!(o.dsq > maxDistSq)) {
				R.l[e] = orca_line(a, o, tau);
				R.valid |= 1u << e;
			}
		}
	}
}
// lp1 over the lines before slot `no` (a constant after unrolling)
template <int K>
UD bool lp1_reg(const RegLines<K>& R, int no, float radius, V2 opt, bool dirOpt, V2& res) {
	const Line Ln = R.l[no];
	const float dp = dot(Ln.point, Ln.dir);
	const float disc = dp * dp + radius * radius - sqr_mag(Ln.point);
	if (disc < 0.0f) return false;
	const float sd = sqrtf(disc);
	float tl = -dp - sd, tr = -dp + sd;
	bool ok = true;
#pragma unroll
	for (int i = 0; i < K; ++i) {
		if (i < no && R.has(i) && ok) {
			const float den = det(Ln.dir, R.l[i].dir);
			const float num = det(R.l[i].dir, Ln.point - R.l[i].point);
			if (fabsf(den) <= kOrcaEps) {
				if (num < 0.0f) ok = false;
			} else {
				const float t = num / den;
				if (den >= 0.0f) tr = smin(tr, t);
				else tl = smax(tl, t);
				if (tl > tr) ok = false;
			}
		}
	}
	if (!ok) return false;
	if (dirOpt) {
		res = dot(opt, Ln.dir) > 0.0f ? Ln.point + (Ln.dir * tr) : Ln.point + (Ln.dir * tl);
	} else {
		const float t = dot(Ln.dir, opt - Ln.point);
		if (t < tl) res = Ln.point + (Ln.dir * tl);
		else if (t > tr) res = Ln.point + (Ln.dir * tr);
		else res = Ln.point + (Ln.dir * t);
	}
	return true;
}
// lp2: returns the slot of the line that failed, or K
template <int K>
UD int lp2_reg(const RegLines<K>& R, float radius, V2 opt, bool dirOpt, V2& res) {
	if (dirOpt) res = opt * radius;
	else if (sqr_mag(opt) > radius * radius) res = normalized(opt) * radius;
	else res = opt;
	int fail = K;
#pragma unroll
	for (int i = 0; i < K; ++i) {
		if (fail == K && R.has(i) && det(R.l[i].dir, R.l[i].point - res) > 0.0f) {
			const V2 tmp = res;
			if (!lp1_reg<K>(R, i, radius, opt, dirOpt, res)) { res = tmp; fail = i; }
		}
	}
	return fail;
}
// lp2 over projected lines in memory, out of line: lp3 calls it from each unrolled slot
__device__ __noinline__ int lp2_mem(Lines L, int n, float radius, V2 opt, V2* res) {
	V2 r = *res;
	const int f = lp2(L, n, radius, opt, true, r);
	*res = r;
	return f;
}
// lp3 from slot `begin`; the projected lines of a slot go to private memory (rare path)
template <int K>
UD void lp3_reg(const RegLines<K>& R, int begin, float radius, V2& res, Lines proj) {
	float distance = 0.0f;
#pragma unroll
	for (int i = 0; i < K; ++i) {
		if (i >= begin && R.has(i) && det(R.l[i].dir, R.l[i].point - res) > distance) {
			int np = 0;
#pragma unroll
			for (int j = 0; j < i; ++j) {
				if (!R.has(j)) continue;
				Line line;
				const float dt = det(R.l[i].dir, R.l[j].dir);
				if (fabsf(dt) <= kOrcaEps) {
					if (dot(R.l[i].dir, R.l[j].dir) > 0.0f) continue;
					line.point = (R.l[i].point + R.l[j].point) * 0.5f;
				} else {
					line.point = R.l[i].point + R.l[i].dir * (det(R.l[j].dir, R.l[i].point - R.l[j].point) / dt);
				}
				line.dir = normalized(R.l[j].dir - R.l[i].dir);
				proj[np++] = line;
			}
			V2 r = res;
			if (lp2_mem(proj, np, radius, V2{-R.l[i].dir.y, R.l[i].dir.x}, &r) >= np) res = r;
			distance = det(R.l[i].dir, R.l[i].point - res);
		}
	}
}
template <int K>
UD V2 orca_solve_reg(const RegLines<K>& R, const Self& a, Lines proj) {
	V2 res;
	const int fail = lp2_reg<K>(R, a.max_speed, a.pref, false, res);
	if (fail < K) lp3_reg<K>(R, fail, a.max_speed, res, proj);
	return res;
}
// ORCA::GetCost (ORCA.cpp:49-86)
UD float orca_cost(Lines L, int n, const Self& a, V2 v) {
	float maxD = -kMaxFloat;
	for (int i = 0; i < n; ++i) {
		const V2 pmv = L[i].point - v;
		maxD = smax(maxD, L[i].dir.x * pmv.y - L[i].dir.y * pmv.x);
	}
	if (maxD <= 0.0f) return mag(v - a.pref);
	return maxD + 2.0f * a.max_speed;
}

// ----------------------------------------------------------- sampled cost models
// Karamouzas::GetCost with getMinSpeed / getMaxSpeed / getMaxDeviationAngle
// (Karamouzas.cpp:28-113). tc_max stays 8 whatever t_max is (Karamouzas.h:53-58).
// ttcp = TTC at the preferred velocity, which does not depend on the candidate.
UD float karamouzas_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v, float ttcp) {
	const float alpha = prm[0], beta = prm[1], gamma = prm[2], delta = prm[3], t_max = prm[4];
	const float tc_min = 2.5f, tc_mid = 6.0f, tc_max = 8.0f;
	const float d_min = 0.05f, d_mid = (float)(kPI / 6.0f), d_max = (float)(kPI / 2.0f);
	const float speed = mag(v);
	const float maxSpeed = a.max_speed;
	float maxDev;
	if (ttcp <= 0.0f || ttcp >= tc_max) maxDev = d_min;
	else if (ttcp < tc_min) {
		const double q = (double)((tc_min - ttcp) / tc_min);
		maxDev = (float)((double)d_mid + q * q * (double)(d_max - d_mid));
	} else if (ttcp < tc_mid) maxDev = d_mid;
	else maxDev = d_min + d_mid - (tc_max - ttcp) / (tc_max - tc_mid);
	if (angle_between(v, a.pref) > maxDev) return kMaxFloat;
	const float pref = a.pref_speed, dS = fabsf(maxSpeed - pref), eps = 0.05f;
	float minS, maxS;
	if (ttcp <= 0.0f || ttcp >= tc_max) { minS = pref - eps; maxS = pref + eps; }
	else if (ttcp > 0.0f && ttcp < tc_min) { minS = 0.0f; maxS = maxSpeed + eps; }
	else { minS = pref - dS - eps; maxS = pref + dS + eps; }
	if (speed < minS || speed > maxS) return kMaxFloat;
	const float ttc = time_to_first_collision(L, a.pos, v, a.radius, range, true);
	const float A = alpha * (1.0f - cos_angle(v, a.pref) / 2.0f);
	const float B = beta * fabsf(speed - mag(a.vel)) / maxSpeed;
	const float C = gamma * mag(v - a.pref) / (2.0f * maxSpeed);
	const float D = delta * smax(0.0f, t_max - ttc) / t_max;
	return A + B + C + D;
}
// RVO::GetCost (RVO.cpp:28-44)
UD float rvo_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	if (mag(v) > a.max_speed) return kMaxFloat;
	const V2 rv = 2.0f * v - a.vel;
	const V2 vd = a.pref - v;
	const float ttc = time_to_first_collision(L, a.pos, rv, a.radius, range, true);
	return prm[0] / ttc + mag(vd);
}
// Moussaid::GetCost (Moussaid.cpp:28-62)
UD float moussaid_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	const float d_max = prm[0];
	const float speed = mag(v);
	const float prefSpeed = a.pref_speed;
	const V2 dir = normalized(v);
	const V2 vp = dir * prefSpeed;
	const float ttc = time_to_first_collision(L, a.pos, vp, a.radius, range, false);
	float DC = (ttc == kMaxFloat ? kMaxFloat : ttc * prefSpeed);
	if (DC > d_max) DC = d_max;
	float K = d_max * d_max + DC * DC - 2.0f * d_max * DC * cos_angle(dir, a.pref);
	K += 1.0f;
	const float S = smin(prefSpeed, DC / a.relaxation);
	const double q = (double)((S - speed) / S);
	return (float)((double)K * (1.0 + q * q));
}
// VanToll::GetCost (VanToll.cpp:29-64)
UD float vantoll_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	const float maxd = prm[0];
	const float prefSpeed = mag(a.pref);
	const float speed = mag(v);
	const float ttc = time_to_first_collision(L, a.pos, v, a.radius, range, false);
	float dist = (ttc == kMaxFloat ? kMaxFloat : ttc * speed);
	if (dist > maxd) dist = maxd;
	const float devPA = angle_between(v, a.pref);
	const float devPS = fabsf(speed - prefSpeed) / prefSpeed;
	const float devCA = angle_between(v, a.vel);
	return 1.0f * (maxd - dist) + 1.0f * devPA + 1.0f * devPS + 1.0f * devCA;
}
// TtcaDca helpers (TtcaDca.cpp:34-48) and CostFunction::ComputeTimeAndDistanceToClosestApproach
// (costFunction.cpp:354-380)
UD float ttca_pair_cost(float sT, float sD, float ttca, float dca) {
	const double x = (double)(ttca / sT), y = (double)(dca / sD);
	return (float)exp(-0.5 * (x * x + y * y));
}
UD void ttca_dca(V2 p1, V2 v1, float r1, V2 p2, V2 v2_, float r2, float& ttca, float& dca) {
	const V2 dpc = p2 - p1;
	const V2 dvc = v2_ - v1;
	const float dpMag = mag(dpc);
	const V2 dp = dpc / dpMag * (dpMag - r1 - r2);
	const V2 nrm{-dpc.y, dpc.x};
	const V2 dv = dvc + nrm * dot(nrm, dvc);
	ttca = (sqr_mag(dv) == 0.0f ? 0.0f : -dot(dp, dv) / sqr_mag(dv));
	if (ttca < 0.0f) ttca = 0.0f;
	dca = mag(dp + dv * ttca);
	if (dca < 0.0f) dca = 0.0f;
}
// TtcaDca::GetCost (TtcaDca.cpp:62-129)
UD float ttcadca_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	const float sA = prm[0], sS = prm[1], sT = prm[2], sD = prm[3];
	const float rangeSq = range * range;
	const float alpha = angle_between(v, a.pref);
	const float ds = mag(v) - a.pref_speed;
	const double qa = (double)(alpha / sA), qs = (double)(ds / sS);
	const float Ca = (float)exp(-0.5 * (qa * qa));
	const float Cs = (float)exp(-0.5 * (qs * qs));
	const float movement = 1.0f - (Ca + Cs) / 2.0f;
	float obst = 0.0f, scaleSum = 0.0f;
	const float half = (float)(kPI / 2.0);
	for (int e = 0; e < L.count; ++e) {
		const Nb o = L.get(e);
		if (o.dsq >= rangeSq) continue;
		const V2 relPos = o.pos - a.pos;
		if (angle_between(relPos, a.vel) > half) continue;
		float ttca, dca;
		ttca_dca(a.pos, v, a.radius, o.pos, o.vel, o.radius, ttca, dca);
		const float distance = mag(relPos) - a.radius - o.radius;
		const float scale = 1.0f / (distance * distance);
		scaleSum += scale;
		obst += ttca_pair_cost(sT, sD, ttca, dca) * scale;
	}
	if (scaleSum > 0.0f) obst /= scaleSum;
	return obst + movement;
}
// TtcaDca::GetGradient (TtcaDca.cpp:131-262) and RotateGradientToEuclideanCoordinates
// (costFunction.cpp:30-34)
UD V2 ttcadca_gradient(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	const float sA = prm[0], sS = prm[1], sT = prm[2], sD = prm[3];
	const float eps = 0.001f;
	const float rangeSq = range * range;
	const float half = (float)(kPI / 2.0);
	float GTh = 0.0f, GS = 0.0f, tot = 0.0f;
	for (int e = 0; e < L.count; ++e) {
		const Nb o = L.get(e);
		if (o.dsq > rangeSq) continue;
		const V2 dpc = o.pos - a.pos;
		const V2 dvc = o.vel - v;
		if (angle_between(dpc, a.vel) > half) continue;
		const float dpMag = mag(dpc);
		const V2 dp = dpc / dpMag * (dpMag - a.radius - o.radius);
		const V2 nrm{-dpc.y, dpc.x};
		const V2 dv = dvc + nrm * dot(nrm, dvc);
		const float dvSq = sqr_mag(dv);
		float ttca = (dvSq == 0.0f ? 0.0f : -dot(dp, dv) / dvSq);
		if (ttca < 0.0f) ttca = 0.0f;
		const V2 vdca = dp + dv * ttca;
		float dca = mag(vdca);
		if (dca < 0.0f) dca = 0.0f;
		if (dvSq <= eps) continue;
		const V2 velRot{dv.y, -dv.x};
		const float velMag = sqrtf(dvSq);
		const V2 velNorm = dv / velMag;
		float gTA = 0.0f, gTS = 0.0f;
		if (fabsf(ttca) > eps) {
			const V2 rel = dp + 2.0f * ttca * dv;
			gTA = -dot(rel, velRot) / dvSq;
			gTS = dot(rel, velNorm) / dvSq;
		}
		float gDA = 0.0f, gDS = 0.0f;
		if (fabsf(dca) > eps) {
			gDA = dot(vdca, gTA * dv + ttca * velRot) / dca;
			gDS = dot(vdca, gTS * dv - ttca * velNorm) / dca;
		}
		const float cst = ttca_pair_cost(sT, sD, ttca, dca);
		const float tf = ttca / (sT * sT);
		const float df = dca / (sD * sD);
		const float gCS = -cst * (gTS * tf + gDS * df);
		const float gCA = -cst * (gTA * tf + gDA * df);
		const float scale = 1.0f / sqr_mag(dp);
		tot += scale;
		GTh += gCA * scale;
		GS += gCS * scale;
	}
	if (tot > 0.0f) { GTh /= tot; GS /= tot; }
	const float alpha = ccw_angle(v, a.pref);
	const float ds = mag(v) - a.pref_speed;
	const double qa = (double)(alpha / sA), qs = (double)(ds / sS);
	const double DCa = (double)(alpha / (2.0f * sA * sA)) * exp(-0.5 * (qa * qa));
	const double DCs = (double)(ds / (2.0f * sS * sS)) * exp(-0.5 * (qs * qs));
	GTh += (float)DCa;
	GS += (float)DCs;
	V2 dir; float speed;
	if (sqr_mag(v) < 0.01f) { dir = normalized(a.pref); speed = mag(a.pref); }
	else { dir = normalized(v); speed = mag(v); }
	const V2 w = V2{sinf(GTh), 1.0f - cosf(GTh)} * speed + V2{0.0f, GS};
	return V2{dir.y * w.x + dir.x * w.y, -dir.x * w.x + dir.y * w.y};
}
// Paris::GetCost and helpers (Paris.cpp:28-213): one pass over the neighbours that
// accumulates both per-neighbour sums in list order.
UD void paris_t12(float t_max, V2 origin, V2 vel, V2 point, float radius, float& t1, float& t2) {
	t1 = time_to_collision(origin, vel, radius, point, V2{0.0f, 0.0f}, 0.0f);
	if (t1 > t_max) t1 = t_max;
	t2 = t1 + radius / mag(vel);
	if (t2 > t_max) t2 = 0.0f;
}
UD float paris_cost(const float* prm, float range, const Self& a, const NbList& L, V2 v) {
	const float w_a = prm[0], w_b = prm[1], t_max = prm[2];
	const V2 dir = normalized(v);
	const float rangeSq = range * range;
	const float prefSpeed = mag(a.pref);
	const float speed = mag(v);
	float sum = 0.0f, dev = 0.0f;
	int nR = 0;
	for (int e = 0; e < L.count; ++e) {
		const Nb o = L.get(e);
		if (o.dsq > rangeSq) continue;
		const float r1 = a.radius, r2 = o.radius;
		float t1, t2, s1, s2;
		V2 X;
		if (!line_intersection(a.pos, a.pos + dir, o.pos, o.pos + o.vel, X)) {
			const float ttc = time_to_collision(a.pos, dir, r1, o.pos, o.vel, r2);
			if (ttc > t_max) { t1 = t_max; t2 = 0.0f; s1 = 0.0f; s2 = kMaxFloat; }
			else { t1 = 0.0f; t2 = t_max; s1 = kMaxFloat; s2 = 0.0f; }
		} else {
			float ta1, ta2;
			paris_t12(t_max, a.pos, a.vel, X, r1 + r2, ta1, ta2);
			if (ta1 == t_max) { t1 = t_max; t2 = 0.0f; s1 = 0.0f; s2 = kMaxFloat; }
			else {
				paris_t12(t_max, o.pos, o.vel, X, r1 + r2, t1, t2);
				const float distToX = mag(X - a.pos);
				s1 = distToX / t1;
				s2 = distToX / t2;
			}
		}
		const float cs = smin(smax(0.0f, 1.0f - s1 / prefSpeed), smax(0.0f, (s2 - prefSpeed) / prefSpeed));
		const float cp = 1.0f - t1 / (t_max + w_b);
		sum += cs * cp;
		dev += smax(0.0f, smax((s1 - speed) / prefSpeed, (speed - s2) / prefSpeed));
		++nR;
	}
	if (nR > 0) sum /= (float)nR;
	const float cangle = (float)((double)(1.0f - cos_angle(dir, a.pref)) / 2.0);
	const float dirCost = w_a * sum + (1.0f - w_a) * cangle;
	const float diff = mag(v - a.pref);
	return 1000.0f * dirCost + dev + 0.001f * diff;
}

} // namespace umans
