// CPU oracle for the UMANS per-step agent loop — TEST INFRASTRUCTURE ONLY.
// See umans_oracle.h for scope. Every function cites the reference code it
// restates (paths relative to the reference repository root).
//
// Template parameter T is the arithmetic type. With T=float the code follows
// the reference's precision map exactly: values the reference computes in
// double are computed in double here and cast back at the same points.
// With T=double every such cast is a no-op, which gives the FP64 oracle.

#include "umans_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPI = 3.1415926535897;                       // vector2D.h:28
const float kMaxFloat = std::numeric_limits<float>::max(); // costFunction.h:43

template <typename T> struct V2 {
	T x, y;
	V2() : x(0), y(0) {}
	V2(T a, T b) : x(a), y(b) {}
	T sqr() const { return x * x + y * y; }              // vector2D.h:52
	T mag() const { return std::sqrt(x * x + y * y); }   // vector2D.h:48
	T dot(const V2& o) const { return x * o.x + y * o.y; }
	V2 normalized() const {                              // vector2D.h:55-72 (true division)
		V2 r(x, y);
		T m = mag();
		if (m > 0) { r.x /= m; r.y /= m; }
		return r;
	}
	bool zero() const { return x == 0 && y == 0; }
};
template <typename T> V2<T> operator+(V2<T> a, V2<T> b) { return V2<T>(a.x + b.x, a.y + b.y); }
template <typename T> V2<T> operator-(V2<T> a, V2<T> b) { return V2<T>(a.x - b.x, a.y - b.y); }
template <typename T> V2<T> operator-(V2<T> a) { return V2<T>(-a.x, -a.y); }
template <typename T> V2<T> operator*(T s, V2<T> v) { return V2<T>(v.x * s, v.y * s); }
template <typename T> V2<T> operator*(V2<T> v, T s) { return V2<T>(v.x * s, v.y * s); }
// vector2D.h:132-135: division multiplies by the reciprocal
template <typename T> V2<T> operator/(V2<T> v, T s) { return (T(1) / s) * v; }

// vector2D.h:166-170: `cos(radians)` of a float binds to the float overload in the
// compiled reference (its costFunction.o calls sincosf), so cos/sin are float, widened
// to double for the products, and the result is cast back. With T=double: double trig.
template <typename T> V2<T> rotate_ccw(V2<T> v, T radians) {
	const double c = std::cos(radians), s = std::sin(radians);
	return V2<T>((T)(c * v.x - s * v.y), (T)(s * v.x + c * v.y));
}
// vector2D.h:177-190
template <typename T> T angle_between(V2<T> a, V2<T> b) {
	const double lengths = a.mag() * b.mag();   // product in T, then widened
	if (lengths == 0) return T(0);
	const double frac = (a.dot(b) / lengths);
	if (frac < -1 || frac > 1) return T(0);
	return (T)std::acos(frac);
}
// vector2D.h:192-199
template <typename T> T cos_angle(V2<T> a, V2<T> b) {
	const double lengths = a.mag() * b.mag();
	if (lengths == 0) return T(0);
	return (T)(a.dot(b) / lengths);
}
// vector2D.h:201-219
template <typename T> T ccw_angle(V2<T> a, V2<T> b) {
	T ang = angle_between(a, b);
	T cross = a.x * b.y - a.y * b.x;
	return cross < 0 ? -ang : ang;
}
// vector2D.h:241-247
template <typename T> V2<T> clamp_vec(V2<T> v, T maxLength) {
	T m = v.mag();
	if (m <= maxLength) return v;
	return v.normalized() * maxLength;
}
// vector2D.h:249-272: coordinate differences in the vector type, the rest in double
template <typename T> bool line_intersection(V2<T> a, V2<T> b, V2<T> c, V2<T> d, V2<T>& res) {
	double a1 = (double)(b.y - a.y), b1 = (double)(a.x - b.x), c1 = a1 * a.x + b1 * a.y;
	double a2 = (double)(d.y - c.y), b2 = (double)(c.x - d.x), c2 = a2 * c.x + b2 * c.y;
	double det = a1 * b2 - a2 * b1;
	if (std::fabs(det) < 0.00001) return false;
	res = V2<T>((T)((b2 * c1 - b1 * c2) / det), (T)((a1 * c2 - a2 * c1) / det));
	return true;
}

// ---------------------------------------------------------------- agents
template <typename T> struct Agent {
	int32_t id;
	V2<T> pos, vel, goal, pref_vel;
	T radius, pref_speed, max_speed, max_accel;
	int policy;
	uint32_t rng;
};

// PhantomAgent (worldBase.h:40-79): real agent, offset position, float distSqr
template <typename T> struct Seg { V2<T> a, b; };   // LineSegment2D

// nearestPointOnLine (vector2D.h:274-296); on_seg reports 0 <= k <= 1 of the unclamped k
template <typename T> V2<T> nearest_point(V2<T> pt, V2<T> la, V2<T> lb, bool onSegment, bool* on_seg = nullptr) {
	if (la.x == lb.x && la.y == lb.y) return la;
	const V2<T> line = la - lb;
	T k = (pt - lb).dot(line) / line.sqr();
	if (onSegment) {
		if (k > 1) k = 1;
		else if (k < 0) k = 0;
	}
	if (on_seg) *on_seg = (k >= 0 && k <= 1);
	return (la * k) + (lb * (T(1) - k));
}
// distanceToLineSquared (vector2D.h:303-306)
template <typename T> T dist_line_sq(V2<T> pt, V2<T> la, V2<T> lb, bool onSegment) {
	return (pt - nearest_point(pt, la, lb, onSegment)).sqr();
}

template <typename T> struct Phantom {
	int idx;
	V2<T> pos;
	T dist_sqr;
	int seg;
	double d2;   // selection metric (for output only)
};

struct Ctx;

// --------------------------------------------------- neighbour selection
struct Cand { double d2; int32_t id; int idx; };

struct CellGrid {
	// Plain hash grid used only to enumerate candidates; membership is decided
	// by the exact predicate, so any candidate superset gives the same set.
	double cs;
	std::unordered_map<long long, std::vector<int>> cells;
	static long long key(long long cx, long long cy) { return (cx << 32) ^ (cy & 0xffffffffLL); }
	void build(const float* pos, int n, double cell) {
		cs = cell;
		cells.clear();
		cells.reserve(n * 2 + 1);
		for (int i = 0; i < n; ++i) {
			long long cx = (long long)std::floor((double)pos[2 * i] / cs);
			long long cy = (long long)std::floor((double)pos[2 * i + 1] / cs);
			cells[key(cx, cy)].push_back(i);
		}
	}
};

// One radius query (AgentKDTree.cpp:35-59 + nanoflann radiusSearch):
// d2 in double from float coordinates, strict < against double(float(r*r)),
// self excluded by ID, canonical order (d2, id).
static void radius_query(const CellGrid* grid, const float* pos, const int32_t* ids, int n,
                         float qx, float qy, float r, int32_t self_id, std::vector<Cand>& out) {
	out.clear();
	const double r2 = (double)(float)(r * r);
	auto test = [&](int j) {
		if (ids[j] == self_id) return;
		const double dx = (double)qx - (double)pos[2 * j];
		const double dy = (double)qy - (double)pos[2 * j + 1];
		const double d2 = dx * dx + dy * dy;
		if (d2 < r2) out.push_back({d2, ids[j], j});
	};
	if (grid == nullptr) {
		for (int j = 0; j < n; ++j) test(j);
	} else {
		const double cs = grid->cs;
		long long x0 = (long long)std::floor(((double)qx - r) / cs) - 1;
		long long x1 = (long long)std::floor(((double)qx + r) / cs) + 1;
		long long y0 = (long long)std::floor(((double)qy - r) / cs) - 1;
		long long y1 = (long long)std::floor(((double)qy + r) / cs) + 1;
		for (long long cy = y0; cy <= y1; ++cy)
			for (long long cx = x0; cx <= x1; ++cx) {
				auto it = grid->cells.find(CellGrid::key(cx, cy));
				if (it == grid->cells.end()) continue;
				for (int j : it->second) test(j);
			}
	}
	std::sort(out.begin(), out.end(), [](const Cand& a, const Cand& b) {
		return a.d2 < b.d2 || (a.d2 == b.d2 && a.id < b.id);
	});
}

// Policy::getInteractionRange (policy.cpp:34-41)
static float interaction_range(const oracle_policy& p) {
	float r = 0;
	for (int c = 0; c < p.n_cost; ++c) r = std::max(r, p.cost_range[c]);
	return r;
}

// WorldInfinite::ComputeNeighbors / WorldToric::ComputeNeighbors
// (worldInfinite.cpp:31-52, worldToric.cpp:32-79): segments (0,0), +W, -W, +H, -H,
// concatenated; the cap keeps the first k entries of the concatenation.
template <typename T>
static void compute_neighbors(const oracle_world& w, const CellGrid* grid, const float* fpos,
                              const int32_t* ids, int n, int i, float range,
                              std::vector<Phantom<T>>& out, int* full_count,
                              const std::vector<Agent<T>>& agents) {
	out.clear();
	std::vector<Cand> cand;
	const float px = fpos[2 * i], py = fpos[2 * i + 1];
	float disp[5][2] = {{0, 0}, {w.width, 0}, {-w.width, 0}, {0, w.height}, {0, -w.height}};
	bool use[5] = {true, false, false, false, false};
	if (w.world_type == 1) {
		use[1] = px - range < -0.5 * w.width;
		use[2] = px + range > 0.5 * w.width;
		use[3] = py - range < -0.5 * w.height;
		use[4] = py + range > 0.5 * w.height;
	}
	int total = 0;
	const int cap = w.max_neighbors > 0 ? w.max_neighbors : std::numeric_limits<int>::max();
	for (int s = 0; s < 5; ++s) {
		if (!use[s]) continue;
		const float qx = px + disp[s][0], qy = py + disp[s][1];
		radius_query(grid, fpos, ids, n, qx, qy, range, ids[i], cand);
		total += (int)cand.size();
		for (const Cand& c : cand) {
			if ((int)out.size() >= cap) break;
			Phantom<T> ph;
			ph.idx = c.idx;
			ph.seg = s;
			ph.d2 = c.d2;
			// PhantomAgent::UpdatePositionAndDistance (worldBase.h:73-77)
			const V2<T> off((T)(-disp[s][0]), (T)(-disp[s][1]));
			ph.pos = agents[c.idx].pos + off;
			const V2<T> q = agents[i].pos;
			ph.dist_sqr = (q - ph.pos).sqr();
			out.push_back(ph);
		}
	}
	if (full_count) *full_count = total;
}

// WorldBase::computeNeighboringObstacles_Flat (worldBase.cpp:62-77) per query segment of
// WorldInfinite / WorldToric::ComputeNeighbors (worldToric.cpp:32-79): every edge whose
// distance to float(position + displacement) is within the range (float arithmetic, as
// in the reference in both precisions), moved back by the displacement.
template <typename T>
static void compute_obstacles(const oracle_world& w, float px, float py, float range, std::vector<Seg<T>>& out) {
	out.clear();
	if (w.n_edges <= 0) return;
	float disp[5][2] = {{0, 0}, {w.width, 0}, {-w.width, 0}, {0, w.height}, {0, -w.height}};
	bool use[5] = {true, false, false, false, false};
	if (w.world_type == 1) {
		use[1] = px - range < -0.5 * w.width;
		use[2] = px + range > 0.5 * w.width;
		use[3] = py - range < -0.5 * w.height;
		use[4] = py + range > 0.5 * w.height;
	}
	const float r2 = range * range;
	for (int s = 0; s < 5; ++s) {
		if (!use[s]) continue;
		const V2<float> q(px + disp[s][0], py + disp[s][1]);
		for (int e = 0; e < w.n_edges; ++e) {
			const float* E = w.edges + 4 * e;
			const V2<float> a(E[0], E[1]), b(E[2], E[3]);
			if (dist_line_sq(q, a, b, true) <= r2) {
				const V2<float> d(disp[s][0], disp[s][1]);
				const V2<float> a2 = a - d, b2 = b - d;
				out.push_back(Seg<T>{V2<T>((T)a2.x, (T)a2.y), V2<T>((T)b2.x, (T)b2.y)});
			}
		}
	}
}

// --------------------------------------------------------- shared numerics
struct Params { const oracle_policy* pol; int c; };

template <typename T> struct AgentView {
	const Agent<T>* self;
	const std::vector<Agent<T>>* all;
	const std::vector<Phantom<T>>* nbs;
	T dt;
	float relaxation;   // Policy::getRelaxationTime
	const std::vector<Seg<T>>* obs;   // neighbouring obstacle edges (NeighborList::second)
};


// CostFunction::SolveQuadraticEquation (costFunction.cpp:173-191), incl. the
// D == 0 branch that evaluates (-b/2)*a.
template <typename T> int solve_quadratic(T a, T b, T c, T& r1, T& r2) {
	const T D = b * b - 4 * a * c;
	if (D < 0) return 0;
	if (D == 0) { r1 = -b / 2 * a; return 1; }
	const T m = -b / (2 * a);
	const T s = std::sqrt(D) / (2 * a);
	r1 = m + s;
	r2 = m - s;
	return 2;
}
// CostFunction::ComputeTimeToCollision (costFunction.cpp:193-227)
template <typename T> T time_to_collision(V2<T> p1, V2<T> v1, T r1, V2<T> p2, V2<T> v2, T r2) {
	const V2<T> P = p1 - p2;
	const T R = r1 + r2;
	const T RSq = R * R;
	if (P.sqr() <= RSq) return T(0);
	const V2<T> V = v1 - v2;
	const T a = V.dot(V);
	const T b = 2 * P.dot(V);
	const T c = P.dot(P) - RSq;
	T t1 = (T)kMaxFloat, t2 = (T)kMaxFloat;
	solve_quadratic(a, b, c, t1, t2);
	if (t1 < 0) t1 = (T)kMaxFloat;
	if (t2 < 0) t2 = (T)kMaxFloat;
	return std::min(t1, t2);
}
// CostFunction::ComputeTimeToCollision_LineSegmentInterior (costFunction.cpp:293-325)
template <typename T> T ttc_segment_interior(V2<T> pos, V2<T> vel, T radius, const Seg<T>& sg) {
	if (dist_line_sq(pos, sg.a, sg.b, true) <= radius * radius) return T(0);
	V2<T> X;
	if (!line_intersection(pos, pos + vel, sg.a, sg.b, X)) return (T)kMaxFloat;
	if (vel.dot(X - pos) < 0) return (T)kMaxFloat;
	bool on = false;
	nearest_point(X, sg.a, sg.b, false, &on);
	if (!on) return (T)kMaxFloat;
	const T speed = vel.mag();
	const T ttc = (X - pos).mag() / speed;
	const T lineAngle = angle_between(vel, sg.b - sg.a);
	const T sub = radius / std::sin(lineAngle);
	const T tsub = sub / speed;
	return std::max(T(0), ttc - tsub);
}
// CostFunction::ComputeTimeToCollision_LineSegment (costFunction.cpp:270-291)
template <typename T> T ttc_segment(V2<T> pos, V2<T> vel, T radius, const Seg<T>& sg) {
	const T t1 = time_to_collision(pos, vel, radius, sg.a, V2<T>(), T(0));
	if (t1 == 0) return T(0);
	if (sg.a.x == sg.b.x && sg.a.y == sg.b.y) return t1;
	const T t2 = time_to_collision(pos, vel, radius, sg.b, V2<T>(), T(0));
	if (t2 == 0) return T(0);
	const T ti = ttc_segment_interior(pos, vel, radius, sg);
	return std::min(std::min(t1, t2), ti);
}
// CostFunction::ComputeTimeToFirstCollision (costFunction.cpp:229-268)
template <typename T>
T time_to_first_collision(const AgentView<T>& av, V2<T> pos, V2<T> vel, T radius, T maxDist, bool ignoreCurrent) {
	T minTTC = (T)kMaxFloat;
	const T maxDistSq = maxDist * maxDist;
	for (const auto& nb : *av.nbs) {
		if (nb.dist_sqr > maxDistSq) continue;
		const Agent<T>& o = (*av.all)[nb.idx];
		T ttc = time_to_collision(pos, vel, radius, nb.pos, o.vel, o.radius);
		if (ignoreCurrent && ttc == 0) continue;
		if (ttc < minTTC) minTTC = ttc;
	}
	if (av.obs) {
		for (const Seg<T>& sg : *av.obs) {
			if (dist_line_sq(pos, sg.a, sg.b, true) > maxDistSq) continue;
			T ttc = ttc_segment(pos, vel, radius, sg);
			if (ignoreCurrent && ttc == 0) continue;
			if (ttc < minTTC) minTTC = ttc;
		}
	}
	return minTTC;
}
// CostFunction::ComputeTimeAndDistanceToClosestApproach (costFunction.cpp:354-380)
template <typename T> std::pair<T, T> ttca_dca(V2<T> p1, V2<T> v1, T r1, V2<T> p2, V2<T> v2, T r2) {
	const V2<T> dpc = p2 - p1;
	const V2<T> dvc = v2 - v1;
	T dpMag = dpc.mag();
	const V2<T> dp = dpc / dpMag * (dpMag - r1 - r2);
	V2<T> nrm(-dpc.y, dpc.x);
	const V2<T> dv = dvc + nrm * nrm.dot(dvc);
	T ttca = (dv.sqr() == 0 ? T(0) : -dp.dot(dv) / dv.sqr());
	if (ttca < 0) ttca = 0;
	T dca = (dp + dv * ttca).mag();
	if (dca < 0) dca = 0;
	return {ttca, dca};
}

// ----------------------------------------------------------- cost functions
template <typename T> T P(const Params& p, int k) { return (T)p.pol->cost_params[p.c][k]; }
template <typename T> T RANGE(const Params& p) { return (T)p.pol->cost_range[p.c]; }

// PowerLaw::ComputeAgentInteractionForce (PowerLaw.cpp:28-66)
template <typename T> V2<T> power_law_pair(const Params& pp, const Agent<T>& a, const Phantom<T>& nb, const Agent<T>& o) {
	const T k = P<T>(pp, 1), tau0 = P<T>(pp, 2);
	T R = o.radius + a.radius;
	const V2<T> x = a.pos - nb.pos;
	const V2<T> v = a.vel - o.vel;
	T aa = v.dot(v);
	T b = -x.dot(v);
	T c = x.dot(x) - R * R;
	if (c <= 0) return V2<T>();
	T d = b * b - aa * c;
	if (d <= 0) return V2<T>();
	T sqrtD = std::sqrt(d);
	T tau = (b - sqrtD) / aa;
	if (tau < (T)0.001f) return V2<T>();
	if (tau > tau0) return V2<T>();
	T comp1 = -k * std::exp(-tau / tau0) * (2 / tau + 1 / tau0) / (aa * tau * tau);
	const V2<T> comp2 = v - (aa * x + b * v) / sqrtD;
	return comp1 * comp2;
}

// SocialForcesAvoidance::ComputeAgentInteractionForce (SocialForcesAvoidance.cpp:26-110)
template <typename T> V2<T> social_force_pair(const Params& pp, const Agent<T>& a, const Phantom<T>& nb, const Agent<T>& o) {
	const T sdt = P<T>(pp, 1), V0 = P<T>(pp, 2), sigma = P<T>(pp, 3);
	const V2<T> R = a.pos - nb.pos;
	T magR = R.mag();
	if (magR <= a.radius + o.radius) return V2<T>();
	T vb = o.vel.mag();
	const V2<T> Eb = o.pref_vel.normalized();
	const V2<T> V = vb * sdt * Eb;
	T magV = V.mag();
	const V2<T> R2 = R - V;
	T magR2 = R2.mag();
	T bSq2 = (T)(std::pow((double)(magR + magR2), 2) - std::pow((double)(vb * sdt), 2));
	if (bSq2 <= 0) return V2<T>();
	T b = T(0.5f) * std::sqrt(bSq2);
	if (b < (T)0.001f) return V2<T>();
	const T d_magR2 = 1 / (2 * magR2) * (2 * magR - 2 * magV * cos_angle(R, V));
	const T d_bSq2 = 2 * (magR + magR2) * (1 + d_magR2);
	const T d_b = (1 / (8 * b)) * d_bSq2;
	const T d_V = -V0 / sigma * std::exp(-b / sigma) * d_b;
	const V2<T> F = -R * d_V;
	const T half = (T)(float)(100.0 * kPI / 180.0);
	T scale = (angle_between(a.vel, -R) < half ? T(1) : T(0.5f));
	return scale * F;
}

// SocialForcesAvoidance::ComputeObstacleInteractionForce (SocialForcesAvoidance.cpp:112-140).
// U0 keeps its default 10: parseParameters reads the "U0" attribute into V0
// (SocialForcesAvoidance.cpp:142-150); R is parameter slot 4.
template <typename T> V2<T> social_force_obstacle(const Params& pp, const Agent<T>& a, const Seg<T>& sg) {
	const T U0 = T(10.0f), R = P<T>(pp, 4);
	const V2<T> nearest = nearest_point(a.pos, sg.a, sg.b, true);
	const V2<T> diff = a.pos - nearest;
	const T dist = diff.mag();
	if (dist < (T)0.001f) return V2<T>();
	const T d_U = -U0 / R * std::exp(-dist / R);
	return -diff * d_U;
}

// ObjectInteractionForces::ComputeForce (ObjectInteractionForces.cpp:26-48) and
// GoalReachingForce::ComputeForce (GoalReachingForce.cpp:26-29); PowerLaw's obstacle
// force is zero (PowerLaw.cpp:68-72)
template <typename T> V2<T> force_of(const Params& pp, const AgentView<T>& av) {
	const Agent<T>& a = *av.self;
	const int kind = pp.pol->cost_kind[pp.c];
	if (kind == ORC_GOAL_REACHING)
		return (a.pref_vel - a.vel) / std::max((T)av.relaxation, av.dt);
	const T rangeSq = RANGE<T>(pp) * RANGE<T>(pp);
	V2<T> F;
	for (const auto& nb : *av.nbs) {
		if (nb.dist_sqr <= rangeSq) {
			const Agent<T>& o = (*av.all)[nb.idx];
			F = F + (kind == ORC_POWER_LAW ? power_law_pair(pp, a, nb, o) : social_force_pair(pp, a, nb, o));
		}
	}
	V2<T> Fo;
	if (av.obs) {
		for (const Seg<T>& sg : *av.obs) {
			if (dist_line_sq(a.pos, sg.a, sg.b, true) <= rangeSq)
				Fo = Fo + (kind == ORC_SOCIAL_FORCES ? social_force_obstacle(pp, a, sg) : V2<T>());
		}
	}
	return F + Fo;
}
// ForceBasedFunction::ComputeTargetVelocity (ForceBasedFunction.cpp:49-52)
template <typename T> V2<T> target_velocity(const Params& pp, const AgentView<T>& av) {
	const T mass = P<T>(pp, 0);
	return av.self->vel + force_of(pp, av) / mass * av.dt;
}

// ---------------------------------------------------------------- ORCA
struct OrcaLineT;
template <typename T> struct Line { V2<T> point, dir; };
template <typename T> T det2(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }  // ORCASolver.cpp:55-58
const float kOrcaEps = 0.00001f;  // ORCASolver.h:62

// linearProgram1 (ORCASolver.cpp:173-244)
template <typename T>
bool lp1(const std::vector<Line<T>>& L, size_t no, T radius, V2<T> opt, bool dirOpt, V2<T>& res) {
	const T dp = L[no].point.dot(L[no].dir);
	const T disc = dp * dp + radius * radius - L[no].point.sqr();
	if (disc < 0) return false;
	const T sd = std::sqrt(disc);
	T tl = -dp - sd, tr = -dp + sd;
	for (size_t i = 0; i < no; ++i) {
		const T den = det2(L[no].dir, L[i].dir);
		const T num = det2(L[i].dir, L[no].point - L[i].point);
		if (std::fabs(den) <= (T)kOrcaEps) {
			if (num < 0) return false;
			continue;
		}
		const T t = num / den;
		if (den >= 0) tr = std::min(tr, t);
		else tl = std::max(tl, t);
		if (tl > tr) return false;
	}
	if (dirOpt) {
		res = opt.dot(L[no].dir) > 0 ? L[no].point + L[no].dir * tr : L[no].point + L[no].dir * tl;
	} else {
		const T t = L[no].dir.dot(opt - L[no].point);
		if (t < tl) res = L[no].point + L[no].dir * tl;
		else if (t > tr) res = L[no].point + L[no].dir * tr;
		else res = L[no].point + L[no].dir * t;
	}
	return true;
}
// linearProgram2 (ORCASolver.cpp:246-277)
template <typename T>
size_t lp2(const std::vector<Line<T>>& L, T radius, V2<T> opt, bool dirOpt, V2<T>& res) {
	if (dirOpt) res = opt * radius;
	else if (opt.sqr() > radius * radius) res = opt.normalized() * radius;
	else res = opt;
	for (size_t i = 0; i < L.size(); ++i) {
		if (det2(L[i].dir, L[i].point - res) > 0) {
			const V2<T> tmp = res;
			if (!lp1(L, i, radius, opt, dirOpt, res)) { res = tmp; return i; }
		}
	}
	return L.size();
}
// linearProgram3 (ORCASolver.cpp:279-326), no obstacle lines
template <typename T>
void lp3(const std::vector<Line<T>>& L, size_t begin, T radius, V2<T>& res) {
	T distance = 0;
	for (size_t i = begin; i < L.size(); ++i) {
		if (det2(L[i].dir, L[i].point - res) > distance) {
			std::vector<Line<T>> proj;
			for (size_t j = 0; j < i; ++j) {
				Line<T> line;
				T determinant = det2(L[i].dir, L[j].dir);
				if (std::fabs(determinant) <= (T)kOrcaEps) {
					if (L[i].dir.dot(L[j].dir) > 0) continue;
					line.point = (L[i].point + L[j].point) * T(0.5f);
				} else {
					line.point = L[i].point + L[i].dir * (det2(L[j].dir, L[i].point - L[j].point) / determinant);
				}
				line.dir = (L[j].dir - L[i].dir).normalized();
				proj.push_back(line);
			}
			const V2<T> tmp = res;
			if (lp2(proj, radius, V2<T>(-L[i].dir.y, L[i].dir.x), true, res) < proj.size()) res = tmp;
			distance = det2(L[i].dir, L[i].point - res);
		}
	}
}
// Solver::createAgentOrcaLines (ORCASolver.cpp:96-171), incl. the collision
// branch that normalises by the squared length.
template <typename T>
void orca_lines(const Params& pp, const AgentView<T>& av, std::vector<Line<T>>& L) {
	const Agent<T>& a = *av.self;
	const T tau = P<T>(pp, 0);
	const T maxD = RANGE<T>(pp);
	const T maxDistSq = maxD * maxD;
	L.clear();
	for (const auto& nb : *av.nbs) {
		if (nb.dist_sqr > maxDistSq) continue;
		const Agent<T>& o = (*av.all)[nb.idx];
		const V2<T> relPos = nb.pos - a.pos;
		const V2<T> relVel = a.vel - o.vel;
		const T distSq = relPos.sqr();
		const T cr = a.radius + o.radius;
		const T crSq = cr * cr;
		Line<T> line;
		V2<T> u;
		if (distSq > crSq) {
			const V2<T> w = relVel - (relPos / tau);
			const T wLenSq = w.sqr();
			const T dp1 = w.dot(relPos);
			if (dp1 < 0 && dp1 * dp1 > crSq * wLenSq) {
				const T wLen = std::sqrt(wLenSq);
				const V2<T> unitW = w / wLen;
				line.dir = V2<T>(unitW.y, -unitW.x);
				u = unitW * (cr / tau - wLen);
			} else {
				const T leg = std::sqrt(distSq - crSq);
				if (det2(relPos, w) > 0)
					line.dir = V2<T>(relPos.x * leg - relPos.y * cr, relPos.x * cr + relPos.y * leg) / distSq;
				else
					line.dir = V2<T>(relPos.x * leg + relPos.y * cr, -relPos.x * cr + relPos.y * leg) * T(-1) / distSq;
				const T dp2 = relVel.dot(line.dir);
				u = (line.dir * dp2) - relVel;
			}
		} else {
			const V2<T> w = relVel - (relPos / av.dt);
			const T wLen = w.sqr();
			const V2<T> unitW = w / wLen;
			line.dir = V2<T>(unitW.y, -unitW.x);
			u = unitW * (cr / av.dt - wLen);
		}
		line.point = a.vel + (u * T(0.5f));
		L.push_back(line);
	}
}
// Solver::solveOrcaProgram (ORCASolver.cpp:61-94)
template <typename T> V2<T> orca_solve(const Params& pp, const AgentView<T>& av, std::vector<Line<T>>& L) {
	orca_lines(pp, av, L);
	V2<T> res;
	const T maxSpeed = av.self->max_speed;
	size_t fail = lp2(L, maxSpeed, av.self->pref_vel, false, res);
	if (fail != L.size()) lp3(L, fail, maxSpeed, res);
	return res;
}

// Cost of one function at a velocity: CostFunction::GetCost of each class.
template <typename T> struct CostEval {
	const oracle_policy* pol;
	const AgentView<T>* av;
	// per-step caches (values are pure functions of the agent's state)
	std::vector<V2<T>> target;          // force-based: target velocity
	std::vector<std::vector<Line<T>>> orca;  // ORCA lines
	std::vector<V2<T>> orca_vel;

	void init() {
		const int nc = pol->n_cost;
		target.assign(nc, V2<T>());
		orca.assign(nc, {});
		orca_vel.assign(nc, V2<T>());
		for (int c = 0; c < nc; ++c) {
			Params pp{pol, c};
			const int k = pol->cost_kind[c];
			if (k == ORC_GOAL_REACHING || k == ORC_POWER_LAW || k == ORC_SOCIAL_FORCES)
				target[c] = target_velocity(pp, *av);
			else if (k == ORC_ORCA)
				orca_vel[c] = orca_solve(pp, *av, orca[c]);
		}
	}

	T cost(int c, V2<T> v) const {
		Params pp{pol, c};
		const Agent<T>& a = *av->self;
		const int kind = pol->cost_kind[c];
		switch (kind) {
		case ORC_GOAL_REACHING: case ORC_POWER_LAW: case ORC_SOCIAL_FORCES:
			// ForceBasedFunction::GetCost (ForceBasedFunction.cpp:26-30)
			return T(0.5f) * (v - target[c]).sqr() / av->dt;
		case ORC_ORCA: {
			// ORCA::GetCost (ORCA.cpp:49-86)
			T maxD = -(T)kMaxFloat;
			for (const auto& l : orca[c]) {
				const V2<T> pmv = l.point - v;
				maxD = std::max(maxD, l.dir.x * pmv.y - l.dir.y * pmv.x);
			}
			if (maxD <= 0) return (v - a.pref_vel).mag();
			return maxD + 2 * a.max_speed;
		}
		case ORC_KARAMOUZAS: return karamouzas(pp, v);
		case ORC_RVO: {
			// RVO::GetCost (RVO.cpp:28-44)
			if (v.mag() > a.max_speed) return (T)kMaxFloat;
			const V2<T> rv = T(2) * v - a.vel;
			const V2<T> vd = a.pref_vel - v;
			T ttc = time_to_first_collision(*av, a.pos, rv, a.radius, RANGE<T>(pp), true);
			return P<T>(pp, 0) / ttc + vd.mag();
		}
		case ORC_MOUSSAID: return moussaid(pp, v);
		case ORC_TTCA_DCA: return ttca_dca_cost(pp, v);
		case ORC_PARIS: return paris(pp, v);
		case ORC_VAN_TOLL: return van_toll(pp, v);
		}
		return 0;
	}

	// Karamouzas::GetCost and helpers (Karamouzas.cpp:28-113); tc_max stays 8
	T karamouzas(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T alpha = P<T>(pp, 0), beta = P<T>(pp, 1), gamma = P<T>(pp, 2), delta = P<T>(pp, 3), t_max = P<T>(pp, 4);
		const T tc_min = T(2.5f), tc_mid = T(6), tc_max = T(8);
		const T d_min = T(0.05f), d_mid = (T)(float)(kPI / 6.0f), d_max = (T)(float)(kPI / 2.0f);
		const T speed = v.mag();
		const T maxSpeed = a.max_speed;
		T ttcp = time_to_first_collision(*av, a.pos, a.pref_vel, a.radius, RANGE<T>(pp), true);
		T maxDev;
		if (ttcp <= 0 || ttcp >= tc_max) maxDev = d_min;
		else if (ttcp < tc_min) maxDev = (T)(d_mid + std::pow((double)((tc_min - ttcp) / tc_min), 2) * (d_max - d_mid));
		else if (ttcp < tc_mid) maxDev = d_mid;
		else maxDev = d_min + d_mid - (tc_max - ttcp) / (tc_max - tc_mid);
		if (angle_between(v, a.pref_vel) > maxDev) return (T)kMaxFloat;
		const T pref = a.pref_speed, dS = std::abs(maxSpeed - pref), eps = T(0.05f);
		T minS, maxS;
		if (ttcp <= 0 || ttcp >= tc_max) { minS = pref - eps; maxS = pref + eps; }
		else if (ttcp > 0 && ttcp < tc_min) { minS = 0; maxS = maxSpeed + eps; }
		else { minS = pref - dS - eps; maxS = pref + dS + eps; }
		if (speed < minS || speed > maxS) return (T)kMaxFloat;
		T ttc = time_to_first_collision(*av, a.pos, v, a.radius, RANGE<T>(pp), true);
		T A = alpha * (1 - cos_angle(v, a.pref_vel) / T(2.0f));
		T B = beta * std::abs(speed - a.vel.mag()) / maxSpeed;
		T C = gamma * (v - a.pref_vel).mag() / (2 * maxSpeed);
		T D = delta * std::max(T(0), t_max - ttc) / t_max;
		return A + B + C + D;
	}
	// Moussaid::GetCost (Moussaid.cpp:28-62)
	T moussaid(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T d_max = P<T>(pp, 0);
		const T speed = v.mag();
		const T prefSpeed = a.pref_speed;
		const V2<T> dir = v.normalized();
		const V2<T> vp = dir * prefSpeed;
		T ttc = time_to_first_collision(*av, a.pos, vp, a.radius, RANGE<T>(pp), false);
		T DC = (ttc == (T)kMaxFloat ? (T)kMaxFloat : ttc * prefSpeed);
		if (DC > d_max) DC = d_max;
		T K = d_max * d_max + DC * DC - 2 * d_max * DC * cos_angle(dir, a.pref_vel);
		K += 1;
		T S = std::min(prefSpeed, DC / (T)av->relaxation);
		return (T)(K * (1 + std::pow((double)((S - speed) / S), 2.0)));
	}
	// VanToll::GetCost (VanToll.cpp:29-64)
	T van_toll(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T maxd = P<T>(pp, 0);
		const T prefSpeed = a.pref_vel.mag();
		const T speed = v.mag();
		const T ttc = time_to_first_collision(*av, a.pos, v, a.radius, RANGE<T>(pp), false);
		T dist = (ttc == (T)kMaxFloat ? (T)kMaxFloat : ttc * speed);
		if (dist > maxd) dist = maxd;
		const T devPA = angle_between(v, a.pref_vel);
		const T devPS = std::abs(speed - prefSpeed) / prefSpeed;
		const T devCA = angle_between(v, a.vel);
		return T(1) * (maxd - dist) + T(1) * devPA + T(1) * devPS + T(1) * devCA;
	}
	// TtcaDca::GetCost (TtcaDca.cpp:34-48, 62-129)
	T ttca_pair_cost(const Params& pp, T ttca, T dca) const {
		const double sT = P<T>(pp, 2), sD = P<T>(pp, 3);
		return (T)std::exp(-0.5 * (std::pow((double)(ttca / (T)sT), 2) + std::pow((double)(dca / (T)sD), 2)));
	}
	T ttca_dca_cost(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T sA = P<T>(pp, 0), sS = P<T>(pp, 1);
		const T rangeSq = RANGE<T>(pp) * RANGE<T>(pp);
		T alpha = angle_between(v, a.pref_vel);
		T ds = v.mag() - a.pref_speed;
		T Ca = (T)std::exp(-0.5 * std::pow((double)(alpha / sA), 2.0));
		T Cs = (T)std::exp(-0.5 * std::pow((double)(ds / sS), 2.0));
		T movement = 1 - (Ca + Cs) / T(2.0f);
		T obst = 0, scaleSum = 0;
		const T half = (T)(float)(kPI / 2.0);
		for (const auto& nb : *av->nbs) {
			if (nb.dist_sqr >= rangeSq) continue;
			const Agent<T>& o = (*av->all)[nb.idx];
			const V2<T> relPos = nb.pos - a.pos;
			if (angle_between(relPos, a.vel) > half) continue;
			auto td = ttca_dca(a.pos, v, a.radius, nb.pos, o.vel, o.radius);
			T distance = relPos.mag() - a.radius - o.radius;
			T scale = 1 / (distance * distance);
			scaleSum += scale;
			obst += ttca_pair_cost(pp, td.first, td.second) * scale;
		}
		if (scaleSum > 0) obst /= scaleSum;
		return obst + movement;
	}
	// Paris::GetCost and helpers (Paris.cpp:28-213)
	struct Range4 { T t1, t2, s1, s2; };
	std::pair<T, T> paris_t12(const Params& pp, V2<T> origin, V2<T> vel, V2<T> point, T radius) const {
		const T t_max = P<T>(pp, 2);
		T t1 = time_to_collision(origin, vel, radius, point, V2<T>(), T(0));
		if (t1 > t_max) t1 = t_max;
		T t2 = t1 + radius / vel.mag();
		if (t2 > t_max) t2 = 0;
		return {t1, t2};
	}
	T paris(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T w_a = P<T>(pp, 0), w_b = P<T>(pp, 1), t_max = P<T>(pp, 2);
		const T MF = (T)kMaxFloat;
		const V2<T> dir = v.normalized();
		const T rangeSq = RANGE<T>(pp) * RANGE<T>(pp);
		std::vector<Range4> R;
		for (const auto& nb : *av->nbs) {
			if (nb.dist_sqr > rangeSq) continue;
			const Agent<T>& o = (*av->all)[nb.idx];
			const T r1 = a.radius, r2 = o.radius;
			V2<T> X;
			if (!line_intersection(a.pos, a.pos + dir, nb.pos, nb.pos + o.vel, X)) {
				T ttc = time_to_collision(a.pos, dir, r1, nb.pos, o.vel, r2);
				if (ttc > t_max) R.push_back({t_max, 0, 0, MF});
				else R.push_back({0, t_max, MF, 0});
				continue;
			}
			auto ta = paris_t12(pp, a.pos, a.vel, X, r1 + r2);
			if (ta.first == t_max) { R.push_back({t_max, 0, 0, MF}); continue; }
			auto t12 = paris_t12(pp, nb.pos, o.vel, X, r1 + r2);
			T distToX = (X - a.pos).mag();
			R.push_back({t12.first, t12.second, distToX / t12.first, distToX / t12.second});
		}
		const T prefSpeed = a.pref_vel.mag();
		T sum = 0;
		for (const auto& r : R) {
			T cs = std::min(std::max(T(0), T(1.0f) - r.s1 / prefSpeed), std::max(T(0), (r.s2 - prefSpeed) / prefSpeed));
			T cp = 1 - r.t1 / (t_max + w_b);
			sum += cs * cp;
		}
		if (!R.empty()) sum /= (T)R.size();
		T cangle = (T)((1 - cos_angle(dir, a.pref_vel)) / 2.0);
		T dirCost = w_a * sum + (1 - w_a) * cangle;
		const T speed = v.mag();
		T dev = 0;
		for (const auto& r : R)
			dev += std::max(T(0), std::max((r.s1 - speed) / prefSpeed, (speed - r.s2) / prefSpeed));
		T diff = (v - a.pref_vel).mag();
		return 1000 * dirCost + dev + T(0.001f) * diff;
	}

	// CostFunction::GetGradient default (costFunction.cpp:36-54) and overrides
	V2<T> gradient(int c, V2<T> v) const {
		Params pp{pol, c};
		const int kind = pol->cost_kind[c];
		if (kind == ORC_GOAL_REACHING || kind == ORC_POWER_LAW || kind == ORC_SOCIAL_FORCES)
			return (v - target[c]) / av->dt;   // ForceBasedFunction.cpp:32-36
		if (kind == ORC_TTCA_DCA) return ttca_dca_gradient(pp, v);
		const T d = T(0.1f);
		V2<T> dx(d, 0), dy(0, d);
		T cl = cost(c, v - dx), cr = cost(c, v + dx);
		T gx = (cr - cl) / (2 * d);
		T cb = cost(c, v - dy), ct = cost(c, v + dy);
		T gy = (ct - cb) / (2 * d);
		return V2<T>(gx, gy);
	}
	// TtcaDca::GetGradient (TtcaDca.cpp:131-262)
	V2<T> ttca_dca_gradient(const Params& pp, V2<T> v) const {
		const Agent<T>& a = *av->self;
		const T sA = P<T>(pp, 0), sS = P<T>(pp, 1), sT = P<T>(pp, 2), sD = P<T>(pp, 3);
		const T eps = T(0.001f);
		const T rangeSq = RANGE<T>(pp) * RANGE<T>(pp);
		const T half = (T)(float)(kPI / 2.0);
		T GTh = 0, GS = 0, tot = 0;
		for (const auto& nb : *av->nbs) {
			if (nb.dist_sqr > rangeSq) continue;
			const Agent<T>& o = (*av->all)[nb.idx];
			const V2<T> dpc = nb.pos - a.pos;
			const V2<T> dvc = o.vel - v;
			if (angle_between(dpc, a.vel) > half) continue;
			T dpMag = dpc.mag();
			const V2<T> dp = dpc / dpMag * (dpMag - a.radius - o.radius);
			V2<T> nrm(-dpc.y, dpc.x);
			const V2<T> dv = dvc + nrm * nrm.dot(dvc);
			T dvSq = dv.sqr();
			T ttca = (dvSq == 0 ? T(0) : -dp.dot(dv) / dvSq);
			if (ttca < 0) ttca = 0;
			const V2<T> vdca(dp + dv * ttca);
			T dca = vdca.mag();
			if (dca < 0) dca = 0;
			if (dvSq <= eps) continue;
			const V2<T> VelRot(dv.y, -dv.x);
			const T VelMag = std::sqrt(dvSq);
			const V2<T> VelNorm = dv / VelMag;
			T gTA = 0, gTS = 0;
			if (std::abs(ttca) > eps) {
				const V2<T> rel = (dp + T(2) * ttca * dv);
				gTA = -rel.dot(VelRot) / dvSq;
				gTS = rel.dot(VelNorm) / dvSq;
			}
			T gDA = 0, gDS = 0;
			if (std::abs(dca) > eps) {
				gDA = vdca.dot(gTA * dv + ttca * VelRot) / dca;
				gDS = vdca.dot(gTS * dv - ttca * VelNorm) / dca;
			}
			T cst = ttca_pair_cost(pp, ttca, dca);
			T tf = ttca / (sT * sT);
			T df = dca / (sD * sD);
			T gCS = -cst * (gTS * tf + gDS * df);
			T gCA = -cst * (gTA * tf + gDA * df);
			T scale = 1 / dp.sqr();
			tot += scale;
			GTh += gCA * scale;
			GS += gCS * scale;
		}
		if (tot > 0) { GTh /= tot; GS /= tot; }
		// getMovementCostGradient (TtcaDca.cpp:50-60)
		T alpha = ccw_angle(v, a.pref_vel);
		T ds = v.mag() - a.pref_speed;
		double DCa = alpha / (2 * sA * sA) * std::exp(-0.5 * std::pow((double)(alpha / sA), 2.0));
		double DCs = ds / (2 * sS * sS) * std::exp(-0.5 * std::pow((double)(ds / sS), 2.0));
		GTh += (T)DCa;
		GS += (T)DCs;
		// RotateGradientToEuclideanCoordinates (costFunction.cpp:30-34)
		auto rot = [&](V2<T> dir, T speed) {
			V2<T> w = V2<T>(std::sin(GTh), 1 - std::cos(GTh)) * speed + V2<T>(0, GS);
			return V2<T>(dir.y * w.x + dir.x * w.y, -dir.x * w.x + dir.y * w.y);
		};
		if (v.sqr() < 0.01) return rot(a.pref_vel.normalized(), a.pref_vel.mag());
		return rot(v.normalized(), v.mag());
	}

	// CostFunction::GetGlobalMinimum overrides: ORCA (ORCA.cpp:88-95), force-based
	// (ForceBasedFunction.cpp:38-41); others sample (costFunction.cpp:56-64)
	bool global_min_closed_form(int c, V2<T>& out) const {
		const int kind = pol->cost_kind[c];
		if (kind == ORC_ORCA) { out = orca_vel[c]; return true; }
		if (kind == ORC_GOAL_REACHING || kind == ORC_POWER_LAW || kind == ORC_SOCIAL_FORCES) { out = target[c]; return true; }
		return false;
	}
};

// Recording URBG: lets std::uniform_real_distribution draw from std::minstd_rand0
// while exposing the state after each draw (agent.cpp:131-135).
struct RecRng {
	using result_type = std::minstd_rand0::result_type;
	std::minstd_rand0 eng;
	result_type last;
	explicit RecRng(uint32_t s) : eng(s), last(s) {}
	static constexpr result_type min() { return std::minstd_rand0::min(); }
	static constexpr result_type max() { return std::minstd_rand0::max(); }
	result_type operator()() { last = eng(); return last; }
};

struct Sampling {
	int type, base, base_dir, radius;
	float angle;
	int speed_samples, angle_samples, random_samples, include_base;
};
static Sampling sampling_of(const oracle_policy& p) {
	return {p.s_type, p.s_base, p.s_base_dir, p.s_radius, p.s_angle, p.s_speed_samples,
	        p.s_angle_samples, p.s_random_samples, p.s_include_base};
}
// SamplingParameters::ApproximateGlobalOptimization (policy.h:74-86)
static Sampling global_sampling() { return {0, 0, 0, 1, 360.0f, 11, 36, 100, 1}; }

// CostFunction::ApproximateGlobalMinimumBySampling (costFunction.cpp:66-164)
template <typename T>
V2<T> sample_best(const CostEval<T>& ce, const Sampling& sp, const std::vector<std::pair<int, T>>& fns, uint32_t& rng) {
	const Agent<T>& a = *ce.av->self;
	V2<T> base;
	if (sp.base == 1) base = a.vel;
	T radius = a.pref_speed;
	if (sp.radius == 1) radius = a.max_speed;
	else if (sp.radius == 2) radius = std::min(T(2.0f) * a.max_speed, a.max_accel * ce.av->dt);
	V2<T> bdir(1, 0);
	if (sp.base_dir == 1) bdir = a.vel.normalized();
	else if (sp.base_dir == 2) bdir = a.pref_vel.normalized();
	T maxAngle = (T)(sp.angle / 360.0 * kPI);
	V2<T> best;
	T bestCost = (T)kMaxFloat;
	auto eval = [&](V2<T> v) {
		T tc = 0;
		for (auto& f : fns) tc += f.second * ce.cost(f.first, v);
		if (tc < bestCost) { best = v; bestCost = tc; }
	};
	if (sp.type == 1) {
		RecRng eng(rng);
		for (int i = 0; i < sp.random_samples; ++i) {
			// agent.cpp:131-135 (distribution constructed per draw)
			T ra, rl;
			{ std::uniform_real_distribution<float> d((float)-maxAngle, (float)maxAngle); ra = (T)d(eng); }
			{ std::uniform_real_distribution<float> d(0.0f, (float)radius); rl = (T)d(eng); }
			eval(base + rotate_ccw(bdir, ra) * rl);
		}
		rng = eng.last;
	} else {
		const T start = -maxAngle, end = maxAngle;
		const T dA = (end - start) / (T)(sp.angle == 360 ? sp.angle_samples : (sp.angle_samples - 1));
		const T dL = radius / (T)(sp.include_base ? (sp.speed_samples - 1) : sp.speed_samples);
		for (int s = 1; s <= sp.speed_samples; ++s) {
			const T len = dL * (T)(sp.include_base ? s - 1 : s);
			T ang = start;
			for (int k = 0; k < sp.angle_samples; ++k, ang += dA) {
				eval(base + rotate_ccw(bdir, ang) * len);
				if (sp.include_base && s == 1) break;
			}
		}
	}
	return best;
}

template <typename T>
int step_impl(const oracle_world& w, const oracle_policy* pols, int npol, int n,
              const oracle_agents_in& in, oracle_agents_out& out, int nthreads) {
	std::vector<Agent<T>> ag(n);
	for (int i = 0; i < n; ++i) {
		Agent<T>& a = ag[i];
		a.id = in.id[i];
		a.pos = V2<T>(in.pos[2 * i], in.pos[2 * i + 1]);
		a.vel = V2<T>(in.vel[2 * i], in.vel[2 * i + 1]);
		a.goal = V2<T>(in.goal[2 * i], in.goal[2 * i + 1]);
		a.radius = in.radius[i];
		a.pref_speed = in.pref_speed[i];
		a.max_speed = in.max_speed[i];
		a.max_accel = in.max_accel[i];
		a.policy = in.policy[i];
		// no state given: the engine as Agent::Agent seeds it, with the ID (agent.cpp:38)
		a.rng = in.rng ? in.rng[i] : oracle_rng_seed((uint32_t)in.id[i]);
		if (a.policy < 0 || a.policy >= npol) return -2;
	}
	// Agent::ComputePreferredVelocity (agent.cpp:52-59, 99-102)
	for (auto& a : ag) {
		if ((a.goal - a.pos).sqr() <= a.radius * a.radius) a.pref_vel = V2<T>();
		else a.pref_vel = (a.goal - a.pos).normalized() * a.pref_speed;
	}
	float maxRange = 0;
	for (int p = 0; p < npol; ++p) maxRange = std::max(maxRange, interaction_range(pols[p]));
	CellGrid grid;
	grid.build(in.pos, n, std::max(1.0, (double)maxRange));
	const T dt = (T)w.delta_time;

#ifdef _OPENMP
	if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
	for (int i = 0; i < n; ++i) {
		const Agent<T>& a = ag[i];
		const oracle_policy& pol = pols[a.policy];
		std::vector<Phantom<T>> nbs;
		int full = 0;
		compute_neighbors<T>(w, &grid, in.pos, in.id, n, i, interaction_range(pol), nbs, &full, ag);
		std::vector<Seg<T>> obs;
		compute_obstacles<T>(w, in.pos[2 * i], in.pos[2 * i + 1], interaction_range(pol), obs);
		if (out.nb_count) {
			const int m = std::min((int)nbs.size(), out.list_stride);
			out.nb_count[i] = (int)nbs.size();
			for (int e = 0; e < m; ++e) {
				out.nb_id[(long long)i * out.list_stride + e] = ag[nbs[e].idx].id;
				if (out.nb_d2) out.nb_d2[(long long)i * out.list_stride + e] = nbs[e].d2;
				if (out.nb_seg) out.nb_seg[(long long)i * out.list_stride + e] = nbs[e].seg;
			}
		}
		if (out.nb_full) out.nb_full[i] = full;

		AgentView<T> av{&a, &ag, &nbs, dt, pol.relaxation_time, &obs};
		uint32_t rng = a.rng;
		V2<T> next;
		// Policy::ComputeNewVelocity (policy.cpp:43-76)
		if (!a.pref_vel.zero()) {
			CostEval<T> ce{&pol, &av, {}, {}, {}};
			ce.init();
			V2<T> acc;
			if (pol.method == 0) {
				// getAccelerationFromGradient (policy.cpp:78-89)
				V2<T> tg;
				for (int c = 0; c < pol.n_cost; ++c) tg = tg + (T)pol.cost_coeff[c] * ce.gradient(c, a.vel);
				acc = T(-1) * tg;
				if (pol.gradient_iterations >= 2) {
					// descent extension (no counterpart in the reference): each iteration's
					// gradient at the current iterate, the first one being tg
					const T eta = (T)pol.gradient_step * dt;
					V2<T> v = a.vel - tg * eta;
					for (int it = 1; it < pol.gradient_iterations; ++it) {
						V2<T> g;
						for (int c = 0; c < pol.n_cost; ++c) g = g + (T)pol.cost_coeff[c] * ce.gradient(c, v);
						v = v - g * eta;
					}
					acc = (v - a.vel) / dt;
				}
			} else {
				V2<T> best;
				std::vector<std::pair<int, T>> fns;
				for (int c = 0; c < pol.n_cost; ++c) fns.push_back({c, (T)pol.cost_coeff[c]});
				if (pol.method == 2) {
					// getBestVelocityGlobal (policy.cpp:91-100)
					if (pol.n_cost == 1) {
						if (!ce.global_min_closed_form(0, best))
							best = sample_best(ce, global_sampling(), {{0, T(1.0f)}}, rng);
					} else {
						best = sample_best(ce, global_sampling(), fns, rng);
					}
				} else {
					best = sample_best(ce, sampling_of(pol), fns, rng);
				}
				acc = (best - a.vel) / std::max((T)pol.relaxation_time, dt);
			}
			acc = clamp_vec(acc, a.max_accel);
			next = clamp_vec(a.vel + acc * dt, a.max_speed);
		}
		// Policy::ComputeContactForces (policy.cpp:107-141)
		V2<T> contact;
		if (pol.contact_force_scale > 0) {
			for (const auto& nb : nbs) {
				const V2<T> diff = a.pos - nb.pos;
				const T inter = a.radius + ag[nb.idx].radius - diff.mag();
				if (inter > 0) contact = contact + diff.normalized() * ((T)pol.contact_force_scale * inter);
			}
			for (const Seg<T>& sg : obs) {
				const V2<T> diff = a.pos - nearest_point(a.pos, sg.a, sg.b, true);
				const T inter = a.radius - diff.mag();
				if (inter > 0) contact = contact + diff.normalized() * ((T)pol.contact_force_scale * inter);
			}
		}
		// Agent::UpdateVelocityAndPosition (agent.cpp:71-93)
		const T mass = T(80.0f);
		V2<T> vel = next + (contact / mass) * dt;
		V2<T> view(in.viewdir[2 * i], in.viewdir[2 * i + 1]);
		const V2<T> vd = (T(2) * vel + a.pref_vel) / T(3);
		if (vd.sqr() > 0.01) view = vd.normalized();
		V2<T> pos = a.pos + vel * dt;
		// WorldToric::DoStep_MoveAllAgents (worldToric.cpp:81-108)
		if (w.world_type == 1) {
			const T hw = T(0.5f) * (T)w.width, hh = T(0.5f) * (T)w.height;
			T x = pos.x, y = pos.y;
			if (x > hw) x -= (T)w.width; else if (x < -hw) x += (T)w.width;
			if (y > hh) y -= (T)w.height; else if (y < -hh) y += (T)w.height;
			pos = V2<T>(x, y);
		}
		out.pos[2 * i] = (float)pos.x; out.pos[2 * i + 1] = (float)pos.y;
		out.vel[2 * i] = (float)vel.x; out.vel[2 * i + 1] = (float)vel.y;
		if (out.viewdir) { out.viewdir[2 * i] = (float)view.x; out.viewdir[2 * i + 1] = (float)view.y; }
		if (out.pref_vel) { out.pref_vel[2 * i] = (float)a.pref_vel.x; out.pref_vel[2 * i + 1] = (float)a.pref_vel.y; }
		if (out.next_vel) { out.next_vel[2 * i] = (float)next.x; out.next_vel[2 * i + 1] = (float)next.y; }
		if (out.contact) { out.contact[2 * i] = (float)contact.x; out.contact[2 * i + 1] = (float)contact.y; }
		if (out.vel64) { out.vel64[2 * i] = (double)vel.x; out.vel64[2 * i + 1] = (double)vel.y; }
		if (out.next_vel64) { out.next_vel64[2 * i] = (double)next.x; out.next_vel64[2 * i + 1] = (double)next.y; }
		if (out.rng) out.rng[i] = rng;
	}
	return 0;
}

} // namespace

extern "C" int oracle_step(const oracle_world* world, const oracle_policy* policies, int n_policies,
                           int n, const oracle_agents_in* in, oracle_agents_out* out,
                           int precision, int nthreads) {
	if (!world || !policies || !in || !out || n < 0 || n_policies <= 0) return -1;
	if (precision == 1) return step_impl<double>(*world, policies, n_policies, n, *in, *out, nthreads);
	return step_impl<float>(*world, policies, n_policies, n, *in, *out, nthreads);
}

extern "C" int oracle_neighbors_bruteforce(const oracle_world* world, const oracle_policy* policies,
                                           int n, const oracle_agents_in* in, int list_stride,
                                           int32_t* nb_count, int32_t* nb_id, double* nb_d2, int32_t* nb_seg) {
	std::vector<Agent<float>> ag(n);
	for (int i = 0; i < n; ++i) ag[i].pos = V2<float>(in->pos[2 * i], in->pos[2 * i + 1]);
	for (int i = 0; i < n; ++i) {
		std::vector<Phantom<float>> nbs;
		compute_neighbors<float>(*world, nullptr, in->pos, in->id, n, i,
		                         interaction_range(policies[in->policy[i]]), nbs, nullptr, ag);
		nb_count[i] = (int)nbs.size();
		for (int e = 0; e < (int)nbs.size() && e < list_stride; ++e) {
			nb_id[(long long)i * list_stride + e] = in->id[nbs[e].idx];
			nb_d2[(long long)i * list_stride + e] = nbs[e].d2;
			nb_seg[(long long)i * list_stride + e] = nbs[e].seg;
		}
	}
	return 0;
}

// Agent::Agent seeds its engine with the ID (agent.cpp:38). A minstd_rand0
// seeded with s holds x0 = s mod m, or 1 if that is 0; its next output is
// 16807*x0 mod m. Returning x0 lets RecRng(x0) continue the same stream.
extern "C" uint32_t oracle_rng_seed(uint32_t id) {
	const uint32_t m = 2147483647u;
	const uint32_t x0 = id % m;
	return x0 == 0 ? 1u : x0;
}

extern "C" uint32_t oracle_rng_uniform(uint32_t state, float lo, float hi, int count, float* out) {
	RecRng eng(state);
	for (int i = 0; i < count; ++i) {
		std::uniform_real_distribution<float> d(lo, hi);
		out[i] = d(eng);
	}
	return eng.last;
}
