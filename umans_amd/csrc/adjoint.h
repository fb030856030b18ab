// Reverse mode of the smooth gradient step (no counterpart in the reference).
//
// Forward map of one step for an agent i whose policy uses the gradient method with
// force-model costs (PowerLaw, SocialForcesAvoidance, GoalReachingForce):
//   pref_i = ps_i * (g_i - p_i) / |g_i - p_i|                     (agent.cpp:52-59)
//   F_c    = sum_j f_c(i, j)  (PowerLaw / SocialForces pairs)       (ObjectInteractionForces.cpp:26-48)
//   F_GR   = (pref_i - v_i) / max(relax, dt)                        (GoalReachingForce.cpp:26-29)
//   a      = sum_c coeff_c F_c / m_c        (the gradient method, policy.cpp:78-89 with
//                                            ForceBasedFunction.cpp:32-36,49-52)
//   next   = clamp(v_i + clamp(a, a_max) dt, v_max)                 (policy.cpp:72-75)
//   C_i    = sum_j scale (r_i + r_j - |d|) d/|d|, d = p_i - p_j      (policy.cpp:107-141)
//   v_i'   = next + C_i dt / 80,  p_i' = p_i + v_i' dt              (agent.cpp:71-93)
// The reverse kernel applies the transposed Jacobian of this map (the exact derivative
// of each smooth piece, with the neighbour lists held fixed) to the adjoint of (p', v')
// and scatters the contributions to the agent itself, to its neighbours (atomics) and to
// the force parameters of its policy. Adjoint vectors live in upload-row space, so they
// survive the per-step reordering of the agents.
//
// Precision. Every branch of the step — the range filter, the pair models' overlap and
// cut-offs, the view factor, contact, the acceleration and speed clamps — is decided in
// float with the forward kernel's own expressions, so the pieces differentiated are
// those the FP32 step took. The values, Jacobians and adjoint vectors are computed in
// double from the float state: crowds in contact are unstable (a contact force of
// 5000 N/m moves an agent's next position by ~1.6x a perturbation of its position, per
// step), and float rounding of the adjoint excites those growing modes, which over 100
// steps took an FP32 reverse pass 5-15 % away from the exact derivative.
#pragma once
#include "kernels.h"

namespace umans {

// ---------------------------------------------------------------- double vectors
struct D2 {
	double x, y;
};
UD D2 dv2(V2 a) { return D2{(double)a.x, (double)a.y}; }
UD D2 dv2(float2 a) { return D2{(double)a.x, (double)a.y}; }
UD D2 dv2(double2 a) { return D2{a.x, a.y}; }
UD D2 operator+(D2 a, D2 b) { return D2{a.x + b.x, a.y + b.y}; }
UD D2 operator-(D2 a, D2 b) { return D2{a.x - b.x, a.y - b.y}; }
UD D2 operator-(D2 a) { return D2{-a.x, -a.y}; }
UD D2 operator*(double s, D2 v) { return D2{s * v.x, s * v.y}; }
UD D2 operator*(D2 v, double s) { return D2{s * v.x, s * v.y}; }
UD double ddot(D2 a, D2 b) { return a.x * b.x + a.y * b.y; }
UD double dmag(D2 a) { return sqrt(a.x * a.x + a.y * a.y); }

struct AdjArgs {
	State in;                     // sorted input state of the step (as the forward step read it)
	const int* nb_count;          // per sorted agent
	const unsigned* nb;           // [k * stride + e] = sorted index | segment << kSegShift
	int stride;
	const DevPolicy* pol;
	const WorldInfo* worlds;
	int n;
	const double2* adj_pos_out;   // adjoint of the step's output position / velocity (row space)
	const double2* adj_vel_out;
	double2* adj_pos_in;          // accumulated adjoint of the input position / velocity / preferred
	double2* adj_vel_in;          // velocity (row space; cleared before the launch)
	double2* adj_pref;
	// [copy][policy][2]: (V0, sigma) of SocialForces or (k, tau0) of PowerLaw; kStatCopies
	// copies param_stride doubles apart (whole 128-byte lines), each wave adding into the
	// copy of its wave index (atomics on one line are serialised); the host sums them
	double* adj_params;
	int param_stride;
	int param_kind;
	// optional (row space): the clamp branches held fixed, bit 0 acceleration clamped,
	// bit 1 speed clamped (policy.cpp:72-75)
	unsigned char* clamps;
};

UD void atomic_add2(double2* p, D2 v) {
	atomicAdd(&p->x, v.x);
	atomicAdd(&p->y, v.y);
}

// Transposed Jacobian of clamp_vec(w, L) applied to g, on the branch `clamped`.
UD D2 clamp_vjp(D2 w, double L, bool clamped, D2 g) {
	if (!clamped) return g;
	const double m = dmag(w);
	const D2 u = w * (1.0 / m);
	const double gu = ddot(g, u);
	return (L / m) * (g - gu * u);
}

// The branches of SocialForcesAvoidance::ComputeAgentInteractionForce
// (SocialForcesAvoidance.cpp:26-110) with social_force_pair's float expressions: does the
// pair exert a force, and its view scale.
UD bool sf_active(const float* prm, const Self& a, const Nb& o, V2 o_pref, double view_cos, float& scale) {
	const float sdt = prm[1];
	const V2 R = a.pos - o.pos;
	const float magR = mag(R);
	if (magR <= a.radius + o.radius) return false;
	const float vb = mag(o.vel);
	const V2 Eb = normalized(o_pref);
	const V2 V = (vb * sdt) * Eb;
	const V2 R2 = R - V;
	const float magR2 = mag(R2);
	const double s1 = (double)(magR + magR2), s2 = (double)(vb * sdt);
	const float bSq2 = (float)(s1 * s1 - s2 * s2);
	if (bSq2 <= 0.0f) return false;
	const float b = 0.5f * sqrtf(bSq2);
	if (b < 0.001f) return false;
	scale = in_view(a.vel, -R, view_cos) ? 1.0f : 0.5f;
	return true;
}

// The branches of PowerLaw::ComputeAgentInteractionForce (PowerLaw.cpp:28-66) with
// power_law_pair's float expressions.
UD bool pl_active(const float* prm, const Self& a, const Nb& o) {
	const float tau0 = prm[2];
	const float R = o.radius + a.radius;
	const V2 x = a.pos - o.pos;
	const V2 v = a.vel - o.vel;
	const float aa = dot(v, v);
	const float b = -dot(x, v);
	const float c = dot(x, x) - R * R;
	if (c <= 0.0f) return false;
	const float d = b * b - aa * c;
	if (d <= 0.0f) return false;
	const float tau = (b - sqrtf(d)) / aa;
	return !(tau < 0.001f) && !(tau > tau0);
}

// SocialForces pair force F = g(R, V; V0, sigma) R of an active pair in double, R the
// agent's position minus the neighbour's, vj the neighbour's velocity, prefj its
// preferred velocity (SocialForcesAvoidance.cpp:26-110; the directional derivative
// d_magR2 = (|R| - R.V / |R|) / |R - V| is the reference's 2|R| - 2|V| cos(R, V) over 2|R - V|).
struct SFd {
	D2 R, V, R2, Eb;
	double magR, magR2, vb, q, S, P, b, E, dB2, dm, g, pref_mag;
};
UD void sf_eval(const float* prm, D2 R, D2 vj, D2 prefj, double scale, SFd& s) {
	const double sdt = prm[1], V0 = prm[2], sigma = prm[3];
	s.R = R;
	s.magR = dmag(R);
	s.vb = dmag(vj);
	s.pref_mag = dmag(prefj);
	s.Eb = s.pref_mag > 0.0 ? prefj * (1.0 / s.pref_mag) : D2{0.0, 0.0};
	s.q = s.vb * sdt;
	s.V = s.q * s.Eb;
	s.R2 = R - s.V;
	s.magR2 = dmag(s.R2);
	s.S = s.magR + s.magR2;
	s.b = 0.5 * sqrt(s.S * s.S - s.q * s.q);
	s.P = ddot(R, s.V);
	s.dm = (s.magR - s.P / s.magR) / s.magR2;
	s.dB2 = 2.0 * s.S * (1.0 + s.dm);
	s.E = exp(-s.b / sigma);
	s.g = V0 / sigma * s.E * (s.dB2 / (8.0 * s.b)) * scale;
}
// Adjoint of F = g R for one pair, given Fbar.
UD void sf_vjp(const float* prm, const SFd& s, D2 vj, D2 Fbar, D2& Rbar, D2& vjbar, D2& prefjbar, double& V0bar,
               double& sigbar) {
	const double sdt = prm[1], V0 = prm[2], sigma = prm[3];
	const double gbar = ddot(Fbar, s.R);
	Rbar = s.g * Fbar;
	V0bar = gbar * s.g / V0;
	sigbar = gbar * s.g * (s.b / (sigma * sigma) - 1.0 / sigma);
	// g = K E dB2 / (8 b), K = V0 scale / sigma
	const double bbar = gbar * s.g * (-1.0 / sigma - 1.0 / s.b);
	const double dB2bar = gbar * s.g / s.dB2;
	const double bSq2bar = bbar / (8.0 * s.b);
	const double Sbar = 2.0 * s.S * bSq2bar + 2.0 * (1.0 + s.dm) * dB2bar;
	double qbar = -2.0 * s.q * bSq2bar;
	const double dmbar = 2.0 * s.S * dB2bar;
	const double magRbar = dmbar * (1.0 + s.P / (s.magR * s.magR)) / s.magR2 + Sbar;
	const double Pbar = -dmbar / (s.magR * s.magR2);
	const double magR2bar = -dmbar * s.dm / s.magR2 + Sbar;
	D2 Vbar = Pbar * s.R;
	Rbar = Rbar + Pbar * s.V;
	const D2 R2bar = (magR2bar / s.magR2) * s.R2;
	Rbar = Rbar + R2bar;
	Vbar = Vbar - R2bar;
	Rbar = Rbar + (magRbar / s.magR) * s.R;
	// V = q Eb, q = vb * sdt, vb = |v_j|, Eb = pref_j / |pref_j|
	qbar += ddot(Vbar, s.Eb);
	const D2 Ebbar = s.q * Vbar;
	const double vbbar = sdt * qbar;
	vjbar = (s.vb > 0.0) ? (vbbar / s.vb) * vj : D2{0.0, 0.0};
	if (s.pref_mag > 0.0) {
		const double eb = ddot(Ebbar, s.Eb);
		prefjbar = (1.0 / s.pref_mag) * (Ebbar - eb * s.Eb);
	} else {
		prefjbar = D2{0.0, 0.0};
	}
}

// PowerLaw pair force of an active pair in double (PowerLaw.cpp:28-66), x the agent's
// position minus the neighbour's, v its velocity minus the neighbour's.
struct PLd {
	D2 x, v, comp2, w;
	double a, b, c, sd, tau, E, H, comp1;
};
UD void pl_eval(const float* prm, D2 x, D2 v, double R, PLd& s) {
	const double k = prm[1], tau0 = prm[2];
	s.x = x;
	s.v = v;
	s.a = ddot(v, v);
	s.b = -ddot(x, v);
	s.c = ddot(x, x) - R * R;
	s.sd = sqrt(s.b * s.b - s.a * s.c);
	s.tau = (s.b - s.sd) / s.a;
	s.E = exp(-s.tau / tau0);
	s.H = 2.0 / s.tau + 1.0 / tau0;
	s.comp1 = -k * s.E * s.H / (s.a * s.tau * s.tau);
	s.w = s.a * x + s.b * v;
	s.comp2 = v - s.w * (1.0 / s.sd);
}
UD void pl_vjp(const float* prm, const PLd& s, D2 Fbar, D2& xbar, D2& vbar, double& kbar, double& tau0bar) {
	const double k = prm[1], tau0 = prm[2];
	const double c1bar = ddot(Fbar, s.comp2);
	const D2 c2bar = s.comp1 * Fbar;
	vbar = c2bar;
	const D2 wbar = (-1.0 / s.sd) * c2bar;
	double sdbar = ddot(c2bar, s.w) / (s.sd * s.sd);
	double abar = ddot(wbar, s.x);
	xbar = s.a * wbar;
	double bbar = ddot(wbar, s.v);
	vbar = vbar + s.b * wbar;
	kbar = c1bar * s.comp1 / k;
	const double dc1_dtau = s.comp1 * (-1.0 / tau0 + (-2.0 / (s.tau * s.tau)) / s.H - 2.0 / s.tau);
	tau0bar = c1bar * s.comp1 * (s.tau / (tau0 * tau0) + (-1.0 / (tau0 * tau0)) / s.H);
	const double taubar = c1bar * dc1_dtau;
	abar += -c1bar * s.comp1 / s.a;
	bbar += taubar / s.a;
	sdbar += -taubar / s.a;
	abar += -taubar * s.tau / s.a;
	const double dbar = sdbar / (2.0 * s.sd);
	bbar += 2.0 * s.b * dbar;
	abar += -s.c * dbar;
	const double cbar = -s.a * dbar;
	xbar = xbar + (2.0 * cbar) * s.x;
	xbar = xbar - bbar * s.v;
	vbar = vbar - bbar * s.x;
	vbar = vbar + (2.0 * abar) * s.v;
}

UD unsigned row_of(const State& S, int j) { return (unsigned)S.meta[j].z; }

// The neighbour of list entry u as the forward step placed it, in double: its position
// plus the toric offset of its segment (worldToric.cpp:36,47), without the float rounding.
UD D2 phantom_pos(const NbList& L, unsigned u) {
	const float2 pj = L.pos[u & kIndexMask];
	const unsigned seg = u >> kSegShift;
	D2 p{(double)pj.x, (double)pj.y};
	if (seg == 1) p.x -= (double)L.W;
	else if (seg == 2) p.x += (double)L.W;
	else if (seg == 3) p.y -= (double)L.H;
	else if (seg == 4) p.y += (double)L.H;
	return p;
}

// Every lane gets the sum over the wave (a butterfly: each lane adds the same operands,
// so the lanes agree bit for bit).
UD double wave_sum(double v) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}
UD float lane_value(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// The longest list the adjoint keeps (the step's saved copy of a list is at most this
// long), and the entries per lane: entry e of an agent's list is handled by lane e mod 64.
constexpr int kAdjListMax = 128;
constexpr int kAdjSlots = kAdjListMax / 64;

// One wave per agent: the lanes share the agent's neighbour list (entries lane,
// lane + 64), so a 10k-agent crowd with ~80 neighbours each fills 10k waves instead of
// 157, and each lane evaluates at most kAdjSlots pairs per pass. The float pass that
// decides the clamp branches adds the pair forces in list order, as the forward step
// does; the double values are summed over the wave.
__global__ __launch_bounds__(256) void k_step_adjoint(AdjArgs A) {
	const int lane = threadIdx.x & 63;
	const int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
	if (k >= A.n) return;                       // whole waves
	const State& S = A.in;
	const int4 meta = S.meta[k];
	const int pi = meta.y & kPolicyMask;
	const WorldInfo wi = A.worlds[meta.y >> kPolicyBits];
	const DevPolicy& P = A.pol[pi];
	const float4 prm = S.prm[k];
	Self a;
	a.pos = v2(S.pos[k]);
	a.vel = v2(S.vel[k]);
	a.pref = v2(S.pref[k]);
	a.radius = prm.x; a.pref_speed = prm.y; a.max_speed = prm.z; a.max_accel = prm.w;
	a.dt = wi.dt;
	a.relaxation = P.relaxation;
	const int cnt = A.nb_count[k];             // at most A.stride <= kUncappedMax
	const NbList L{A.nb + (size_t)k * A.stride, 1, cnt, S.pos, S.vel, S.prm, S.pref, wi.W, wi.H, a.pos};
	const unsigned row = (unsigned)meta.z;
	const double dt = (double)a.dt;
	const D2 pi_d = dv2(a.pos), vi_d = dv2(a.vel);

	// this lane's entries
	bool has[kAdjSlots];
	Nb o[kAdjSlots];
	D2 opos[kAdjSlots];
	unsigned orow[kAdjSlots];
#pragma unroll
	for (int s = 0; s < kAdjSlots; ++s) {
		const int e = lane + 64 * s;
		has[s] = e < cnt;
		if (has[s]) {
			o[s] = L.get(e);
			opos[s] = phantom_pos(L, L.raw(e));
			orow[s] = row_of(S, o[s].j);
		}
	}

	const D2 pbar_out = dv2(A.adj_pos_out[row]);
	const D2 vbar_out = dv2(A.adj_vel_out[row]);
	// p' = p + v' dt; v' = next + C dt / 80
	const D2 vtot = vbar_out + pbar_out * dt;   // adjoint of v'
	const D2 nextbar = vtot;
	const D2 Cbar = vtot * (dt / 80.0);
	// lane-local parts of the agent's own adjoints (summed over the wave at the end) and
	// the wave-uniform parts (added once)
	D2 pbar_l{0.0, 0.0}, vbar_l{0.0, 0.0};
	double p0bar_l = 0.0, p1bar_l = 0.0;
	D2 pbar_u = pbar_out, vbar_u{0.0, 0.0}, prefbar_u{0.0, 0.0};

	// contact forces (policy.cpp:107-141): C_j = f (R - s) u, u = d / s, s = |d|
	if (P.contact_scale > 0.0f) {
		const double f = (double)P.contact_scale;
#pragma unroll
		for (int s = 0; s < kAdjSlots; ++s) {
			if (!has[s]) continue;
			const V2 d32 = a.pos - o[s].pos;
			const float s32 = mag(d32);
			const float inter32 = a.radius + o[s].radius - s32;
			if (!(inter32 > 0.0f) || !(s32 > 0.0f)) continue;
			const D2 d = pi_d - opos[s];
			const double sd = dmag(d);
			const D2 u = d * (1.0 / sd);
			const double inter = (double)a.radius + (double)o[s].radius - sd;
			const double cu = ddot(Cbar, u);
			// J = f [ (R - s)/s (I - u u^T) - u u^T ], symmetric
			const D2 jd = f * ((inter / sd) * (Cbar - cu * u) - cu * u);
			pbar_l = pbar_l + jd;
			atomic_add2(&A.adj_pos_in[orow[s]], -jd);
		}
	}

	if (!is_zero(a.pref) && P.mode == MODE_GRADIENT) {
		// the forward step's float acceleration and its clamp decisions (k_step, force
		// gradient class: pair forces added in list order, target = v + F / m dt,
		// gradient = (v - target) / dt), and which pairs exert a force
		V2 F32[4];
		unsigned active = 0;                    // bit c * kAdjSlots + s
		float sfscale[4][kAdjSlots];
		for (int c = 0; c < P.n_cost; ++c) {
			F32[c] = V2{0.0f, 0.0f};
			const int kind = P.kind[c];
			if (kind == KIND_GOAL_REACHING) continue;
			const float rsq = P.range[c] * P.range[c];
			V2 f[kAdjSlots];
#pragma unroll
			for (int s = 0; s < kAdjSlots; ++s) {
				f[s] = V2{0.0f, 0.0f};
				sfscale[c][s] = 1.0f;
				if (!has[s] || !(o[s].dsq <= rsq)) continue;
				bool act;
				if (kind == KIND_POWER_LAW) {
					f[s] = power_law_pair(P.params[c], a, o[s]);
					act = pl_active(P.params[c], a, o[s]);
				} else {
					const V2 opref = L.pref_of(o[s].j);
					f[s] = social_force_pair(P.params[c], a, o[s], opref, P.view_cos);
					act = sf_active(P.params[c], a, o[s], opref, P.view_cos, sfscale[c][s]);
				}
				if (act) active |= 1u << (c * kAdjSlots + s);
			}
			// list order (a pair outside the range adds +0, which changes no float sum
			// that starts at +0)
			V2 F{0.0f, 0.0f};
#pragma unroll
			for (int s = 0; s < kAdjSlots; ++s) {
				const int m = min(64, cnt - 64 * s);
				for (int l = 0; l < m; ++l) F = F + V2{lane_value(f[s].x, l), lane_value(f[s].y, l)};
			}
			F32[c] = F;
		}
		V2 tg32{0.0f, 0.0f};
		for (int c = 0; c < P.n_cost; ++c) {
			const V2 f = (P.kind[c] == KIND_GOAL_REACHING) ? (a.pref - a.vel) / smax(a.relaxation, a.dt)
			                                                : F32[c] + V2{0.0f, 0.0f};
			const V2 target = a.vel + f / P.params[c][0] * a.dt;
			tg32 = tg32 + P.coeff[c] * ((a.vel - target) / a.dt);
		}
		const V2 acc32 = -1.0f * tg32;
		const bool acc_clamped = !(mag(acc32) <= a.max_accel);
		const V2 w32 = a.vel + clamp_vec(acc32, a.max_accel) * a.dt;
		const bool speed_clamped = !(mag(w32) <= a.max_speed);
		if (A.clamps && lane == 0) A.clamps[row] = (unsigned char)((acc_clamped ? 1 : 0) | (speed_clamped ? 2 : 0));

		// the same map in double: forces of the active pairs, acceleration, clamps
		const double tr = (double)smax(a.relaxation, a.dt);
		D2 acc{0.0, 0.0};
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			D2 Fc{0.0, 0.0};
			if (kind == KIND_GOAL_REACHING) {
				Fc = (dv2(a.pref) - vi_d) * (1.0 / tr);
			} else {
				D2 part{0.0, 0.0};
#pragma unroll
				for (int s = 0; s < kAdjSlots; ++s) {
					if (!(active >> (c * kAdjSlots + s) & 1u)) continue;
					if (kind == KIND_SOCIAL_FORCES) {
						SFd q;
						sf_eval(P.params[c], pi_d - opos[s], dv2(o[s].vel), dv2(L.pref_of(o[s].j)), sfscale[c][s], q);
						part = part + q.g * q.R;
					} else {
						PLd q;
						pl_eval(P.params[c], pi_d - opos[s], vi_d - dv2(o[s].vel),
						        (double)o[s].radius + (double)a.radius, q);
						part = part + q.comp1 * q.comp2;
					}
				}
				Fc = D2{wave_sum(part.x), wave_sum(part.y)};
			}
			acc = acc + ((double)P.coeff[c] / (double)P.params[c][0]) * Fc;
		}
		const double amax = (double)a.max_accel;
		const D2 accc = acc_clamped ? acc * (amax / dmag(acc)) : acc;
		const D2 w = vi_d + accc * dt;
		// reverse
		const D2 wbar = clamp_vjp(w, (double)a.max_speed, speed_clamped, nextbar);
		vbar_u = vbar_u + wbar;
		const D2 accbar = clamp_vjp(acc, amax, acc_clamped, wbar * dt);
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			const D2 Fbar = ((double)P.coeff[c] / (double)P.params[c][0]) * accbar;
			if (kind == KIND_GOAL_REACHING) {
				prefbar_u = prefbar_u + Fbar * (1.0 / tr);
				vbar_u = vbar_u - Fbar * (1.0 / tr);
				continue;
			}
#pragma unroll
			for (int s = 0; s < kAdjSlots; ++s) {
				if (!(active >> (c * kAdjSlots + s) & 1u)) continue;
				if (kind == KIND_SOCIAL_FORCES) {
					SFd q;
					const D2 vj = dv2(o[s].vel);
					sf_eval(P.params[c], pi_d - opos[s], vj, dv2(L.pref_of(o[s].j)), sfscale[c][s], q);
					D2 Rbar, vjbar, prefjbar;
					double V0b, sgb;
					sf_vjp(P.params[c], q, vj, Fbar, Rbar, vjbar, prefjbar, V0b, sgb);
					pbar_l = pbar_l + Rbar;
					atomic_add2(&A.adj_pos_in[orow[s]], -Rbar);
					atomic_add2(&A.adj_vel_in[orow[s]], vjbar);
					atomic_add2(&A.adj_pref[orow[s]], prefjbar);
					if (A.param_kind == KIND_SOCIAL_FORCES) { p0bar_l += V0b; p1bar_l += sgb; }
				} else {
					PLd q;
					pl_eval(P.params[c], pi_d - opos[s], vi_d - dv2(o[s].vel), (double)o[s].radius + (double)a.radius,
					        q);
					D2 xbar, vrbar;
					double kb, t0b;
					pl_vjp(P.params[c], q, Fbar, xbar, vrbar, kb, t0b);
					pbar_l = pbar_l + xbar;
					vbar_l = vbar_l + vrbar;
					atomic_add2(&A.adj_pos_in[orow[s]], -xbar);
					atomic_add2(&A.adj_vel_in[orow[s]], -vrbar);
					if (A.param_kind == KIND_POWER_LAW) { p0bar_l += kb; p1bar_l += t0b; }
				}
			}
		}
	}
	const D2 pbar = pbar_u + D2{wave_sum(pbar_l.x), wave_sum(pbar_l.y)};
	const D2 vbar = vbar_u + D2{wave_sum(vbar_l.x), wave_sum(vbar_l.y)};
	const double p0bar = wave_sum(p0bar_l), p1bar = wave_sum(p1bar_l);
	if (lane == 0) {
		if (p0bar != 0.0 || p1bar != 0.0) {
			double* pc = A.adj_params + (size_t)stat_copy() * A.param_stride;
			if (p0bar != 0.0) atomicAdd(&pc[2 * pi + 0], p0bar);
			if (p1bar != 0.0) atomicAdd(&pc[2 * pi + 1], p1bar);
		}
		atomic_add2(&A.adj_pos_in[row], pbar);
		atomic_add2(&A.adj_vel_in[row], vbar);
		atomic_add2(&A.adj_pref[row], prefbar_u);
	}
}

// Minimum-image difference p - t in a toric world, plain difference otherwise.
UD D2 loss_diff(const WorldInfo& wi, float2 p, float2 t) {
	double dx = (double)p.x - (double)t.x, dy = (double)p.y - (double)t.y;
	if (wi.world_type == 1) {
		if (dx > 0.5 * (double)wi.W) dx -= (double)wi.W; else if (dx < -0.5 * (double)wi.W) dx += (double)wi.W;
		if (dy > 0.5 * (double)wi.H) dy -= (double)wi.H; else if (dy < -0.5 * (double)wi.H) dy += (double)wi.H;
	}
	return D2{dx, dy};
}

// After the reverse step of step t, per agent of its sorted input state `in`:
//  * the adjoint of the preferred velocity, pref = ps (g - p) / |g - p| unless the goal
//    is reached (agent.cpp:52-59, decided as k_reorder decides it), moved to the position;
//  * with `target` (the targets of step t - 1), the loss gradient scale * 2 (p - target)
//    of step t - 1's output added: that output is step t's input state;
//  * the preferred-velocity adjoint and the buffers clr0, clr1 (the next reverse step
//    accumulates into them) cleared.
// Every row is one agent's, handled by one thread.
__global__ __launch_bounds__(256) void k_adj_link(State in, int n, const WorldInfo* worlds, double2* adj_pref,
                                                 double2* adj_pos, const float2* target, double scale, double2* clr0,
                                                 double2* clr1) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const int4 meta = in.meta[k];
	const unsigned row = (unsigned)meta.z;
	const float4 prm = in.prm[k];
	const float2 p32 = in.pos[k];
	D2 p = dv2(adj_pos[row]);
	if (!(sqr_mag(v2(in.goal[k]) - v2(p32)) <= prm.x * prm.x)) {
		const D2 pb = dv2(adj_pref[row]);
		const D2 d = dv2(in.goal[k]) - dv2(p32);
		const double m = dmag(d);
		const D2 u = d * (1.0 / m);
		const double pu = ddot(pb, u);
		const double s = (double)prm.y / m;
		// d pref / d p = -(ps / m) (I - u u^T)
		p = p - s * (pb - pu * u);
	}
	if (target) {
		const D2 d = loss_diff(worlds[meta.y >> kPolicyBits], p32, target[row]);
		p = p + (scale * 2.0) * d;
	}
	adj_pos[row] = make_double2(p.x, p.y);
	adj_pref[row] = make_double2(0.0, 0.0);
	clr0[row] = make_double2(0.0, 0.0);
	clr1[row] = make_double2(0.0, 0.0);
}

// Loss of one step's output: sum over agents of |p - target|^2 with the minimum-image
// difference in toric worlds; optionally its gradient scale * 2 (p - target) added to
// grad (row space).
__global__ __launch_bounds__(256) void k_loss(State out, int n, const WorldInfo* worlds, const float2* target,
                                             double* loss, double2* grad, double scale) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	double l = 0.0;
	if (k < n) {
		const int4 meta = out.meta[k];
		const unsigned row = (unsigned)meta.z;
		const D2 d = loss_diff(worlds[meta.y >> kPolicyBits], out.pos[k], target[row]);
		l = d.x * d.x + d.y * d.y;
		if (grad) {
			double2 g = grad[row];
			g.x += scale * 2.0 * d.x;
			g.y += scale * 2.0 * d.y;
			grad[row] = g;
		}
	}
	for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
	if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

__global__ void k_fill_f(float* p, long long n, float v) {
	const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// the adjoint vectors handed out in float
__global__ void k_to_float2(const double2* x, float2* y, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) y[i] = make_float2((float)x[i].x, (float)x[i].y);
}

} // namespace umans
