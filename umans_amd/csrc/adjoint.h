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

__global__ __launch_bounds__(128) void k_step_adjoint(AdjArgs A) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= A.n) return;
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
	const int cnt = A.nb_count[k];
	const NbList L{A.nb + (size_t)k * A.stride, 1, cnt, S.pos, S.vel, S.prm, S.pref, wi.W, wi.H, a.pos};
	const unsigned row = (unsigned)meta.z;
	const double dt = (double)a.dt;
	const D2 pi_d = dv2(a.pos), vi_d = dv2(a.vel);

	const D2 pbar_out = dv2(A.adj_pos_out[row]);
	const D2 vbar_out = dv2(A.adj_vel_out[row]);
	// p' = p + v' dt
	D2 pbar = pbar_out;
	const D2 vtot = vbar_out + pbar_out * dt;   // adjoint of v'
	D2 vbar{0.0, 0.0}, prefbar{0.0, 0.0};
	// v' = next + C dt / 80
	const D2 nextbar = vtot;
	const D2 Cbar = vtot * (dt / 80.0);

	// contact forces (policy.cpp:107-141): C_j = f (R - s) u, u = d / s, s = |d|
	if (P.contact_scale > 0.0f) {
		const double f = (double)P.contact_scale;
		for (int e = 0; e < cnt; ++e) {
			const Nb o = L.get(e);
			const V2 d32 = a.pos - o.pos;
			const float s32 = mag(d32);
			const float inter32 = a.radius + o.radius - s32;
			if (!(inter32 > 0.0f) || !(s32 > 0.0f)) continue;
			const D2 d = pi_d - phantom_pos(L, L.raw(e));
			const double s = dmag(d);
			const D2 u = d * (1.0 / s);
			const double inter = (double)a.radius + (double)o.radius - s;
			const double cu = ddot(Cbar, u);
			// J = f [ (R - s)/s (I - u u^T) - u u^T ], symmetric
			const D2 jd = f * ((inter / s) * (Cbar - cu * u) - cu * u);
			pbar = pbar + jd;
			atomic_add2(&A.adj_pos_in[row_of(S, o.j)], -jd);
		}
	}

	if (!is_zero(a.pref) && P.mode == MODE_GRADIENT) {
		// the forward step's float acceleration and its clamp decisions (k_step, force
		// gradient class: target = v + F / m dt, gradient = (v - target) / dt)
		V2 F32[4];
		for (int c = 0; c < P.n_cost; ++c) F32[c] = V2{0.0f, 0.0f};
		for (int e = 0; e < cnt; ++e) {
			const Nb o = L.get(e);
			for (int c = 0; c < P.n_cost; ++c) {
				const int kind = P.kind[c];
				if (kind == KIND_GOAL_REACHING || !(o.dsq <= P.range[c] * P.range[c])) continue;
				F32[c] = F32[c] + ((kind == KIND_POWER_LAW) ? power_law_pair(P.params[c], a, o)
				                                            : social_force_pair(P.params[c], a, o, L.pref_of(o.j),
				                                                                P.view_cos));
			}
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
		if (A.clamps) A.clamps[row] = (unsigned char)((acc_clamped ? 1 : 0) | (speed_clamped ? 2 : 0));

		// the same map in double: forces of the active pairs, acceleration, clamps
		const double tr = (double)smax(a.relaxation, a.dt);
		D2 acc{0.0, 0.0};
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			D2 Fc{0.0, 0.0};
			if (kind == KIND_GOAL_REACHING) {
				Fc = (dv2(a.pref) - vi_d) * (1.0 / tr);
			} else {
				for (int e = 0; e < cnt; ++e) {
					const Nb o = L.get(e);
					if (!(o.dsq <= P.range[c] * P.range[c])) continue;
					const unsigned u = L.raw(e);
					if (kind == KIND_SOCIAL_FORCES) {
						float scale;
						if (!sf_active(P.params[c], a, o, L.pref_of(o.j), P.view_cos, scale)) continue;
						SFd s;
						sf_eval(P.params[c], pi_d - phantom_pos(L, u), dv2(o.vel), dv2(L.pref_of(o.j)), scale, s);
						Fc = Fc + s.g * s.R;
					} else if (kind == KIND_POWER_LAW) {
						if (!pl_active(P.params[c], a, o)) continue;
						PLd s;
						pl_eval(P.params[c], pi_d - phantom_pos(L, u), vi_d - dv2(o.vel),
						        (double)o.radius + (double)a.radius, s);
						Fc = Fc + s.comp1 * s.comp2;
					}
				}
			}
			acc = acc + ((double)P.coeff[c] / (double)P.params[c][0]) * Fc;
		}
		const double amax = (double)a.max_accel;
		const D2 accc = acc_clamped ? acc * (amax / dmag(acc)) : acc;
		const D2 w = vi_d + accc * dt;
		// reverse
		const D2 wbar = clamp_vjp(w, (double)a.max_speed, speed_clamped, nextbar);
		vbar = vbar + wbar;
		const D2 accbar = clamp_vjp(acc, amax, acc_clamped, wbar * dt);
		double p0bar = 0.0, p1bar = 0.0;
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			const D2 Fbar = ((double)P.coeff[c] / (double)P.params[c][0]) * accbar;
			if (kind == KIND_GOAL_REACHING) {
				prefbar = prefbar + Fbar * (1.0 / tr);
				vbar = vbar - Fbar * (1.0 / tr);
				continue;
			}
			for (int e = 0; e < cnt; ++e) {
				const Nb o = L.get(e);
				if (!(o.dsq <= P.range[c] * P.range[c])) continue;
				const unsigned u = L.raw(e);
				const unsigned rj = row_of(S, o.j);
				if (kind == KIND_SOCIAL_FORCES) {
					float scale;
					if (!sf_active(P.params[c], a, o, L.pref_of(o.j), P.view_cos, scale)) continue;
					SFd s;
					const D2 vj = dv2(o.vel);
					sf_eval(P.params[c], pi_d - phantom_pos(L, u), vj, dv2(L.pref_of(o.j)), scale, s);
					D2 Rbar, vjbar, prefjbar;
					double V0b, sgb;
					sf_vjp(P.params[c], s, vj, Fbar, Rbar, vjbar, prefjbar, V0b, sgb);
					pbar = pbar + Rbar;
					atomic_add2(&A.adj_pos_in[rj], -Rbar);
					atomic_add2(&A.adj_vel_in[rj], vjbar);
					atomic_add2(&A.adj_pref[rj], prefjbar);
					if (A.param_kind == KIND_SOCIAL_FORCES) { p0bar += V0b; p1bar += sgb; }
				} else if (kind == KIND_POWER_LAW) {
					if (!pl_active(P.params[c], a, o)) continue;
					PLd s;
					pl_eval(P.params[c], pi_d - phantom_pos(L, u), vi_d - dv2(o.vel),
					        (double)o.radius + (double)a.radius, s);
					D2 xbar, vrbar;
					double kb, t0b;
					pl_vjp(P.params[c], s, Fbar, xbar, vrbar, kb, t0b);
					pbar = pbar + xbar;
					vbar = vbar + vrbar;
					atomic_add2(&A.adj_pos_in[rj], -xbar);
					atomic_add2(&A.adj_vel_in[rj], -vrbar);
					if (A.param_kind == KIND_POWER_LAW) { p0bar += kb; p1bar += t0b; }
				}
			}
		}
		double* pc = A.adj_params + (size_t)stat_copy() * A.param_stride;
		if (p0bar != 0.0) atomicAdd(&pc[2 * pi + 0], p0bar);
		if (p1bar != 0.0) atomicAdd(&pc[2 * pi + 1], p1bar);
	}
	atomic_add2(&A.adj_pos_in[row], pbar);
	atomic_add2(&A.adj_vel_in[row], vbar);
	atomic_add2(&A.adj_pref[row], prefbar);
}

// pref = ps (g - p) / |g - p| unless the goal is reached: adjoint to the position.
__global__ __launch_bounds__(256) void k_pref_adjoint(State in, int n, const double2* adj_pref, double2* adj_pos) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const unsigned row = (unsigned)in.meta[k].z;
	const float4 prm = in.prm[k];
	// the forward step's decision (k_reorder)
	if (sqr_mag(v2(in.goal[k]) - v2(in.pos[k])) <= prm.x * prm.x) return;
	const D2 pb = dv2(adj_pref[row]);
	const D2 d = dv2(in.goal[k]) - dv2(in.pos[k]);
	const double m = dmag(d);
	const D2 u = d * (1.0 / m);
	const double pu = ddot(pb, u);
	const double s = (double)prm.y / m;
	// d pref / d p = -(ps / m) (I - u u^T)
	atomic_add2(&adj_pos[row], -s * (pb - pu * u));
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
		const WorldInfo wi = worlds[meta.y >> kPolicyBits];
		const unsigned row = (unsigned)meta.z;
		const float2 p = out.pos[k], t = target[row];
		double dx = (double)p.x - (double)t.x, dy = (double)p.y - (double)t.y;
		if (wi.world_type == 1) {
			if (dx > 0.5 * (double)wi.W) dx -= (double)wi.W; else if (dx < -0.5 * (double)wi.W) dx += (double)wi.W;
			if (dy > 0.5 * (double)wi.H) dy -= (double)wi.H; else if (dy < -0.5 * (double)wi.H) dy += (double)wi.H;
		}
		l = dx * dx + dy * dy;
		if (grad) {
			double2 g = grad[row];
			g.x += scale * 2.0 * dx;
			g.y += scale * 2.0 * dy;
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

// y += x (adjoint vectors)
__global__ void k_axpy2(double2* y, const double2* x, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) { y[i].x += x[i].x; y[i].y += x[i].y; }
}

// the adjoint vectors handed out in float
__global__ void k_to_float2(const double2* x, float2* y, int n) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) y[i] = make_float2((float)x[i].x, (float)x[i].y);
}

} // namespace umans
