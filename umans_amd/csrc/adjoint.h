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
// of each smooth piece; the neighbour lists, filters, clamp branches and view factors
// of the forward step are held fixed) to the adjoint of (p', v') and scatters the
// contributions to the agent itself, to its neighbours (atomics) and to the force
// parameters of its policy. Adjoint vectors live in upload-row space, so they survive
// the per-step reordering of the agents.
#pragma once
#include "kernels.h"

namespace umans {

struct AdjArgs {
	State in;                     // sorted input state of the step (as the forward step read it)
	const int* nb_count;          // per sorted agent
	const unsigned* nb;           // [k * stride + e] = sorted index | segment << kSegShift
	int stride;
	const DevPolicy* pol;
	const WorldInfo* worlds;
	int n;
	const float2* adj_pos_out;    // adjoint of the step's output position / velocity (row space)
	const float2* adj_vel_out;
	float2* adj_pos_in;           // accumulated adjoint of the input position / velocity / preferred
	float2* adj_vel_in;           // velocity (row space; cleared before the launch)
	float2* adj_pref;
	// [copy][policy][2]: (V0, sigma) of SocialForces or (k, tau0) of PowerLaw; kStatCopies
	// copies param_stride floats apart (whole 128-byte lines), each wave adding into the
	// copy of its wave index (float atomics on one line are serialised); the host sums them
	float* adj_params;
	int param_stride;
	int param_kind;
};

UD void atomic_add2(float2* p, V2 v) {
	atomicAdd(&p->x, v.x);
	atomicAdd(&p->y, v.y);
}

// Transposed Jacobian of clamp_vec(w, L) applied to g.
UD V2 clamp_vjp(V2 w, float L, V2 g) {
	const float m = mag(w);
	if (m <= L) return g;
	const V2 u = w / m;
	const float s = L / m;
	const float gu = dot(g, u);
	return V2{s * (g.x - gu * u.x), s * (g.y - gu * u.y)};
}

// Forward SocialForces pair force with the intermediates the reverse pass needs
// (SocialForcesAvoidance.cpp:26-110). Returns false when the pair exerts no force.
struct SFPair {
	V2 R, V, R2, Eb;
	float magR, magR2, vb, q, S, P, b, E, dB2, dm, g, pref_mag;
};
UD bool sf_forward(const float* prm, const Self& a, const Nb& o, V2 o_pref, double view_cos, SFPair& s) {
	const float sdt = prm[1], V0 = prm[2], sigma = prm[3];
	s.R = a.pos - o.pos;
	s.magR = mag(s.R);
	if (s.magR <= a.radius + o.radius) return false;
	s.vb = mag(o.vel);
	s.pref_mag = mag(o_pref);
	s.Eb = normalized(o_pref);
	s.q = s.vb * sdt;
	s.V = s.q * s.Eb;
	s.R2 = s.R - s.V;
	s.magR2 = mag(s.R2);
	s.S = s.magR + s.magR2;
	const float bSq2 = s.S * s.S - s.q * s.q;
	if (bSq2 <= 0.0f) return false;
	s.b = 0.5f * sqrtf(bSq2);
	if (s.b < 0.001f) return false;
	s.P = dot(s.R, s.V);
	s.dm = (s.magR - s.P / s.magR) / s.magR2;
	s.dB2 = 2.0f * s.S * (1.0f + s.dm);
	s.E = expf(-s.b / sigma);
	const float scale = in_view(a.vel, -s.R, view_cos) ? 1.0f : 0.5f;
	s.g = V0 / sigma * s.E * (s.dB2 / (8.0f * s.b)) * scale;   // F = g R
	return true;
}

// Adjoint of F = g(R, V; V0, sigma) R for one pair, given Fbar.
UD void sf_vjp(const float* prm, const SFPair& s, V2 Fbar, V2& Rbar, V2& vjbar, V2& prefjbar, V2& vj,
               float& V0bar, float& sigbar) {
	const float sdt = prm[1], V0 = prm[2], sigma = prm[3];
	const float gbar = dot(Fbar, s.R);
	Rbar = s.g * Fbar;
	V0bar = gbar * s.g / V0;
	sigbar = gbar * s.g * (s.b / (sigma * sigma) - 1.0f / sigma);
	// g = K E dB2 / (8 b), K = V0 scale / sigma
	const float bbar = gbar * s.g * (-1.0f / sigma - 1.0f / s.b);
	const float dB2bar = gbar * s.g / s.dB2;
	const float bSq2bar = bbar / (8.0f * s.b);
	float Sbar = 2.0f * s.S * bSq2bar + 2.0f * (1.0f + s.dm) * dB2bar;
	float qbar = -2.0f * s.q * bSq2bar;
	const float dmbar = 2.0f * s.S * dB2bar;
	float magRbar = dmbar * (1.0f + s.P / (s.magR * s.magR)) / s.magR2 + Sbar;
	const float Pbar = -dmbar / (s.magR * s.magR2);
	const float magR2bar = -dmbar * s.dm / s.magR2 + Sbar;
	V2 Vbar = Pbar * s.R;
	Rbar = Rbar + Pbar * s.V;
	const V2 R2bar = (magR2bar / s.magR2) * s.R2;
	Rbar = Rbar + R2bar;
	Vbar = Vbar - R2bar;
	Rbar = Rbar + (magRbar / s.magR) * s.R;
	// V = q Eb, q = vb * sdt, vb = |v_j|, Eb = pref_j / |pref_j|
	qbar += dot(Vbar, s.Eb);
	const V2 Ebbar = s.q * Vbar;
	const float vbbar = sdt * qbar;
	vjbar = (s.vb > 0.0f) ? (vbbar / s.vb) * vj : V2{0.0f, 0.0f};
	if (s.pref_mag > 0.0f) {
		const float eb = dot(Ebbar, s.Eb);
		prefjbar = V2{(Ebbar.x - eb * s.Eb.x) / s.pref_mag, (Ebbar.y - eb * s.Eb.y) / s.pref_mag};
	} else {
		prefjbar = V2{0.0f, 0.0f};
	}
}

// Forward PowerLaw pair with intermediates (PowerLaw.cpp:28-66).
struct PLPair {
	V2 x, v, comp2, w;
	float a, b, c, sd, tau, E, H, comp1;
};
UD bool pl_forward(const float* prm, const Self& a, const Nb& o, PLPair& s) {
	const float k = prm[1], tau0 = prm[2];
	const float R = o.radius + a.radius;
	s.x = a.pos - o.pos;
	s.v = a.vel - o.vel;
	s.a = dot(s.v, s.v);
	s.b = -dot(s.x, s.v);
	s.c = dot(s.x, s.x) - R * R;
	if (s.c <= 0.0f) return false;
	const float d = s.b * s.b - s.a * s.c;
	if (d <= 0.0f) return false;
	s.sd = sqrtf(d);
	s.tau = (s.b - s.sd) / s.a;
	if (s.tau < 0.001f || s.tau > tau0) return false;
	s.E = expf(-s.tau / tau0);
	s.H = 2.0f / s.tau + 1.0f / tau0;
	s.comp1 = -k * s.E * s.H / (s.a * s.tau * s.tau);
	s.w = s.a * s.x + s.b * s.v;
	s.comp2 = s.v - s.w / s.sd;
	return true;
}
UD void pl_vjp(const float* prm, const PLPair& s, V2 Fbar, V2& xbar, V2& vbar, float& kbar, float& tau0bar) {
	const float k = prm[1], tau0 = prm[2];
	const float c1bar = dot(Fbar, s.comp2);
	const V2 c2bar = s.comp1 * Fbar;
	vbar = c2bar;
	const V2 wbar = (-1.0f / s.sd) * c2bar;
	float sdbar = dot(c2bar, s.w) / (s.sd * s.sd);
	float abar = dot(wbar, s.x);
	xbar = s.a * wbar;
	float bbar = dot(wbar, s.v);
	vbar = vbar + s.b * wbar;
	kbar = c1bar * s.comp1 / k;
	const float dc1_dtau = s.comp1 * (-1.0f / tau0 + (-2.0f / (s.tau * s.tau)) / s.H - 2.0f / s.tau);
	tau0bar = c1bar * s.comp1 * (s.tau / (tau0 * tau0) + (-1.0f / (tau0 * tau0)) / s.H);
	const float taubar = c1bar * dc1_dtau;
	abar += -c1bar * s.comp1 / s.a;
	bbar += taubar / s.a;
	sdbar += -taubar / s.a;
	abar += -taubar * s.tau / s.a;
	const float dbar = sdbar / (2.0f * s.sd);
	bbar += 2.0f * s.b * dbar;
	abar += -s.c * dbar;
	const float cbar = -s.a * dbar;
	xbar = xbar + (2.0f * cbar) * s.x;
	xbar = xbar - bbar * s.v;
	vbar = vbar - bbar * s.x;
	vbar = vbar + (2.0f * abar) * s.v;
}

UD unsigned row_of(const State& S, int j) { return (unsigned)S.meta[j].z; }

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

	const V2 pbar_out = v2(A.adj_pos_out[row]);
	const V2 vbar_out = v2(A.adj_vel_out[row]);
	// p' = p + v' dt
	V2 pbar = pbar_out;
	const V2 vtot = vbar_out + pbar_out * a.dt;   // adjoint of v'
	V2 vbar{0.0f, 0.0f}, prefbar{0.0f, 0.0f};
	// v' = next + C dt / 80
	const V2 nextbar = vtot;
	const V2 Cbar = vtot * (a.dt / 80.0f);

	// contact forces (policy.cpp:107-141)
	if (P.contact_scale > 0.0f) {
		for (int e = 0; e < cnt; ++e) {
			const Nb o = L.get(e);
			const V2 d = a.pos - o.pos;
			const float s = mag(d);
			const float inter = a.radius + o.radius - s;
			if (!(inter > 0.0f) || !(s > 0.0f)) continue;
			const V2 u = d / s;
			const float cu = dot(Cbar, u);
			// J = scale [ (R - s)/s (I - u u^T) - u u^T ], symmetric
			const float f = P.contact_scale;
			const V2 jd = V2{f * (inter / s * (Cbar.x - cu * u.x) - cu * u.x),
			                 f * (inter / s * (Cbar.y - cu * u.y) - cu * u.y)};
			pbar = pbar + jd;
			atomic_add2(&A.adj_pos_in[row_of(S, o.j)], -jd);
		}
	}

	if (!is_zero(a.pref) && P.mode == MODE_GRADIENT) {
		// forward: forces, acceleration, clamps
		V2 F[4];
		V2 acc{0.0f, 0.0f};
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			if (kind == KIND_GOAL_REACHING) F[c] = (a.pref - a.vel) / smax(a.relaxation, a.dt);
			else {
				const float rangeSq = P.range[c] * P.range[c];
				V2 Fs{0.0f, 0.0f};
				for (int e = 0; e < cnt; ++e) {
					const Nb o = L.get(e);
					if (!(o.dsq <= rangeSq)) continue;
					if (kind == KIND_SOCIAL_FORCES) {
						SFPair sp;
						if (sf_forward(P.params[c], a, o, L.pref_of(o.j), P.view_cos, sp)) Fs = Fs + sp.g * sp.R;
					} else if (kind == KIND_POWER_LAW) {
						PLPair pp;
						if (pl_forward(P.params[c], a, o, pp)) Fs = Fs + pp.comp1 * pp.comp2;
					}
				}
				F[c] = Fs;
			}
			acc = acc + (P.coeff[c] / P.params[c][0]) * F[c];
		}
		const V2 accc = clamp_vec(acc, a.max_accel);
		const V2 w = a.vel + accc * a.dt;
		// reverse
		const V2 wbar = clamp_vjp(w, a.max_speed, nextbar);
		vbar = vbar + wbar;
		const V2 accbar = clamp_vjp(acc, a.max_accel, wbar * a.dt);
		float p0bar = 0.0f, p1bar = 0.0f;
		for (int c = 0; c < P.n_cost; ++c) {
			const int kind = P.kind[c];
			const V2 Fbar = (P.coeff[c] / P.params[c][0]) * accbar;
			if (kind == KIND_GOAL_REACHING) {
				const float tr = smax(a.relaxation, a.dt);
				prefbar = prefbar + Fbar / tr;
				vbar = vbar - Fbar / tr;
				continue;
			}
			const float rangeSq = P.range[c] * P.range[c];
			for (int e = 0; e < cnt; ++e) {
				const Nb o = L.get(e);
				if (!(o.dsq <= rangeSq)) continue;
				const unsigned rj = row_of(S, o.j);
				if (kind == KIND_SOCIAL_FORCES) {
					SFPair sp;
					if (!sf_forward(P.params[c], a, o, L.pref_of(o.j), P.view_cos, sp)) continue;
					V2 Rbar, vjbar, prefjbar;
					float V0b, sgb;
					V2 vj = o.vel;
					sf_vjp(P.params[c], sp, Fbar, Rbar, vjbar, prefjbar, vj, V0b, sgb);
					pbar = pbar + Rbar;
					atomic_add2(&A.adj_pos_in[rj], -Rbar);
					atomic_add2(&A.adj_vel_in[rj], vjbar);
					atomic_add2(&A.adj_pref[rj], prefjbar);
					if (A.param_kind == KIND_SOCIAL_FORCES) { p0bar += V0b; p1bar += sgb; }
				} else if (kind == KIND_POWER_LAW) {
					PLPair pp;
					if (!pl_forward(P.params[c], a, o, pp)) continue;
					V2 xbar, vrbar;
					float kb, t0b;
					pl_vjp(P.params[c], pp, Fbar, xbar, vrbar, kb, t0b);
					pbar = pbar + xbar;
					vbar = vbar + vrbar;
					atomic_add2(&A.adj_pos_in[rj], -xbar);
					atomic_add2(&A.adj_vel_in[rj], -vrbar);
					if (A.param_kind == KIND_POWER_LAW) { p0bar += kb; p1bar += t0b; }
				}
			}
		}
		float* pc = A.adj_params + (size_t)stat_copy() * A.param_stride;
		if (p0bar != 0.0f) atomicAdd(&pc[2 * pi + 0], p0bar);
		if (p1bar != 0.0f) atomicAdd(&pc[2 * pi + 1], p1bar);
	}
	atomic_add2(&A.adj_pos_in[row], pbar);
	atomic_add2(&A.adj_vel_in[row], vbar);
	atomic_add2(&A.adj_pref[row], prefbar);
}

// pref = ps (g - p) / |g - p| unless the goal is reached: adjoint to the position.
__global__ __launch_bounds__(256) void k_pref_adjoint(State in, int n, const float2* adj_pref, float2* adj_pos) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const unsigned row = (unsigned)in.meta[k].z;
	const V2 pb = v2(adj_pref[row]);
	const V2 d = v2(in.goal[k]) - v2(in.pos[k]);
	const float4 prm = in.prm[k];
	if (sqr_mag(d) <= prm.x * prm.x) return;
	const float m = mag(d);
	const V2 u = d / m;
	const float pu = dot(pb, u);
	const float s = prm.y / m;
	// d pref / d p = -(ps / m) (I - u u^T)
	const V2 g{-s * (pb.x - pu * u.x), -s * (pb.y - pu * u.y)};
	atomic_add2(&adj_pos[row], g);
}

// Loss of one step's output: sum over agents of |p - target|^2 with the minimum-image
// difference in toric worlds; optionally the gradient scale * 2 (p - target) (row space).
__global__ __launch_bounds__(256) void k_loss(State out, int n, const WorldInfo* worlds, const float2* target,
                                             double* loss, float2* grad, float scale) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	double l = 0.0;
	if (k < n) {
		const int4 meta = out.meta[k];
		const WorldInfo wi = worlds[meta.y >> kPolicyBits];
		const unsigned row = (unsigned)meta.z;
		const float2 p = out.pos[k], t = target[row];
		float dx = p.x - t.x, dy = p.y - t.y;
		if (wi.world_type == 1) {
			if (dx > 0.5f * wi.W) dx -= wi.W; else if (dx < -0.5f * wi.W) dx += wi.W;
			if (dy > 0.5f * wi.H) dy -= wi.H; else if (dy < -0.5f * wi.H) dy += wi.H;
		}
		l = (double)dx * dx + (double)dy * dy;
		if (grad) {
			float2 g = grad[row];
			g.x += scale * 2.0f * dx;
			g.y += scale * 2.0f * dy;
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

} // namespace umans
