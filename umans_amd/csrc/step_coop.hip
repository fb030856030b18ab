// The cooperative-sampling kernel's instances (k_step_coop) for every list capacity, both
// agents-per-wave counts and the single-cost specialisations.
#define UMANS_STEP_KERNELS_ONLY
#include "kernels.h"
#include "step_launch.h"

namespace umans {
namespace {

// Agents per wave of the cooperative kernel: 64 while that makes at least ~4 waves per
// SIMD (1024 SIMDs), else 32 (twice the waves, half the candidates per lane).
int coop_agents_per_wave(int n) { return n >= 262144 ? 64 : (n >= 131072 ? 32 : 16); }

template <int KMAX, int KIND>
void launch_coop_k(const StepArgs& a, hipStream_t st) {
	const int apw = coop_agents_per_wave(a.n);
	const int blocks = (a.n + apw - 1) / apw;
	if (blocks <= 0) return;
	if (apw == 64) hipLaunchKernelGGL((k_step_coop<KMAX, 64, KIND>), dim3(blocks), dim3(64), 0, st, a);
	else if (apw == 32) hipLaunchKernelGGL((k_step_coop<KMAX, 32, KIND>), dim3(blocks), dim3(64), 0, st, a);
	else hipLaunchKernelGGL((k_step_coop<KMAX, 16, KIND>), dim3(blocks), dim3(64), 0, st, a);
}

// kind: the one cost of every cooperative policy, or -1; instances of single kinds for
// lists of up to 16 entries only
template <int KMAX>
void launch_coop_t(const StepArgs& a, hipStream_t st, int kind) {
	if constexpr (KMAX <= 16) {
		switch (kind) {
		case KIND_KARAMOUZAS: launch_coop_k<KMAX, KIND_KARAMOUZAS>(a, st); return;
		case KIND_RVO: launch_coop_k<KMAX, KIND_RVO>(a, st); return;
		case KIND_MOUSSAID: launch_coop_k<KMAX, KIND_MOUSSAID>(a, st); return;
		case KIND_TTCA_DCA: launch_coop_k<KMAX, KIND_TTCA_DCA>(a, st); return;
		case KIND_PARIS: launch_coop_k<KMAX, KIND_PARIS>(a, st); return;
		case KIND_VAN_TOLL: launch_coop_k<KMAX, KIND_VAN_TOLL>(a, st); return;
		default: break;
		}
	}
	launch_coop_k<KMAX, -1>(a, st);
}

} // namespace

void launch_step_coop(int kmax, const StepArgs& a, hipStream_t st, int kind) {
	switch (kmax) {
	case 12: launch_coop_t<12>(a, st, kind); break;
	case 16: launch_coop_t<16>(a, st, kind); break;
	case 32: launch_coop_t<32>(a, st, kind); break;
	case 64: launch_coop_t<64>(a, st, kind); break;
	default: launch_coop_t<128>(a, st, kind); break;
	}
}

} // namespace umans
