// The step kernel's instances of one policy class (UMANS_CLS, set by the build) for half
// of the list capacities (UMANS_PART) and every block size the launch tuning offers.
#define UMANS_STEP_KERNELS_ONLY
#include "kernels.h"
#include "step_launch.h"

#if !defined(UMANS_CLS) || !defined(UMANS_PART)
#error "UMANS_CLS (the policy class's number) and UMANS_PART (list capacities) select the instances"
#endif

namespace umans {
namespace {

template <int KMAX, int BS>
void launch_t(const StepArgs& a, hipStream_t st) {
	const int apw = KMAX == 0 && a.long_apw > 0 ? a.long_apw : BS;   // agents per block
	const int blocks = (a.n + apw - 1) / apw;
	if (blocks > 0) hipLaunchKernelGGL((k_step<KMAX, BS, UMANS_CLS>), dim3(blocks), dim3(BS), 0, st, a);
}

// UMANS_PART 0: the short lists (caps 12 and 16); 1: the rest
void launch_class(int kmax, int bs, const StepArgs& a, hipStream_t st) {
#if UMANS_PART == 0
	switch (kmax) {
	case 12:
		if (bs == 64) launch_t<12, 64>(a, st);
		else if (bs == 256) launch_t<12, 256>(a, st);
		else launch_t<12, 128>(a, st);
		break;
	default:
		if (bs == 64) launch_t<16, 64>(a, st);
		else if (bs == 256) launch_t<16, 256>(a, st);
		else launch_t<16, 128>(a, st);
		break;
	}
#else
	switch (kmax) {
	case 32:
		if (bs == 64) launch_t<32, 64>(a, st);
		else launch_t<32, 128>(a, st);
		break;
	case 64: launch_t<64, 64>(a, st); break;
	case 0: launch_t<0, 64>(a, st); break;     // long lists, ranked per wave
	default: launch_t<128, 64>(a, st); break;
	}
#endif
}

} // namespace

static_assert(UMANS_CLS >= 0 && UMANS_CLS < CLS_COOP && (UMANS_PART == 0 || UMANS_PART == 1), "unit");
UMANS_STEP_LAUNCHER(UMANS_CLS, UMANS_PART) { launch_class(kmax, bs, a, st); }

} // namespace umans
