// Launchers of the step kernels' instances. Each policy class's instances live in two
// translation units of their own (step_class.hip, built once per class and half of the
// list capacities; step_coop.hip), so the library's kernels compile in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include "step_types.h"

namespace umans {

// k_step<kmax, bs, cls> over a.n agents, one lane each: launch_step_c<cls>_p<part>, part 0
// for kmax 12 and 16 (sorted lists in LDS), part 1 for kmax 32, 64, 128 and 0 (long
// lists ranked per wave); bs 64, 128 or 256 as the capacity allows.
#define UMANS_STEP_LAUNCHER_NAME(cls, part) launch_step_c##cls##_p##part
#define UMANS_STEP_LAUNCHER_(cls, part) \
	void UMANS_STEP_LAUNCHER_NAME(cls, part)(int kmax, int bs, const StepArgs& a, hipStream_t st)
#define UMANS_STEP_LAUNCHER(cls, part) UMANS_STEP_LAUNCHER_(cls, part)
UMANS_STEP_LAUNCHER(0, 0);
UMANS_STEP_LAUNCHER(0, 1);
UMANS_STEP_LAUNCHER(1, 0);
UMANS_STEP_LAUNCHER(1, 1);
UMANS_STEP_LAUNCHER(2, 0);
UMANS_STEP_LAUNCHER(2, 1);
UMANS_STEP_LAUNCHER(3, 0);
UMANS_STEP_LAUNCHER(3, 1);
// k_step_coop<kmax, agents per wave, kind> over a.n agents (kind: the one cost of every
// cooperative policy, or -1)
void launch_step_coop(int kmax, const StepArgs& a, hipStream_t st, int kind);

} // namespace umans
