// vah_stage3d.hip -- launches of the 3+1D single-lattice stage kernels.  They
// sit in a translation unit of their own because the machine scheduler with
// the AMDGPU register-pressure trackers (-amdgpu-use-amdgpu-trackers) speeds
// them up (201x201x101 FP64 step 2.61 -> 2.56 ms) but slows the 2+1D kernels
// (81x81 step 0.0756 -> 0.0794 ms); the build gives this unit that setting only.
#include <hip/hip_runtime.h>

#include "../../include/vah.h"
#include "vah_stage.h"

namespace vah {

namespace {

template <bool S2, bool SH, typename R>
void launch3(const StageLaunch &L, const StageParams &P) {
    hipLaunchKernelGGL((stage_kernel<true, S2, SH, R>), dim3(L.grid), dim3(NTHREADS), 0, L.stream, P);
}

template <bool S2>
void variant3(const StageLaunch &L, const StageParams &P) {
    if (L.precision == VAH_FP32_SPLIT)
        launch3<S2, true, float>(L, P);
    else if (L.flux_mode == VAH_FLUX_SHARED)
        launch3<S2, true, double>(L, P);
    else
        launch3<S2, false, double>(L, P);
}

}  // namespace

void launch_stage_3d(const StageLaunch &L, int stage, const StageParams &P) {
    if (stage == 1)
        variant3<false>(L, P);
    else
        variant3<true>(L, P);
}

}  // namespace vah
