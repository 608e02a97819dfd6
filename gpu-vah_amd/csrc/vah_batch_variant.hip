// vah_batch_variant.hip -- instantiates the persistent batch stage kernels for
// one tile shape and occupancy target.  Built once per variant with
//   -DVAH_BV_TX=<tile x> -DVAH_BV_TY=<tile y> -DVAH_BV_B=<workgroups per CU>
//   -DVAH_BV_FN=<name of the launch function>
// (the variant table in vah_batch.hip lists the builds).
#include "vah_batch_kernels.h"

#if !defined(VAH_BV_TX) || !defined(VAH_BV_TY) || !defined(VAH_BV_B) || !defined(VAH_BV_FN)
#error "vah_batch_variant.hip needs VAH_BV_TX, VAH_BV_TY, VAH_BV_B and VAH_BV_FN"
#endif

namespace vah {

namespace {

// a 3+1D workgroup carries the eta faces in LDS as well (2 x 15 x 256
// values): at most two fit a CU in double precision
constexpr int kBlocks3D = VAH_BV_B > 2 ? 2 : VAH_BV_B;

template <bool D3, bool S2, typename R>
void launch_one(const BatchLaunch &L) {
    constexpr int B = D3 ? kBlocks3D : VAH_BV_B;
    hipLaunchKernelGGL((batch_stage_kernel<D3, S2, R, VAH_BV_TX, VAH_BV_TY, B>), dim3(D3 ? L.grid3 : L.grid), dim3(NTHREADS),
                       0, L.stream, L.rec, D3 ? L.tiles3 : L.tiles2, L.counts);
}

template <typename R>
void launch_r(const BatchLaunch &L, bool d3, bool stage2) {
    if (d3) {
        if (stage2) launch_one<true, true, R>(L);
        else launch_one<true, false, R>(L);
    } else {
        if (stage2) launch_one<false, true, R>(L);
        else launch_one<false, false, R>(L);
    }
}

}  // namespace

void VAH_BV_FN(const BatchLaunch &L, bool d3, bool stage2, bool fp32) {
    if (fp32)
        launch_r<float>(L, d3, stage2);
    else
        launch_r<double>(L, d3, stage2);
}

}  // namespace vah
