// vah_batch_variant.hip -- instantiates the persistent batch stage kernels for
// one tile shape and occupancy target.  Built twice per variant with
//   -DVAH_BV_TX=<tile x> -DVAH_BV_TY=<tile y> -DVAH_BV_B=<workgroups per CU>
//   -DVAH_BV_FN=<name of the launch function> -DVAH_BV_FN3=<its 3+1D part>
// and -DVAH_BV_PART=2 (the launch function and the 2+1D kernels) or
// -DVAH_BV_PART=3 (the 3+1D kernels), so that the two dimensionalities can be
// compiled with different scheduler settings (the variant table in
// vah_batch.hip lists the builds).
#include "vah_batch_kernels.h"

#if !defined(VAH_BV_TX) || !defined(VAH_BV_TY) || !defined(VAH_BV_B) || !defined(VAH_BV_FN) || \
    !defined(VAH_BV_FN3) || !defined(VAH_BV_PART)
#error "vah_batch_variant.hip needs VAH_BV_TX, VAH_BV_TY, VAH_BV_B, VAH_BV_FN, VAH_BV_FN3 and VAH_BV_PART"
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

template <bool D3, typename R>
void launch_r(const BatchLaunch &L, bool stage2) {
    if (stage2)
        launch_one<D3, true, R>(L);
    else
        launch_one<D3, false, R>(L);
}

}  // namespace

#if VAH_BV_PART == 3
void VAH_BV_FN3(const BatchLaunch &L, bool stage2, bool fp32) {
    if (fp32)
        launch_r<true, float>(L, stage2);
    else
        launch_r<true, double>(L, stage2);
}
#else
void VAH_BV_FN3(const BatchLaunch &L, bool stage2, bool fp32);

void VAH_BV_FN(const BatchLaunch &L, bool d3, bool stage2, bool fp32) {
    if (d3)
        VAH_BV_FN3(L, stage2, fp32);
    else if (fp32)
        launch_r<false, float>(L, stage2);
    else
        launch_r<false, double>(L, stage2);
}
#endif

}  // namespace vah
