// lbm_q27_16.hip -- the D3Q27 step, start state and population import with
// 16-bit population storage (shifted FP16, shifted BF16), simple bounce-back.
#include "lbm_q27_kernel.h"

namespace lbm {

int q27_launch_step16(const q27::StepArgs& a) {
    if (a.storage == LBM_FP16S) q27::launch_storage<q27::StoreF16S27, false>(a);
    else q27::launch_storage<q27::StoreBF16S27, false>(a);
    return q27::last_error();
}

int q27_launch_equilibrium16(int storage, const Grid& g, hipStream_t st, void* f, const double* rho, const double* u,
                             double rho0, const double u0[3]) {
    const dim3 grid((unsigned)nblocks_of(g.N)), block(kBlock);
    if (storage == LBM_FP16S)
        hipLaunchKernelGGL(q27::k27_equilibrium<q27::StoreF16S27>, grid, block, 0, st, (__half*)f, rho, u, rho0, u0[0],
                           u0[1], u0[2], g.N, g.P, g.sh);
    else
        hipLaunchKernelGGL(q27::k27_equilibrium<q27::StoreBF16S27>, grid, block, 0, st, (uint16_t*)f, rho, u, rho0,
                           u0[0], u0[1], u0[2], g.N, g.P, g.sh);
    return q27::last_error();
}

int q27_launch_from_double16(int storage, const Grid& g, hipStream_t st, void* f, const double* src) {
    const dim3 grid((unsigned)nblocks_of(g.N)), block(kBlock);
    if (storage == LBM_FP16S)
        hipLaunchKernelGGL(q27::k27_from_double<q27::StoreF16S27>, grid, block, 0, st, (__half*)f, src, g.N, g.P, g.sh);
    else
        hipLaunchKernelGGL(q27::k27_from_double<q27::StoreBF16S27>, grid, block, 0, st, (uint16_t*)f, src, g.N, g.P,
                           g.sh);
    return q27::last_error();
}

}  // namespace lbm
