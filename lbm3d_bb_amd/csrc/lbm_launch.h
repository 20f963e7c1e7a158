// lbm_launch.h -- launch<S>(): the kernels of one storage format.  Included
// once per storage by lbm_store_*.hip, which instantiates it.
#pragma once

#include "lbm_device.h"

namespace lbm {

template <class S, int MODE, int COLL, bool FORCED>
static void launch_bulk_t(const LaunchReq& r, const Phys<typename S::C>& p) {
    using T = typename S::T;
    const BulkShape bs = bulk_shape<S>(r.g);
    hipLaunchKernelGGL((k_bulk<S, COLL, FORCED, MODE>), dim3(bs.gx, bs.gy, r.g.nz), dim3(bs.bx, bs.by), 0, r.stream,
                       (const T*)r.fin, (T*)r.fout, r.gout, r.rho_out, r.u_out, r.g, p, (const StepGeom*)r.rec,
                       (const uint8_t*)r.flags);
}

template <class S, int MODE, int COLL, bool FORCED, bool IBB = false>
static void launch_step_t(const LaunchReq& r, const Phys<typename S::C>& p) {
    using T = typename S::T;
    const StepGeom* sg = (const StepGeom*)r.rec;
    const DevBody* bn = (const DevBody*)(r.rec + sizeof(StepGeom));
    const DevBody* bo = bn + r.nb;
    const BulkShape bs = bulk_shape<S>(r.g);
    const int64_t nbulk = (int64_t)bs.gx * bs.gy * r.g.nz;
    hipLaunchKernelGGL((k_step<S, COLL, FORCED, MODE, IBB>), dim3((unsigned)(r.grid_blocks + nbulk)), dim3(kBlock),
                       (unsigned)step_lds_bytes(r.nb, IBB), r.stream, (const T*)r.fin, (T*)r.fout, r.gout, r.rho_out, r.u_out, r.g, p, sg, bn, bo, r.flags,
                       r.list, r.list_count, r.partials, r.write_flags, bs, r.grid_blocks);
}

// Step and getter kernels with the interpolated bounce-back (separate
// instantiations, so the simple bounce-back kernels keep their register
// budget): defined in lbm_launch_ibb.h, instantiated by lbm_store_*_ibb.hip.
template <class S>
void launch_step_ibb(const LaunchReq& r, const Phys<typename S::C>& p);

template <class S, int MODE>
static void launch_mode(const LaunchReq& r, const Phys<typename S::C>& p) {
    const bool trt = r.collision == LBM_TRT;
    if (r.kind == K_BULK) {
        if (trt) r.forced ? launch_bulk_t<S, MODE, LBM_TRT, true>(r, p) : launch_bulk_t<S, MODE, LBM_TRT, false>(r, p);
        else r.forced ? launch_bulk_t<S, MODE, LBM_BGK, true>(r, p) : launch_bulk_t<S, MODE, LBM_BGK, false>(r, p);
    } else if ((MODE == MODE_STEP || MODE == MODE_GET) && r.solid_bc == LBM_SOLID_IBB) {
        launch_step_ibb<S>(r, p);
    } else if (MODE == MODE_STEP || MODE == MODE_GET) {
        if (trt) r.forced ? launch_step_t<S, MODE, LBM_TRT, true>(r, p) : launch_step_t<S, MODE, LBM_TRT, false>(r, p);
        else r.forced ? launch_step_t<S, MODE, LBM_BGK, true>(r, p) : launch_step_t<S, MODE, LBM_BGK, false>(r, p);
    }
}

template <class S>
int launch(const LaunchReq& r) {
    using C = typename S::C;
    using T = typename S::T;
    Phys<C> p;
    p.wp = (C)r.p.wp;
    p.wm = (C)r.p.wm;
    for (int k = 0; k < 3; ++k) p.g[k] = (C)r.p.g[k];
    const dim3 grid((unsigned)nblocks_of(r.g.N)), block(kBlock);
    switch (r.kind) {
        case K_BULK:
        case K_STEP:
            switch (r.mode) {
                case MODE_STEP: launch_mode<S, MODE_STEP>(r, p); break;
                case MODE_GET: launch_mode<S, MODE_GET>(r, p); break;
                case MODE_COLLIDE_ONLY: launch_mode<S, MODE_COLLIDE_ONLY>(r, p); break;
                case MODE_GET_RAW: launch_mode<S, MODE_GET_RAW>(r, p); break;
                default: return (int)hipErrorInvalidValue;
            }
            break;
        case K_EQUILIBRIUM:
            hipLaunchKernelGGL(k_equilibrium<S>, grid, block, 0, r.stream, (T*)r.fout, r.src_rho, r.src_u, r.rho0,
                               r.u0[0], r.u0[1], r.u0[2], r.g.N, r.g.P, r.g.sh);
            break;
        case K_FROM_DOUBLE:
            hipLaunchKernelGGL(k_from_double<S>, grid, block, 0, r.stream, (T*)r.fout, r.src_f, r.g.N, r.g.P, r.g.sh);
            break;
        default: return (int)hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace lbm
