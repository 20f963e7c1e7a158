// lbm_launch_ibb.h -- launch_step_ibb<S>(): the step and getter kernels of one
// storage format with the interpolated bounce-back on the body links.
// Included once per storage by lbm_store_*_ibb.hip, which instantiates it.
#pragma once

#include "lbm_launch.h"

namespace lbm {

template <class S, int MODE>
static void launch_ibb_mode(const LaunchReq& r, const Phys<typename S::C>& p) {
    if (r.collision == LBM_TRT)
        r.forced ? launch_step_t<S, MODE, LBM_TRT, true, true>(r, p) : launch_step_t<S, MODE, LBM_TRT, false, true>(r, p);
    else
        r.forced ? launch_step_t<S, MODE, LBM_BGK, true, true>(r, p) : launch_step_t<S, MODE, LBM_BGK, false, true>(r, p);
}

template <class S>
void launch_step_ibb(const LaunchReq& r, const Phys<typename S::C>& p) {
    if (r.mode == MODE_STEP) launch_ibb_mode<S, MODE_STEP>(r, p);
    else if (r.mode == MODE_GET) launch_ibb_mode<S, MODE_GET>(r, p);
}

}  // namespace lbm
