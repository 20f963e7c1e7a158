// lbm_q27_ibb.hip -- the D3Q27 step with the interpolated bounce-back on the
// links (its own translation unit: the instances build in parallel).
#include "lbm_q27_kernel.h"

namespace lbm {

int q27_launch_step_ibb(const q27::StepArgs& a) {
    if (a.storage == LBM_FP64) q27::launch_storage<StoreF64, true>(a);
    else q27::launch_storage<StoreF32, true>(a);
    return q27::last_error();
}

}  // namespace lbm
