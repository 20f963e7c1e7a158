// lbm_q27_16_ibb.hip -- the D3Q27 step with 16-bit population storage and the
// interpolated bounce-back on the links.
#include "lbm_q27_kernel.h"

namespace lbm {

int q27_launch_step16_ibb(const q27::StepArgs& a) {
    if (a.storage == LBM_FP16S) q27::launch_storage<q27::StoreF16S27, true>(a);
    else q27::launch_storage<q27::StoreBF16S27, true>(a);
    return q27::last_error();
}

}  // namespace lbm
