// lbm_store_f16s_ibb.hip -- interpolated bounce-back kernels for StoreF16S populations.
#include "lbm_launch_ibb.h"

namespace lbm {
template void launch_step_ibb<StoreF16S>(const LaunchReq& r, const Phys<StoreF16S::C>& p);
}  // namespace lbm
