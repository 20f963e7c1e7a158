// lbm_store_f64_ibb.hip -- interpolated bounce-back kernels for StoreF64 populations.
#include "lbm_launch_ibb.h"

namespace lbm {
template void launch_step_ibb<StoreF64>(const LaunchReq& r, const Phys<StoreF64::C>& p);
}  // namespace lbm
