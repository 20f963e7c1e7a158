// lbm_store_f32_ibb.hip -- interpolated bounce-back kernels for StoreF32 populations.
#include "lbm_launch_ibb.h"

namespace lbm {
template void launch_step_ibb<StoreF32>(const LaunchReq& r, const Phys<StoreF32::C>& p);
}  // namespace lbm
