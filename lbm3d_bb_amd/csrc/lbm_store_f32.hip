// lbm_store_f32.hip -- kernels for StoreF32 populations.
#include "lbm_launch.h"

namespace lbm {
template int launch<StoreF32>(const LaunchReq& r);
}  // namespace lbm
