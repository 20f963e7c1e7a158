// lbm_store_f64.hip -- kernels for StoreF64 populations.
#include "lbm_launch.h"

namespace lbm {
template int launch<StoreF64>(const LaunchReq& r);
}  // namespace lbm
