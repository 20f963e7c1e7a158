// lbm_store_f16s.hip -- kernels for StoreF16S populations.
#include "lbm_launch.h"

namespace lbm {
template int launch<StoreF16S>(const LaunchReq& r);
}  // namespace lbm
