// lbm_store_bf16s.hip -- kernels for StoreBF16S populations.
#include "lbm_launch.h"

namespace lbm {
template int launch<StoreBF16S>(const LaunchReq& r);
}  // namespace lbm
