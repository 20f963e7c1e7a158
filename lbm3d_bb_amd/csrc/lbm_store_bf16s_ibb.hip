// lbm_store_bf16s_ibb.hip -- interpolated bounce-back kernels for StoreBF16S populations.
#include "lbm_launch_ibb.h"

namespace lbm {
template void launch_step_ibb<StoreBF16S>(const LaunchReq& r, const Phys<StoreBF16S::C>& p);
}  // namespace lbm
