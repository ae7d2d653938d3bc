// FP32 instantiation of the axis-pass kernels.
#include "pass_kernels.inc"

namespace scft {
template hipError_t launch_pass<float>(int, int, const PassArgs<float> &, hipStream_t);
}  // namespace scft
