// FP32 instantiation of the axis-pass kernels.
#include "pass_kernels.inc"

namespace scft {
template hipError_t launch_pass<float>(int, int, int, const PassArgs<float> &, hipStream_t);
template bool plane_form_supported<float>(int, int);
template hipError_t launch_plane_form<float>(int, int, const PassArgs<float> &, const cplx<float>::type *, const float *,
                                             long, hipStream_t);
template hipError_t launch_fused<float>(int, int, int, const PassArgs<float> &, hipStream_t);
template hipError_t launch_plane_fused<float>(int, int, const PassArgs<float> &, const cplx<float>::type *, const float *,
                                              long, hipStream_t);
}  // namespace scft
