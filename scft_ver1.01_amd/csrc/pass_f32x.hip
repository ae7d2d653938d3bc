// FP32-storage, FP64-arithmetic instantiation of the axis-pass kernels
// (namespace scft_f32x): the field, the W factors and the slices move through
// HBM as FP32 -- the bytes of the FP32 form -- while every transform runs in
// FP64 registers and LDS, with FP64 twiddle and k-factor tables, and each
// value is rounded to FP32 once, where it is stored.
#define SCFT_PASS_NS scft_f32x
#define SCFT_PASS_MIXED 1
#include "pass_kernels.inc"

namespace scft_f32x {
template hipError_t launch_pass<float>(int, int, int, const scft::PassArgs<float> &, hipStream_t);
template bool plane_form_supported<float>(int, int);
template hipError_t launch_plane_form<float>(int, int, const scft::PassArgs<float> &, const float2 *, const float *,
                                             long, hipStream_t);
template hipError_t launch_fused<float>(int, int, int, const scft::PassArgs<float> &, hipStream_t);
template hipError_t launch_plane_fused<float>(int, int, const scft::PassArgs<float> &, const float2 *, const float *,
                                              long, hipStream_t);
}  // namespace scft_f32x
