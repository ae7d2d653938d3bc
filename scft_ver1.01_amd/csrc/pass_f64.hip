// FP64 instantiation of the axis-pass kernels, plus the size table queries.
#include "pass_kernels.inc"

namespace scft {

template hipError_t launch_pass<double>(int, int, int, const PassArgs<double> &, hipStream_t);

bool pass_supported(int n)
{
    switch (n) {
#define SCFT_CASE(N_, A_, B_) \
    case N_: return true;
        SCFT_PASS_SIZES(SCFT_CASE)
#undef SCFT_CASE
    default: {
        GenPlan g;
        return gen_plan(n, &g);
    }
    }
}

bool pass3_supported(int n)
{
    switch (n) {
#define SCFT_CASE(N_, A_, B_) \
    case N_: return true;
        SCFT_PASS3_SIZES(SCFT_CASE)
#undef SCFT_CASE
    default: return false;
    }
}

void pass_tiling(int axis, int n, int is_fp32, int *A, int *B, int *lines, int *threads)
{
    (void)is_fp32;
    *A = *B = *lines = *threads = 0;
    switch (n) {
#define SCFT_CASE(N_, A_, B_)                                                        \
    case N_:                                                                         \
        *A = A_;                                                                     \
        *B = B_;                                                                     \
        *lines = (axis == AX_X) ? Lines<B_, AX_X>::value : Lines<B_, AX_Y>::value;   \
        *threads = B_ * *lines;                                                      \
        break;
        SCFT_PASS_SIZES(SCFT_CASE)
#undef SCFT_CASE
    default: break;
    }
}

template bool plane_form_supported<double>(int, int);
template hipError_t launch_plane_form<double>(int, int, const PassArgs<double> &, const cplx<double>::type *, const double *,
                                             long, hipStream_t);
template hipError_t launch_fused<double>(int, int, int, const PassArgs<double> &, hipStream_t);
template hipError_t launch_plane_fused<double>(int, int, const PassArgs<double> &, const cplx<double>::type *, const double *,
                                              long, hipStream_t);
}  // namespace scft
