// kernels.h -- device-side building blocks of the SCFT propagator path and
// the host launchers that the C-ABI layer (scft_lib.hip) calls.
//
// Data layout in HBM (one context, `batch` instances of the same grid):
//   point index   r = (k*ny + j)*nx + i          (x fastest, as init_cuda.cu:425)
//   fields        wA, wB, phiA, phiB : double [batch][N]
//   wdz           T [4][batch][N]       exp(-w*ds/2) of block A, B; the same / N
//   slice store   [Ns+1] complex fields  z(t) = q(t) + i*qdag(Ns - t)
//   work          one complex field      the field between the three axis passes
// A complex field is planar: T [2][batch][N], the real parts of all points,
// then their imaginary parts (host code passes it as a C pointer).  The split
// xy-plane kernel, one workgroup per part, then reads and writes whole cache
// lines; the line kernels load a point's two parts with two coalesced loads.
// where T is the propagator precision (double or float) and C its complex type.
#pragma once

#include <hip/hip_runtime.h>

#include "dft.h"

namespace scft {

enum Axis { AX_X = 0, AX_Y = 1, AX_Z = 2 };

// Entry type of a k-factor table (the propagator precision).
template <typename T> struct kfac { using type = T; };
template <typename C, typename T> __device__ __forceinline__ C kf_apply(C v, T k) { return C{v.x * k, v.y * k}; }

// One batched 1-D convolution pass along an axis:
//   line <- IDFT[ kf(k) * DFT[ line ] ]      (unnormalised DFTs, kf(0) = 1)
// pre:  the loaded value is multiplied part-wise by the pre W factors first,
// post: the result is multiplied part-wise by the post W factors before the
//       store (these carry the 1/N of the inverse transforms).
// W = [pre A, pre B, post A, post B] ([4][BN]); step == 0 takes the four
// factors at the element offsets wo[] (pre real, pre imag, post real, post
// imag) for every line.  Per-instance mode (step > 0): each line's blocks are
// chosen in the kernel from nsab[instance] = (NsA, NsB) at contour step
// `step` (real part: block A while step <= NsA; imaginary part: block B while
// step <= NsB), and lines of instances whose chain is shorter than `step` are
// left untouched.  kf of instance b is kf + b * kf_stride.
// Every element offset is 32-bit: the host keeps 4*BN*sizeof(T) < 2^32.
template <typename T>
struct PassArgs {
    using C = typename cplx<T>::type;
    const C *src;
    C *dst;
    const T *W;          // the four W arrays (FP32X: the two pre arrays hold W - 1)
    unsigned wo[4];      // step == 0: element offsets of the four factors in W
    C *slot;             // step-boundary kernels: z(step) goes here, dst gets step + 1 opened
    unsigned wn[2];      // step-boundary kernels, step == 0: pre factors (real, imag) of step + 1
    T post_scale;        // step-boundary kernels with post_from_pre: post factor = pre factor * post_scale
    int post_from_pre;   //   (exact: 1/N a power of two and no folded length-1 axis)
    const int2 *nsab;    // per-instance mode: (NsA, NsB) per instance
    const C *tw;         // n entries, (cos 2 pi k/n, -sin 2 pi k/n)
    const typename kfac<T>::type *kf;  // n entries per instance (kf_stride apart, 0 = shared)
    unsigned kf_stride;
    unsigned nlines;     // lines in this launch (batch * N / n)
    unsigned N;          // points per instance
    unsigned BN;         // batch * N
    int nx, ny, nz;
    int pre, post;
    int step;            // contour step (per-instance mode) or 0
};

// Radix plan of the LDS Stockham line kernel that serves even 2-3-5-smooth
// lengths without a register tiling: n = r[0] * ... * r[nr-1], radices 4
// first, then 2, 3, 5.  GEN_MAX bounds n (two FP64 complex lines in 64 KB of
// LDS).
constexpr int GEN_MAX = 2048;
struct GenPlan {
    int n, nr;
    int r[16];
};
__host__ __device__ inline bool gen_plan(int n, GenPlan *g)
{
    if (n < 2 || n > GEN_MAX || (n & 1)) return false;
    g->n = n;
    g->nr = 0;
    int m = n;
    while (m % 4 == 0 && g->nr < 16) {
        g->r[g->nr++] = 4;
        m /= 4;
    }
    const int rs[3] = {2, 3, 5};
    for (int k = 0; k < 3; ++k)
        while (m % rs[k] == 0 && g->nr < 16) {
            g->r[g->nr++] = rs[k];
            m /= rs[k];
        }
    return m == 1;
}

// true if the pass kernels have an instantiation for line length n (the
// register tilings, else the LDS Stockham kernel; and the three-level
// tiling for it)
bool pass_supported(int n);
bool pass3_supported(int n);
// launch one pass along `axis` (n = length of that axis) in tiling variant
// `variant`: 0 one line per thread group, 1 two lines per group, 2 the
// three-level tiling (twice the threads per line; lengths 32, 64, 128, 256)
constexpr int PASS_VARIANTS = 3;
template <typename T>
hipError_t launch_pass(int axis, int n, int variant, const PassArgs<T> &a, hipStream_t s);
// xy-plane kernels for nx = ny = n.  form 1: one workgroup per plane; form 2:
// one workgroup per plane and real part; form 3: form 2 with the three-level
// tiling; forms 4 and 5: forms 2 and 3 with each y line on consecutive lanes
// of one wave.  nplanes = batch * nz.
template <typename T>
bool plane_form_supported(int form, int n);
template <typename T>
hipError_t launch_plane_form(int form, int n, const PassArgs<T> &a, const typename cplx<T>::type *twy,
                             const typename kfac<T>::type *kfy, long nplanes, hipStream_t s);
// Step-boundary pass (the march alternates its axis order, so the last axis
// of step t is the first of step t+1): convolve along `axis` (step t), store
// z(t) = result * post W(t) to a.slot, multiply by pre W(t+1), convolve again
// (step t+1) and store to a.dst.  a.pre multiplies the load by pre W(t) (the
// first step).  Per-instance mode: lines of instances whose chain ends at t
// store only z(t); lines of finished instances are left untouched.  variant 2
// selects the three-level tiling (others the one-line two-level one).
template <typename T>
hipError_t launch_fused(int axis, int n, int variant, const PassArgs<T> &a, hipStream_t s);
// the step-boundary forms of the xy-plane kernels
template <typename T>
hipError_t launch_plane_fused(int form, int n, const PassArgs<T> &a, const typename cplx<T>::type *twy,
                              const typename kfac<T>::type *kfy, long nplanes, hipStream_t s);
// the (A, B, lines-per-workgroup, threads) tiling used for (axis, n)
void pass_tiling(int axis, int n, int is_fp32, int *A, int *B, int *lines, int *threads);

}  // namespace scft

// The same launchers over FP32 storage with FP64 arithmetic (pass_f32x.hip):
// PassArgs<float> pointers for the field, W and slices; tw and kf point to
// FP64 tables (double2 / double entries) despite their declared type.
namespace scft_f32x {
template <typename T>
hipError_t launch_pass(int axis, int n, int variant, const scft::PassArgs<T> &a, hipStream_t s);
template <typename T>
bool plane_form_supported(int form, int n);
template <typename T>
hipError_t launch_plane_form(int form, int n, const scft::PassArgs<T> &a, const typename scft::cplx<T>::type *twy,
                             const typename scft::kfac<T>::type *kfy, long nplanes, hipStream_t s);
template <typename T>
hipError_t launch_fused(int axis, int n, int variant, const scft::PassArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_plane_fused(int form, int n, const scft::PassArgs<T> &a, const typename scft::cplx<T>::type *twy,
                              const typename scft::kfac<T>::type *kfy, long nplanes, hipStream_t s);
}  // namespace scft_f32x
