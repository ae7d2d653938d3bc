// kernels.h -- device-side building blocks of the SCFT propagator path and
// the host launchers that the C-ABI layer (scft_lib.hip) calls.
//
// Data layout in HBM (one context, `batch` instances of the same grid):
//   point index   r = (k*ny + j)*nx + i          (x fastest, as init_cuda.cu:425)
//   fields        wA, wB, phiA, phiB : double [batch][N]
//   wdz           T [2][batch][N]       exp(-w*ds/2) of block A (0) and B (1)
//   slice store   C [Ns+1][batch][N]     z(t) = q(t) + i*qdag(Ns - t)
//   work          C [batch][N]           the field between the three axis passes
// where T is the propagator precision (double or float) and C its complex type.
#pragma once

#include <hip/hip_runtime.h>

#include "dft.h"

namespace scft {

enum Axis { AX_X = 0, AX_Y = 1, AX_Z = 2 };

// One batched 1-D convolution pass along an axis:
//   line <- IDFT[ kf(k) * DFT[ line ] ]      (unnormalised DFTs; kf carries 1/n)
// pre:  the loaded value is multiplied by (wre, wim) part-wise first,
// post: the result is multiplied by (wre, wim) part-wise before the store.
template <typename T>
struct PassArgs {
    using C = typename cplx<T>::type;
    const C *src;
    C *dst;
    const T *wre;
    const T *wim;
    const C *tw;   // n entries, (cos 2 pi k/n, -sin 2 pi k/n)
    const T *kf;   // n entries
    long nlines;   // lines in this launch (batch * N / n)
    long N;        // points per instance
    int nx, ny, nz;
    int pre, post;
};

// true if the pass kernels have an instantiation for line length n
bool pass_supported(int n);
// launch one pass along `axis` (n = length of that axis); returns hipError_t
template <typename T>
hipError_t launch_pass(int axis, int n, const PassArgs<T> &a, hipStream_t s);
// the (A, B, lines-per-workgroup, threads) tiling used for (axis, n)
void pass_tiling(int axis, int n, int is_fp32, int *A, int *B, int *lines, int *threads);

}  // namespace scft
