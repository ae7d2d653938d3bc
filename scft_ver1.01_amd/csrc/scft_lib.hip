// scft_lib.hip -- the SCFT solver context, the point-wise and reduction
// kernels around the propagator passes, and the C ABI of include/scft.h.
//
// One SCFT iteration on the device (cuda_scft.cu:167-223 in the reference):
//   k_sums      sums of wA, wB per instance (deterministic two-level reduction)
//   k_center    w -= mean (both fields), partial sums of the new wA
//   k_prep      wA -= mean again (getConc, cuda_scft.cu:71), wdz = exp(-w ds/2),
//               slice 0 = 1 + i
//   Ns x {x pass, y pass, z pass}     the packed march (see below)
//   k_qsum      partial sums of q(Ns)
//   k_contour   Q, then phiA/phiB by the trapezoid rule, FP64 accumulation
//   k_mix_simple / Anderson kernels   field update + free-energy partials
//   k_fe_final  F, F_AB, F_W, F_S, inCompMax per instance
//
// Packed march: the step operator S(q) = W * IFFT[K * FFT[W q]] / N has a
// real, even k-factor, so it maps real fields to real fields and is complex
// linear.  The chain propagator q(s) (s = 0..Ns, block A for s <= NsA) and the
// complementary propagator qdag(s) (from s = Ns down) are therefore marched
// together as z(t) = q(t) + i*qdag(Ns - t), t = 0..Ns, with the real and
// imaginary parts multiplied by their own block's W.  K is separable, so the
// 3-D convolution is three 1-D convolutions (x, y, z passes), each a batched
// in-register FFT (pass_kernels.inc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "scft.h"

namespace scft {

// ---------------------------------------------------------------------------
// reductions (wave64 shuffles, fixed order => bitwise deterministic)
// ---------------------------------------------------------------------------
constexpr int RT = 256;  // threads of every point-wise / reduction kernel

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// block-wide sum of NV values per thread; result broadcast to all threads
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double *sh /* >= NV*RT/64 + NV */)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (l == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i * (RT / 64) + w] = v[i];
    __syncthreads();
    if (threadIdx.x < NV) {
        double r = 0.0;
        for (int k = 0; k < RT / 64; ++k) r += sh[threadIdx.x * (RT / 64) + k];
        sh[NV * (RT / 64) + threadIdx.x] = r;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = sh[NV * (RT / 64) + i];
    __syncthreads();
}

__device__ __forceinline__ double block_max(double v, double *sh)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    v = wave_max(v);
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int k = 1; k < RT / 64; ++k) r = fmax(r, sh[k]);
        sh[RT / 64] = r;
    }
    __syncthreads();
    v = sh[RT / 64];
    __syncthreads();
    return v;
}

// sum of part[0..nblk) (the same tree in every block that calls it)
__device__ __forceinline__ double finalize_sum(const double *part, int nblk, double *sh)
{
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < nblk; k += RT) v[0] += part[k];
    block_sums<1>(v, sh);
    return v[0];
}

// ---------------------------------------------------------------------------
// field kernels (FP64)
// ---------------------------------------------------------------------------
// part[(b*2 + f)*nblk + blk] = partial sum of field f of instance b
__global__ void __launch_bounds__(RT) k_sums(const double *__restrict__ f0, const double *__restrict__ f1,
                                             long N, int nblk, double *__restrict__ part)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    double v[2] = {0.0, 0.0};
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        v[0] += f0[b * N + r];
        if (f1) v[1] += f1[b * N + r];
    }
    block_sums<2>(v, sh);
    if (threadIdx.x == 0) {
        part[(b * 2 + 0) * nblk + blockIdx.x] = v[0];
        part[(b * 2 + 1) * nblk + blockIdx.x] = v[1];
    }
}

// average_value for both fields (minus_average, cuda_aid.cu:376: data - sum/(double)N),
// then the partial sums of the new f0 into part_out[(b*2)*nblk + blk]
__global__ void __launch_bounds__(RT) k_center(double *__restrict__ f0, double *__restrict__ f1, long N, int nblk,
                                               const double *__restrict__ part, double *__restrict__ part_out)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    const double s0 = finalize_sum(part + (b * 2 + 0) * nblk, nblk, sh);
    const double s1 = f1 ? finalize_sum(part + (b * 2 + 1) * nblk, nblk, sh) : 0.0;
    double v[1] = {0.0};
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const double a = f0[b * N + r] - s0 / (double)N;
        f0[b * N + r] = a;
        v[0] += a;
        if (f1) f1[b * N + r] = f1[b * N + r] - s1 / (double)N;
    }
    if (part_out) {
        block_sums<1>(v, sh);
        if (threadIdx.x == 0) part_out[(b * 2) * nblk + blockIdx.x] = v[0];
    }
}

// getConc's own mean subtraction of wA (if part != null), the W = exp(-w ds/2)
// factors of both blocks (initilize_wdz, cuda_aid.cu:254), and slice 0 = 1 + i.
// wdz layout [4][batch][N]: W_A, W_B (applied before the first pass of a step)
// and W_A*post, W_B*post (after the last pass), post = 1/N times the factor of
// the untransformed length-1 axes -- formed in FP64 and rounded once, so the
// k = 0 mode of every pass is multiplied by exactly 1.
template <typename T, bool DELTA>
__global__ void __launch_bounds__(RT) k_prep(double *__restrict__ wA, const double *__restrict__ wB, long N, int nblk,
                                             const double *__restrict__ part, double ds2,
                                             const double *__restrict__ post_b, T *__restrict__ wdz,
                                             typename cplx<T>::type *__restrict__ slot0, int batch)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    const double s0 = part ? finalize_sum(part + (b * 2) * nblk, nblk, sh) : 0.0;
    const double post = post_b[b];
    const long BN = (long)batch * N;
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const long i = b * N + r;
        double a = wA[i];
        if (part) {
            a = a - s0 / (double)N;
            wA[i] = a;
        }
        const double ea = exp(-a * ds2), eb = exp(-wB[i] * ds2);
        wdz[i] = DELTA ? (T)expm1(-a * ds2) : (T)ea;  // DELTA (FP32X): the pre factors as W - 1
        wdz[BN + i] = DELTA ? (T)expm1(-wB[i] * ds2) : (T)eb;
        wdz[2 * BN + i] = (T)(ea * post);
        wdz[3 * BN + i] = (T)(eb * post);
        if (slot0) {  // planar: real parts, then imaginary parts
            reinterpret_cast<T *>(slot0)[i] = T(1);
            reinterpret_cast<T *>(slot0)[BN + i] = T(1);
        }
    }
}

// W (pre) and W*post of one block only (sovDifFft computes W from the field as given)
template <typename T, bool DELTA>
__global__ void __launch_bounds__(RT) k_wdz_one(const double *__restrict__ w, long n, long N, double ds2,
                                                const double *__restrict__ post_b, T *__restrict__ wpre,
                                                T *__restrict__ wpost)
{
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) {
        const double e = exp(-w[i] * ds2);
        wpre[i] = DELTA ? (T)expm1(-w[i] * ds2) : (T)e;  // DELTA (FP32X): the pre factor as W - 1
        wpost[i] = (T)(e * post_b[i / N]);
    }
}

// Complex fields (slices, the work field) are planar: the n = batch*N real
// parts, then the n imaginary parts.
// slice <- (q_init or 1) + 0i
template <typename T>
__global__ void __launch_bounds__(RT) k_seed_real(typename cplx<T>::type *__restrict__ slot,
                                                  const double *__restrict__ q, long n)
{
    T *z = reinterpret_cast<T *>(slot);
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) {
        z[i] = q ? (T)q[i] : T(1);
        z[n + i] = T(0);
    }
}

// FP64 copy of the real part of a slice (its first n entries)
template <typename T>
__global__ void __launch_bounds__(RT) k_take_real(const typename cplx<T>::type *__restrict__ slot,
                                                  double *__restrict__ out, long n)
{
    const T *z = reinterpret_cast<const T *>(slot);
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) out[i] = (double)z[i];
}

// partial sums of Re z(Ns) = q(Ns)   (cal_ql, cuda_aid.cu:119-168).  With
// nsab (per-instance chains) instance b reads slice Ns_b of the store at
// slots; otherwise the slice at slotNs.
template <typename T>
__global__ void __launch_bounds__(RT) k_qsum(const typename cplx<T>::type *__restrict__ slotNs,
                                             const typename cplx<T>::type *__restrict__ slots,
                                             const int2 *__restrict__ nsab, long N, int batch, int nblk,
                                             double *__restrict__ part)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    const typename cplx<T>::type *zs = slotNs;
    if (nsab) zs = slots + (long)(nsab[b].x + nsab[b].y) * batch * N;
    const T *z = reinterpret_cast<const T *>(zs);  // the real parts come first
    double v[1] = {0.0};
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) v[0] += (double)z[b * N + r];
    block_sums<1>(v, sh);
    if (threadIdx.x == 0) part[(b * 2) * nblk + blockIdx.x] = v[0];
}

// Q = sum/N and ffl = ds/Q per instance (cal_ql + host divide, cuda_scft.cu:118-126)
__global__ void __launch_bounds__(RT) k_q_final(const double *__restrict__ part, int nblk, long N, double ds,
                                                double *__restrict__ Qout, double *__restrict__ ffl)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.x;
    const double Q = finalize_sum(part + (b * 2) * nblk, nblk, sh) / N;
    if (threadIdx.x == 0) {
        Qout[b] = Q;
        ffl[b] = ds / Q;
    }
}

// trapezoid weight of contour point s for block A and block B (w_to_phi_go,
// cuda_aid.cu:87-107): A covers s = 0..NsA, B covers s = NsA..Ns, ends 1/2
__device__ __forceinline__ void trap_w(int s, int NsA, int Ns, double &wa, double &wb)
{
    wa = (s < NsA) ? ((s == 0) ? 0.5 : 1.0) : ((s == NsA) ? 0.5 : 0.0);
    wb = (s > NsA) ? ((s == Ns) ? 0.5 : 1.0) : ((s == NsA) ? 0.5 : 0.0);
}

// P consecutive real (or imaginary) parts of a slice, loaded as one access
template <typename T, int P> struct alignas(P * sizeof(T)) RGroup {
    T v[P];
};
// the real and imaginary parts of P consecutive points
template <typename T, int P> struct ZGroup {
    RGroup<T, P> re, im;
};
// The contour reads every slice once: nontemporal 16-byte loads, so that the
// slices streaming through do not evict what the next iteration's march reads
// again (measured 64^3 FP32: contour 75 -> 64 us, iteration 1.69 -> 1.61 ms;
// 128^3 FP64 contour 1253 -> 1154 us)
template <typename T, int P>
__device__ __forceinline__ RGroup<T, P> ld_stream(const RGroup<T, P> *p)
{
    if constexpr (sizeof(RGroup<T, P>) == 16) {
        using V = __attribute__((ext_vector_type(4))) unsigned int;
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
        RGroup<T, P> r;
        __builtin_memcpy(&r, &v, 16);
        return r;
    } else {
        return *p;
    }
}
template <typename T, int P>
__device__ __forceinline__ ZGroup<T, P> ld_group(const RGroup<T, P> *re, long im_off, long gi)
{
    return ZGroup<T, P>{ld_stream<T, P>(re + gi), ld_stream<T, P>(re + im_off + gi)};
}

// phi by the trapezoid rule with FP64 accumulation (w_to_phi_go, cuda_aid.cu:61-118).
// Contour point s pairs q(s) = Re z(s) with qdag(s) = Im z(Ns - s); slices s and
// Ns - s are loaded once for both products.  P consecutive points per thread
// (one 16-byte load per slice: P = 2 in FP32, 1 in FP64; N % P == 0), 4 slice
// pairs in flight.  Each point's sum runs in the order of the reference's
// s-loop restricted to the pairs, the same for every P.  Every block then
// finishes Q = sum/N of k_qsum's block partials qpart (the sum k_q_final
// forms, in the same order) for its factor ds/Q, and block 0 stores Q and
// ds/Q: one launch less per evaluation.
template <typename T, int P, int K = 4>
__global__ void __launch_bounds__(RT) k_contour(const typename cplx<T>::type *__restrict__ slots, long N, int batch,
                                                const int2 *__restrict__ nsab, const double *__restrict__ qpart,
                                                int nblk, double ds, double *__restrict__ Qout,
                                                double *__restrict__ fflout, double *__restrict__ phiA,
                                                double *__restrict__ phiB)
{
    using G = ZGroup<T, P>;
    using RG = RGroup<T, P>;
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    const int NsA = nsab[b].x, Ns = nsab[b].x + nsab[b].y;
    const long g = (long)blockIdx.x * RT + threadIdx.x;
    const bool act = g * P < N;  // the whole block takes part in the Q sum below
    const long BNg = (long)batch * N / P;  // one plane of a slice, in groups
    const RG *zr = reinterpret_cast<const RG *>(slots) + (b * N) / P + g;
    auto zp_at = [&](long s) { return ld_group<T, P>(zr + s * 2 * BNg, BNg, 0); };
    double accA[P], accB[P];
#pragma unroll
    for (int p = 0; p < P; ++p) accA[p] = accB[p] = 0.0;
    if (act) {
        const int half = Ns / 2;  // pairs s = 0..half (s == Ns - s only when Ns is even and s == half)
        int s = 0;
        for (; s + K - 1 <= half; s += K) {
            G zs[K], zu[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                zs[k] = zp_at(s + k);
                zu[k] = zp_at(Ns - s - k);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int s1 = s + k, u1 = Ns - s - k;
                double wa1, wb1, wa2, wb2;
                trap_w(s1, NsA, Ns, wa1, wb1);
                trap_w(u1, NsA, Ns, wa2, wb2);
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const double p1 = (double)zs[k].re.v[p] * (double)zu[k].im.v[p];  // q(s1) qdag(s1)
                    accA[p] += wa1 * p1;
                    accB[p] += wb1 * p1;
                    if (u1 != s1) {
                        const double p2 = (double)zu[k].re.v[p] * (double)zs[k].im.v[p];  // q(u1) qdag(u1)
                        accA[p] += wa2 * p2;
                        accB[p] += wb2 * p2;
                    }
                }
            }
        }
        for (; s <= half; ++s) {
            const int u1 = Ns - s;
            const G a = zp_at(s), c = zp_at(u1);
            double wa1, wb1, wa2, wb2;
            trap_w(s, NsA, Ns, wa1, wb1);
            trap_w(u1, NsA, Ns, wa2, wb2);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double p1 = (double)a.re.v[p] * (double)c.im.v[p];
                accA[p] += wa1 * p1;
                accB[p] += wb1 * p1;
                if (u1 != s) {
                    const double p2 = (double)c.re.v[p] * (double)a.im.v[p];
                    accA[p] += wa2 * p2;
                    accB[p] += wb2 * p2;
                }
            }
        }
    }
    double qv[1] = {0.0};
    for (int k = threadIdx.x; k < nblk; k += RT) qv[0] += qpart[(b * 2) * nblk + k];
    block_sums<1>(qv, sh);
    const double Q = qv[0] / N, f = ds / Q;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Qout[b] = Q;
        fflout[b] = f;
    }
    if (!act) return;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        phiA[b * N + g * P + p] = accA[p] * f;
        phiB[b * N + g * P + p] = accB[p] * f;
    }
}

// Checkpoint mode: the slices of the pairs (tau, Ns - tau), tau = tau0 + k,
// of one k_contour_pairs launch
constexpr int CK_PAIRS = 32;
template <typename C> struct PairSlices {
    const C *zt[CK_PAIRS];
    const C *zu[CK_PAIRS];
};

// phi contributions of the pairs (tau, Ns - tau), tau = tau0 .. tau0 + np - 1,
// summed in FP64 registers and added to phi once (init: assigned); the last
// launch of a contour (ffl non-null) multiplies the total by ds/Q.  P
// consecutive points per thread (one 16-byte load per slice: P = 2 in FP32,
// 1 in FP64; N % P == 0), the slices of 4 pairs in flight.
template <typename T, int P>
__global__ void __launch_bounds__(RT) k_contour_pairs(PairSlices<typename cplx<T>::type> ps, int tau0, int np, int NsA,
                                                      int Ns, long N, long BN, double *__restrict__ phiA,
                                                      double *__restrict__ phiB, int init,
                                                      const double *__restrict__ ffl)
{
    using G = ZGroup<T, P>;
    using RG = RGroup<T, P>;
    const int b = blockIdx.y;
    const long g = (long)blockIdx.x * RT + threadIdx.x;
    if (g * P >= N) return;
    const long gi = b * (N / P) + g;
    const long BNg = BN / P;
    auto at = [&](const typename cplx<T>::type *z) { return ld_group<T, P>(reinterpret_cast<const RG *>(z), BNg, gi); };
    double a[P], c[P];
#pragma unroll
    for (int p = 0; p < P; ++p) a[p] = c[p] = 0.0;
    auto add = [&](int tau, const G &zt, const G &zu) {
        const int u = Ns - tau;
        double wa1, wb1, wa2, wb2;
        trap_w(tau, NsA, Ns, wa1, wb1);
        trap_w(u, NsA, Ns, wa2, wb2);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double p1 = (double)zt.re.v[p] * (double)zu.im.v[p];
            a[p] += wa1 * p1;
            c[p] += wb1 * p1;
            if (u != tau) {
                const double p2 = (double)zu.re.v[p] * (double)zt.im.v[p];
                a[p] += wa2 * p2;
                c[p] += wb2 * p2;
            }
        }
    };
    int k = 0;
    for (; k + 3 < np; k += 4) {
        G zt[4], zu[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            zt[j] = at(ps.zt[k + j]);
            zu[j] = at(ps.zu[k + j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) add(tau0 + k + j, zt[j], zu[j]);
    }
    for (; k < np; ++k) add(tau0 + k, at(ps.zt[k]), at(ps.zu[k]));
    const double f = ffl ? ffl[b] : 1.0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const long i = b * N + g * P + p;
        const double va = (init ? 0.0 : phiA[i]) + a[p], vb = (init ? 0.0 : phiB[i]) + c[p];
        phiA[i] = ffl ? va * f : va;
        phiB[i] = ffl ? vb * f : vb;
    }
}

// partial sums of lamA*phiA + lamB*phiB -> part[(b*2)*nblk + blk]
__global__ void __launch_bounds__(RT) k_dot_lamphi(const double *__restrict__ lamA, const double *__restrict__ lamB,
                                                   const double *__restrict__ phiA, const double *__restrict__ phiB,
                                                   long N, int nblk, double *__restrict__ part)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    double v[1] = {0.0};
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const long i = b * N + r;
        v[0] += lamA[i] * phiA[i] + lamB[i] * phiB[i];
    }
    block_sums<1>(v, sh);
    if (threadIdx.x == 0) part[(b * 2) * nblk + blockIdx.x] = v[0];
}

// L = sum lam.phi + mu ln Q and dL/dQ = (-sum lam.phi + mu)/Q per instance;
// out[b*2] = L, out[b*2+1] = dL/dQ
__global__ void __launch_bounds__(RT) k_adj_scalars(const double *__restrict__ part, int nblk,
                                                    const double *__restrict__ Q, double mu, double *__restrict__ out)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.x;
    const double lphi = finalize_sum(part + (b * 2) * nblk, nblk, sh);
    if (threadIdx.x == 0) {
        out[b * 2 + 0] = lphi + mu * log(Q[b]);
        out[b * 2 + 1] = (-lphi + mu) / Q[b];
    }
}

// Adjoint step tau of the packed march: zb = S(zb_prev) + E(tau) (zb holds the
// propagated part on entry, or nothing when init), then the gradient terms of
// contour points tau and Ns - tau:
//   E(tau) = g(tau) Re z(tau)/Q + i [g(Ns-tau) Im z(tau)/Q + (tau == 0) dL/dQ / N]
//   g(s)   = ds (cA(s) lamA + cB(s) lamB)      (trapezoid weights cA, cB)
//   du_b  += -ds cb(tau)    [Re zb Im z(Ns-tau) - g(tau) Re z(tau) Im z(Ns-tau)/Q]
//   du_b  += -ds cb(Ns-tau)  Im zb Re z(Ns-tau)
// (zb(tau) = qbar_dag(tau) + i qbar(Ns - tau); z(t) = q(t) + i qdag(Ns - t)).
template <typename T>
__global__ void __launch_bounds__(RT) k_adj_point(typename cplx<T>::type *__restrict__ zb,
                                                  const typename cplx<T>::type *__restrict__ zt,
                                                  const typename cplx<T>::type *__restrict__ zu,
                                                  const double *__restrict__ lamA, const double *__restrict__ lamB,
                                                  int tau, int NsA, int Ns, double ds, long N, int batch,
                                                  const double *__restrict__ Q, const double *__restrict__ sc,
                                                  double *__restrict__ gA, double *__restrict__ gB, int init)
{
    using C = typename cplx<T>::type;
    const int b = blockIdx.y;
    const long r = (long)blockIdx.x * RT + threadIdx.x;
    if (r >= N) return;
    const long i = b * N + r;
    const int u = Ns - tau;
    const double q = Q[b];
    const double la = lamA[i], lb = lamB[i];
    double at, bt, au, bu;
    trap_w(tau, NsA, Ns, at, bt);
    trap_w(u, NsA, Ns, au, bu);
    const double gt = ds * (at * la + bt * lb);
    const double gu = ds * (au * la + bu * lb);
    const long BN = (long)batch * N;
    const T *ztp = reinterpret_cast<const T *>(zt), *zup = reinterpret_cast<const T *>(zu);
    T *zbp = reinterpret_cast<T *>(zb);
    const C z_t = C{ztp[i], ztp[BN + i]}, z_u = C{zup[i], zup[BN + i]};
    const double ztr = (double)z_t.x, zti = (double)z_t.y, zur = (double)z_u.x, zui = (double)z_u.y;
    double er = gt * ztr / q;
    double ei = gu * zti / q;
    if (tau == 0) ei += sc[b * 2 + 1] / (double)N;
    double zbr = er, zbi = ei;
    if (!init) {
        zbr += (double)zbp[i];
        zbi += (double)zbp[BN + i];
    }
    zbp[i] = (T)zbr;
    zbp[BN + i] = (T)zbi;
    const double t1 = zbr * zui - gt * ztr * zui / q;  // qbar_dag(tau) qdag(tau) - g q qdag / Q
    const double t2 = zbi * zur;                       // qbar(u) q(u)
    double da = -ds * (at * t1 + au * t2);
    double db = -ds * (bt * t1 + bu * t2);
    if (!init) {
        da += gA[i];
        db += gB[i];
    }
    gA[i] = da;
    gB[i] = db;
}

// free-energy partials with the updated fields: fpart[(b*3 + j)*nblk + blk],
// j = 0: sum chiN*phiA*phiB, 1: -sum(wA phiA + wB phiB), 2: max|1 - phiA - phiB|
__device__ __forceinline__ void fe_partials(double fab, double fw, double inc, double *sh, double *fpart, int b,
                                            int nblk)
{
    double v[2] = {fab, fw};
    block_sums<2>(v, sh);
    const double m = block_max(inc, sh);
    if (threadIdx.x == 0) {
        fpart[(b * 3 + 0) * nblk + blockIdx.x] = v[0];
        fpart[(b * 3 + 1) * nblk + blockIdx.x] = v[1];
        fpart[(b * 3 + 2) * nblk + blockIdx.x] = m;
    }
}

// phi_w (cuda_aid.cu:382-416) + F partials (cuda_scft.cu:202-210)
__global__ void __launch_bounds__(RT) k_mix_simple(double *__restrict__ wA, double *__restrict__ wB,
                                                   const double *__restrict__ phiA, const double *__restrict__ phiB,
                                                   long N, int nblk, const double *__restrict__ chiN, double wopt,
                                                   double wcmp, double *__restrict__ fpart)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.y;
    const double hAB = chiN[b];
    double fab = 0.0, fw = 0.0, inc = 0.0;
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const long i = b * N + r;
        const double pa = phiA[i], pb = phiB[i];
        double a = wA[i], c = wB[i];
        const double eta = (a + c - hAB) / 2;
        const double psum = 1.0 - pa - pb;
        double waDiff = hAB * pb + eta - a;
        double wbDiff = hAB * pa + eta - c;
        waDiff -= wcmp * psum;
        wbDiff -= wcmp * psum;
        a += wopt * waDiff;
        c += wopt * wbDiff;
        wA[i] = a;
        wB[i] = c;
        inc = fmax(inc, fabs(1 - pa - pb));
        fab = fab + hAB * pa * pb;
        fw = fw - (a * pa + c * pb);
    }
    fe_partials(fab, fw, inc, sh, fpart, b, nblk);
}

// fe[b*5 + .] = F, F_AB, F_W, F_S, inCompMax  (cuda_scft.cu:212-218)
__global__ void __launch_bounds__(RT) k_fe_final(const double *__restrict__ fpart, int nblk, const double *__restrict__ Q,
                                                 long N, double *__restrict__ fe, long *and_cnt)
{
    __shared__ double sh[2 * RT / 64 + 2];
    const int b = blockIdx.x;
    const double fab = finalize_sum(fpart + (b * 3 + 0) * nblk, nblk, sh);
    const double fw = finalize_sum(fpart + (b * 3 + 1) * nblk, nblk, sh);
    double inc = 0.0;
    for (int k = threadIdx.x; k < nblk; k += RT) inc = fmax(inc, fpart[(b * 3 + 2) * nblk + k]);
    inc = block_max(inc, sh);
    if (threadIdx.x == 0) {
        const double freeAB = fab / N;
        const double freeW = fw / N;
        const double freeS = -log(Q[b]);
        fe[b * 5 + 0] = freeAB + freeW + freeS;
        fe[b * 5 + 1] = freeAB;
        fe[b * 5 + 2] = freeW;
        fe[b * 5 + 3] = freeS;
        fe[b * 5 + 4] = inc;
        if (and_cnt && b == 0) *and_cnt += 1;  // one Anderson update done (all instances in lockstep)
    }
}

// Up to SEG_MAX contiguous FP64 segments copied by one launch (the iteration
// state snapshots of scft_iterate with a tolerance).
constexpr int SEG_MAX = 12;
struct CopySegs {
    const double *src[SEG_MAX];
    double *dst[SEG_MAX];
    long n[SEG_MAX];
    int count;
};
__global__ void __launch_bounds__(RT) k_copy_segs(CopySegs s)
{
    for (int k = 0; k < s.count; ++k) {
        const double *__restrict__ src = s.src[k];
        double *__restrict__ dst = s.dst[k];
        for (long i = (long)blockIdx.x * RT + threadIdx.x; i < s.n[k]; i += (long)gridDim.x * RT) dst[i] = src[i];
    }
}

// ---------------------------------------------------------------------------
// Anderson mixing (no reference counterpart).  History ring of m+1 slots of
// (w, d) per instance, FP64.  Residual (the direction simple mixing steps in):
//   eta = (wA+wB-chiN)/2, psi = 1-phiA-phiB,
//   dA = chiN phiB + eta - wA - wcmp psi,  dB = chiN phiA + eta - wB - wcmp psi.
// Gram matrix G[i][j] = <d_i, d_j> over both fields; for the newest slot c and
// history slots s_1..s_m':  U_ij = G_cc - G_cs_i - G_cs_j + G_s_i s_j,
// V_i = G_cc - G_cs_i; U c = V by Gaussian elimination with partial pivoting;
// w+ = w_c + sum c_i (w_si - w_c) + mix (d_c + sum c_i (d_si - d_c)).
// ---------------------------------------------------------------------------
constexpr int AND_MAX = 32;

// Ring bookkeeping from the device-side update counter *cnt (updates done so
// far): newest slot s[0] = cnt mod (m+1), history s[k] = (cnt - k) mod (m+1),
// mp = min(cnt, m) history entries.  Reading the counter on the device keeps
// every kernel argument fixed from one iteration to the next.
struct AndRing {
    int mp;
    int s[AND_MAX + 1];
};
__device__ __forceinline__ AndRing and_ring(const long *cnt, int m)
{
    AndRing r;
    const long c = *cnt;
    const int m1 = m + 1;
    r.mp = (int)(c < m ? c : m);
    const int cur = (int)(c % m1);
    for (int k = 0; k <= AND_MAX; ++k) r.s[k] = ((cur - k) % m1 + m1) % m1;
    return r;
}

// residual into the newest ring slot; partial inner products
// part[(b*(MAXM+2) + j)*nblk + blk]: j < mp: <d, d_{s[j+1]}>, j = MAXM: <d,d>, j = MAXM+1: <w,w>
template <int MAXM>
__global__ void __launch_bounds__(RT) k_and_resid(const double *__restrict__ wA, const double *__restrict__ wB,
                                                  const double *__restrict__ phiA, const double *__restrict__ phiB,
                                                  long N, int batch, int nblk, const double *__restrict__ chiN,
                                                  double wcmp, double mix, double *__restrict__ wh,
                                                  double *__restrict__ dh, const long *cnt, int m,
                                                  double *__restrict__ part)
{
    __shared__ double sh[(MAXM + 2) * RT / 64 + MAXM + 2 + 2];
    const int b = blockIdx.y;
    const double hAB = chiN[b];
    const AndRing rg = and_ring(cnt, m);
    const long slot = 2 * (long)batch * N;  // one ring slot: [batch][2][N]
    double acc[MAXM + 2];
#pragma unroll
    for (int j = 0; j < MAXM + 2; ++j) acc[j] = 0.0;
    double *wc = wh + rg.s[0] * slot + b * 2 * N;
    double *dc = dh + rg.s[0] * slot + b * 2 * N;
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const long i = b * N + r;
        const double a = wA[i], c = wB[i], pa = phiA[i], pb = phiB[i];
        const double eta = (a + c - hAB) / 2;
        const double psum = 1.0 - pa - pb;
        const double dA = hAB * pb + eta - a - wcmp * psum;
        const double dB = hAB * pa + eta - c - wcmp * psum;
        // the field ring holds g = w + mix * d, all the update needs of the
        // pair (w, d): w_new = g_0 + sum_k c_k (g_k - g_0), equal to
        // (w_0 + sum c_k (w_k - w_0)) + mix * (d_0 + sum c_k (d_k - d_0))
        wc[r] = a + mix * dA;
        wc[N + r] = c + mix * dB;
        dc[r] = dA;
        dc[N + r] = dB;
        // the history slots in groups of four, each group's loads issued
        // together; a group's slots past the history re-read its first slot
        // (a cache hit) and add nothing.  Measured at 128^3 with m = 20: one
        // round trip per slot 217 us, groups of four 178 us, all 32 slots'
        // loads at once 256 us (the slots past the history read from memory)
#pragma unroll
        for (int j0 = 0; j0 < MAXM; j0 += 4) {
            if (j0 < rg.mp) {
                double xa[4], xb[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int sj = (j0 + u < rg.mp) ? rg.s[j0 + u + 1] : rg.s[j0 + 1];
                    const double *dj = dh + sj * slot + b * 2 * N;
                    xa[u] = dj[r];
                    xb[u] = dj[N + r];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j0 + u < rg.mp) acc[j0 + u] += dA * xa[u] + dB * xb[u];
            }
        }
        acc[MAXM] += dA * dA + dB * dB;
        acc[MAXM + 1] += a * a + c * c;
    }
    block_sums<MAXM + 2>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < MAXM + 2; ++j) part[(b * (MAXM + 2) + j) * nblk + blockIdx.x] = acc[j];
}

// One workgroup per instance: finish the new Gram row, solve, write the
// coefficients.  The Gram-row sums take a wave each and the elements of G
// and U are spread over the workgroup; the elimination runs in the first wave
// (a lane per row, the pivot found by a wave argmax) and the back
// substitution on one lane.  Each element sees the floating-point operations
// of the serial elimination in the same order.
__global__ void __launch_bounds__(RT) k_and_solve(const double *__restrict__ part, int nblk, int maxm, int m,
                                                  const long *cnt, double *__restrict__ G, double *__restrict__ coef,
                                                  double *__restrict__ err, double ridge)
{
    __shared__ double row[AND_MAX + 2];
    __shared__ double Gs[(AND_MAX + 1) * (AND_MAX + 1)];
    __shared__ double U[AND_MAX * AND_MAX];
    __shared__ double V[AND_MAX];
    __shared__ double F[AND_MAX];
    __shared__ int ok_s;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const AndRing rg = and_ring(cnt, m);
    const int mp = rg.mp, m1 = m + 1;
    // the new Gram row (j < mp), |d|^2 (maxm) and |w|^2 (maxm + 1): one wave per sum
    for (int q = wv; q < mp + 2; q += RT / 64) {
        const int j = q < mp ? q : maxm + (q - mp);
        const double *pj = part + (b * (maxm + 2) + j) * nblk;
        double v = 0.0;
#pragma unroll 8
        for (int k = lane; k < nblk; k += 64) v += pj[k];  // eight loads in flight, the sum in order
        v = wave_sum(v);
        if (lane == 0) row[j] = v;
    }
    double *g = G + (long)b * m1 * m1;
    for (int e = tid; e < m1 * m1; e += RT) Gs[e] = g[e];
    __syncthreads();
    const int c = rg.s[0];
    if (tid < mp) {
        const int sj = rg.s[tid + 1];
        Gs[c * m1 + sj] = Gs[sj * m1 + c] = row[tid];
        g[c * m1 + sj] = g[sj * m1 + c] = row[tid];
    }
    if (tid == 0) {
        Gs[c * m1 + c] = g[c * m1 + c] = row[maxm];
        err[b] = row[maxm + 1] > 0.0 ? sqrt(row[maxm] / row[maxm + 1]) : sqrt(row[maxm]);
    }
    __syncthreads();
    const double gcc = Gs[c * m1 + c];
    for (int e = tid; e < mp * mp; e += RT) {
        const int si = rg.s[e / mp + 1], sj = rg.s[e % mp + 1];
        U[e] = gcc - Gs[c * m1 + si] - Gs[c * m1 + sj] + Gs[si * m1 + sj];
    }
    if (tid < mp) V[tid] = gcc - Gs[c * m1 + rg.s[tid + 1]];
    __syncthreads();
    if (ridge > 0.0 && mp > 0) {  // U += ridge * tr(U)/mp * I
        if (tid == 0) {
            double tr = 0.0;
            for (int i = 0; i < mp; ++i) tr += U[i * mp + i];
            F[0] = ridge * tr / mp;
        }
        __syncthreads();
        const double lam = F[0];
        __syncthreads();
        if (tid < mp) U[tid * mp + tid] += lam;
    }
    // Gaussian elimination with partial pivoting (the oracle's steps) in the
    // first wave alone: lane i holds row i, the pivot is a wave argmax and the
    // LDS updates need only wave-scope ordering (with the whole workgroup,
    // every step took four barriers and a serial pivot search)
    __syncthreads();  // U and V are complete
    if (wv == 0) {
        auto wave_sync = [] {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        int ok = 1;
        for (int k = 0; k < mp; ++k) {
            // the largest |U[i][k]|, i >= k; the first such row on ties
            double best = (lane >= k && lane < mp) ? fabs(U[lane * mp + k]) : -1.0;
            int piv = lane;
            for (int off = 32; off >= 1; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int op = __shfl_xor(piv, off);
                if (ob > best || (ob == best && op < piv)) {
                    best = ob;
                    piv = op;
                }
            }
            if (best == 0.0) {  // uniform: every lane holds the reduced value
                ok = 0;
                break;
            }
            if (piv != k) {
                if (lane < mp) {
                    const double t = U[k * mp + lane];
                    U[k * mp + lane] = U[piv * mp + lane];
                    U[piv * mp + lane] = t;
                }
                if (lane == 0) {
                    const double t = V[k];
                    V[k] = V[piv];
                    V[piv] = t;
                }
                wave_sync();
            }
            if (lane > k && lane < mp) {
                const double f = U[lane * mp + k] / U[k * mp + k];
#pragma unroll 4
                for (int j = k; j < mp; ++j) U[lane * mp + j] -= f * U[k * mp + j];
                V[lane] -= f * V[k];
            }
            wave_sync();
        }
        if (ok && lane == 0)
            for (int k = mp - 1; k >= 0; --k) {
                double sum = V[k];
                for (int j = k + 1; j < mp; ++j) sum -= U[k * mp + j] * V[j];
                V[k] = sum / U[k * mp + k];
            }
        if (lane == 0) ok_s = ok;
    }
    __syncthreads();
    const bool ok = ok_s != 0;
    for (int i = tid; i < AND_MAX; i += RT) coef[b * AND_MAX + i] = (ok && i < mp) ? V[i] : 0.0;
}

// w+ from the ring, then F partials with the updated fields
__global__ void __launch_bounds__(RT) k_and_update(double *__restrict__ wA, double *__restrict__ wB,
                                                   const double *__restrict__ phiA, const double *__restrict__ phiB,
                                                   long N, int batch, int nblk, int m, const long *cnt,
                                                   const double *__restrict__ chiN, const double *__restrict__ gh,
                                                   const double *__restrict__ coef, double *__restrict__ fpart)
{
    __shared__ double sh[2 * RT / 64 + 2];
    __shared__ double cs[AND_MAX];
    const int b = blockIdx.y;
    const double hAB = chiN[b];
    const AndRing rg = and_ring(cnt, m);
    const int mp = rg.mp;
    if (threadIdx.x < AND_MAX) cs[threadIdx.x] = coef[b * AND_MAX + threadIdx.x];
    __syncthreads();
    const long slot = 2 * (long)batch * N;
    const double *g0 = gh + rg.s[0] * slot + b * 2 * N;  // g = w + mix * d (k_and_resid)
    double fab = 0.0, fw = 0.0, inc = 0.0;
    for (long r = (long)blockIdx.x * RT + threadIdx.x; r < N; r += (long)nblk * RT) {
        const double ga0 = g0[r], gb0 = g0[N + r];
        double a = ga0, cB = gb0;
        // four history slots' loads in flight per round (the sums keep their order)
#pragma unroll 4
        for (int k = 0; k < mp; ++k) {
            const double *gk = gh + rg.s[k + 1] * slot + b * 2 * N;
            const double c = cs[k];
            a += c * (gk[r] - ga0);
            cB += c * (gk[N + r] - gb0);
        }
        const long i = b * N + r;
        wA[i] = a;
        wB[i] = cB;
        const double pa = phiA[i], pb = phiB[i];
        inc = fmax(inc, fabs(1 - pa - pb));
        fab = fab + hAB * pa * pb;
        fw = fw - (a * pa + cB * pb);
    }
    fe_partials(fab, fw, inc, sh, fpart, b, nblk);
}

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
struct SolverBase {
    virtual ~SolverBase() = default;
};

}  // namespace scft

// ===========================================================================
// C ABI
// ===========================================================================
using namespace scft;

static thread_local std::string g_create_error;

static long slot_count(int Ns, int ckpt, int *K_out, int *nck_out);

struct scft_ctx {
    scft_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    int NsA = 0, NsB = 0, Ns = 0;
    long N = 0;
    int batch = 1;
    int fp32 = 0;   // FP32 storage of the propagator (SCFT_FP32 and SCFT_FP32X)
    int mixed = 0;  // SCFT_FP32X: FP64 transform arithmetic over the FP32 storage
    int n[3] = {1, 1, 1};
    int active[3] = {0, 0, 0};
    int first_axis = -1, last_axis = -1;
    int use_plane = 0;  // x and y passes in one kernel: 0 no, 1 xy-plane, 2 split xy-plane
    int ckpt = 0;       // sqrt(Ns) checkpointing of the slice store
    int ck_K = 1, ck_n = 1;  // checkpoint spacing and count
    // adjoint scratch (allocated on first use): uA, uB, lamA, lamB, gA, gB, phiA, phiB
    double *adj = nullptr;
    // scft_adjoint_loop's trajectory and reduction buffers (grown on demand)
    double *loop = nullptr;
    size_t loop_bytes = 0;
    // per-instance chain parameters (equal entries for a uniform context)
    int2 *nsab = nullptr;      // device (NsA, NsB) per instance
    double *chiN_b = nullptr;  // device chiN per instance
    bool kf_per_inst = false;  // one k table per instance (per-instance box) or one shared
    int post_exact = 0;        // post W factors equal pre / N exactly (see upload_tables)
    double *post_b = nullptr;  // device per-instance post-step factor (1/N x length-1 axis factors)
    int Ns_max = 0;
    std::vector<int2> nsab_h;  // host copy of nsab
    bool blocks_uniform = true;  // every instance has the same (NsA, NsB)
    bool own_stream = true;
    int pass_variant[3] = {0, 0, 0};  // tiling variant per axis (autotuned)
    int nblk = 1;
    size_t csize = 16;  // bytes per complex element
    size_t tsize = 8;   // bytes per real element
    // device buffers
    double *wA = nullptr, *wB = nullptr, *phiA = nullptr, *phiB = nullptr;
    void *wdz = nullptr;    // T [4][batch][N]: W_A, W_B, W_A*post, W_B*post (FP32X: W_A - 1, W_B - 1, ...)
    double post = 1.0;      // 1/N times the factor of the length-1 axes
    void *slots = nullptr;  // C [Ns+1][batch][N]
    void *work = nullptr;   // C [batch][N]
    double *part = nullptr, *part2 = nullptr, *fpart = nullptr;
    double *Q = nullptr, *fe = nullptr, *ffl = nullptr;
    void *tw[3] = {nullptr, nullptr, nullptr};
    void *kf[3] = {nullptr, nullptr, nullptr};
    // FP32-storage contexts: the twiddle and k tables of the other FP32-storage
    // form (FP64 tables for SCFT_FP32, FP32 ones for SCFT_FP32X)
    void *tw_alt[3] = {nullptr, nullptr, nullptr};
    void *kf_alt[3] = {nullptr, nullptr, nullptr};
    // Anderson: rings of g = w + mix * d (wh) and of the residuals d (dh), [m+1][batch][2][N]
    double *wh = nullptr, *dh = nullptr, *G = nullptr, *coef = nullptr, *aerr = nullptr, *apart = nullptr;
    long *and_cnt = nullptr;  // device: Anderson updates done since the last set_fields
    int and_maxm = 0;         // accumulator bucket (8, 16 or 32) >= anderson_m
    // one SCFT iteration captured as a HIP graph, and two in a row (built on
    // first use): runs of iterations replay the pair graph, which saves one
    // graph launch boundary per pair (64^3 FP32 iteration 1.592 -> 1.588 ms)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph2 = nullptr;
    hipGraphExec_t graph2_exec = nullptr;
    int have_conc = 0;
    int have_fe = 0;
    long device_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // scft_iterate(tol > 0): two device snapshots of the iteration state, pinned
    // host slots for two iterations' free energies, and their events
    double *tol_bk[2] = {nullptr, nullptr};
    double *tol_fe = nullptr;
    hipEvent_t tol_ev[2] = {nullptr, nullptr};
    std::vector<void *> allocs;
    std::string err;
};

static int set_err(scft_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPX(ctx, call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return set_err((ctx), SCFT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

static int dalloc(scft_ctx *c, void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return set_err(c, SCFT_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " +
                                            hipGetErrorString(e));
    }
    c->allocs.push_back(*ptr);
    c->device_bytes += (long)bytes;
    return SCFT_OK;
}

extern "C" void scft_default_params(scft_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 32;
    p->lx = p->ly = p->lz = 4.0;
    p->chiN = 15.0;
    p->fA = 0.5;
    p->ds = 0.01;
    p->batch = 1;
    p->precision = SCFT_FP64;
    p->mixing = SCFT_MIX_SIMPLE;
    p->wopt = 0.050;
    p->wcmp = 0.10;
    p->anderson_m = 10;
    p->anderson_mix = 0.1;
    p->pi = SCFT_REFERENCE_PI;
    p->use_graph = 1;
    p->xy_plane = -1;
    p->fuse_steps = 1;
    p->anderson_reg = 0.0;
    p->tiling = -1;
    p->checkpoint = 0;
}

// init_cuda.cu:326-331
extern "C" int scft_contour_counts(double fA, double ds, int *NsA, int *NsB)
{
    if (!NsA || !NsB || !(ds > 0.0) || !std::isfinite(fA) || !std::isfinite(ds)) return SCFT_E_INVALID;
    const double dfB = 1 - fA;
    *NsA = ((int)(fA / ds + 1.0e-8));
    *NsB = ((int)(dfB / ds + 1.0e-8));
    return SCFT_OK;
}

extern "C" const char *scft_last_error(const scft_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

// the captured iteration graphs, dropped when what they launch changes
static void drop_graphs(scft_ctx *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->graph2_exec) (void)hipGraphExecDestroy(c->graph2_exec);
    if (c->graph2) (void)hipGraphDestroy(c->graph2);
    c->graph_exec = c->graph2_exec = nullptr;
    c->graph = c->graph2 = nullptr;
}

static void free_all(scft_ctx *c)
{
    drop_graphs(c);
    for (void *ptr : c->allocs) (void)hipFree(ptr);
    c->allocs.clear();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int k = 0; k < 2; ++k) {
        if (c->tol_ev[k]) (void)hipEventDestroy(c->tol_ev[k]);
        c->tol_ev[k] = nullptr;
        c->tol_bk[k] = nullptr;  // freed with c->allocs
    }
    if (c->tol_fe) (void)hipHostFree(c->tol_fe);
    c->tol_fe = nullptr;
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    c->ev0 = c->ev1 = nullptr;
    c->stream = nullptr;
}

extern "C" int scft_destroy(scft_ctx *ctx)
{
    if (!ctx) return SCFT_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return SCFT_OK;
}

// k^2 of one axis exactly as init_cuda.cu:408-418 (dx = l/(double)n)
static void ksq_axis(int n, double l, double pi, std::vector<double> &ksq)
{
    const double dx = l / (double)n;
    ksq.assign(n, 0.0);
    for (int i = 0; i <= n / 2 - 1; i++) ksq[i] = 2 * pi * i * 1.0 / n / dx;
    for (int i = n / 2; i < n; i++) ksq[i] = 2 * pi * (i - n) * 1.0 / dx / n;
    for (int i = 0; i < n; i++) ksq[i] *= ksq[i];
}

// k-factor and twiddle tables.  inst_L (per instance, may be null) replaces
// the box lengths of instance b by (L_b, L_b, L_b); then every instance gets
// its own k table (kf_stride = n) and its own post factor.
static int upload_tables(scft_ctx *c, const double *inst_L)
{
    const double pi = c->p.pi;
    const int nb = inst_L ? c->batch : 1;
    std::vector<double> post(c->batch);
    std::vector<std::vector<double>> kfa(3);
    for (int b = 0; b < nb; ++b) {
        const double ls[3] = {inst_L ? inst_L[b] : c->p.lx, inst_L ? inst_L[b] : c->p.ly, inst_L ? inst_L[b] : c->p.lz};
        std::vector<double> k2[3];
        for (int a = 0; a < 3; ++a) ksq_axis(c->n[a], ls[a], pi, k2[a]);
        // axes of length 1 are not transformed; their single factor exp(-ds k^2[0])
        // (nonzero: the reference's table gives k = -2 pi/l there) and the inverse
        // transforms' 1/N go into the post-step factor, so kf[0] = 1 exactly
        double fold = 1.0;
        for (int a = 0; a < 3; ++a)
            if (!c->active[a]) fold *= std::exp(-c->p.ds * k2[a][0]);
        if (inst_L) post[b] = fold / (double)c->N;
        else std::fill(post.begin(), post.end(), fold / (double)c->N);
        for (int a = 0; a < 3; ++a)
            if (c->active[a])
                for (int k = 0; k < c->n[a]; ++k) kfa[a].push_back(std::exp(-c->p.ds * k2[a][k]));
    }
    c->post = post[0];
    c->kf_per_inst = inst_L != nullptr;
    // post factor = pre factor / N bit for bit when 1/N is a power of two and no axis is folded in
    c->post_exact = (c->N & (c->N - 1)) == 0 &&
                    std::all_of(post.begin(), post.end(), [&](double v) { return v == 1.0 / (double)c->N; });
    int rc;
    if ((rc = dalloc(c, (void **)&c->post_b, sizeof(double) * c->batch))) return rc;
    HIPX(c, hipMemcpy(c->post_b, post.data(), sizeof(double) * c->batch, hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a) {
        if (!c->active[a]) continue;
        const int n = c->n[a];
        const size_t nk = kfa[a].size();  // n or batch * n
        std::vector<double> twr(n), twi(n);
        for (int k = 0; k < n; ++k) {
            // exact-angle twiddles, long double then rounded
            const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
            twr[k] = (double)cosl(ang);
            twi[k] = (double)-sinl(ang);
        }
        // FP32 arithmetic tables (SCFT_FP32) and FP64 ones (SCFT_FP64, and SCFT_FP32X, whose
        // transforms run in FP64); an FP32-storage context holds both, the other form's in
        // tw_alt / kf_alt (scft_set_precision swaps them)
        void *f_tw = nullptr, *f_kf = nullptr, *d_tw = nullptr, *d_kf = nullptr;
        if (c->fp32) {
            std::vector<float> kff(nk), tw(2 * n);
            for (size_t k = 0; k < nk; ++k) kff[k] = (float)kfa[a][k];
            for (int k = 0; k < n; ++k) {
                tw[2 * k] = (float)twr[k];
                tw[2 * k + 1] = (float)twi[k];
            }
            if ((rc = dalloc(c, &f_kf, sizeof(float) * nk))) return rc;
            if ((rc = dalloc(c, &f_tw, sizeof(float) * 2 * n))) return rc;
            HIPX(c, hipMemcpy(f_kf, kff.data(), sizeof(float) * nk, hipMemcpyHostToDevice));
            HIPX(c, hipMemcpy(f_tw, tw.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
        }
        {
            std::vector<double> tw(2 * n);
            for (int k = 0; k < n; ++k) {
                tw[2 * k] = twr[k];
                tw[2 * k + 1] = twi[k];
            }
            if ((rc = dalloc(c, &d_kf, sizeof(double) * nk))) return rc;
            if ((rc = dalloc(c, &d_tw, sizeof(double) * 2 * n))) return rc;
            HIPX(c, hipMemcpy(d_kf, kfa[a].data(), sizeof(double) * nk, hipMemcpyHostToDevice));
            HIPX(c, hipMemcpy(d_tw, tw.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
        }
        const bool ftab = c->fp32 && !c->mixed;
        c->tw[a] = ftab ? f_tw : d_tw;
        c->kf[a] = ftab ? f_kf : d_kf;
        c->tw_alt[a] = ftab ? d_tw : f_tw;
        c->kf_alt[a] = ftab ? d_kf : f_kf;
    }
    return SCFT_OK;
}

// true if xy-plane form `form` (1 xy-plane, 2 split xy-plane, 3 split with the
// three-level tiling) has a kernel for
// an nx x ny plane in this precision that fits a CU's LDS
static constexpr int PLANE_FORMS = 5;  // xy-plane forms 1..5 (0: separate x and y passes)

// pass-timing codes (scft_time_pass): 0..2 the axis passes, 5..7 their
// step-boundary kernels; xy-plane form f has a pass code and a boundary code
static constexpr int kPlaneCode[PLANE_FORMS + 1][2] = {{-1, -1}, {3, 8}, {4, 9}, {10, 11}, {12, 13}, {14, 15}};
static constexpr int CONTOUR_CODE = 16;  // the contour kernel over the full slice store
static constexpr int PASS_CODES = 17;
static void code_form(int code, int *form, bool *fused)
{
    *form = 0;
    *fused = code >= 5 && code <= 7;
    for (int f = 1; f <= PLANE_FORMS; ++f)
        for (int k = 0; k < 2; ++k)
            if (kPlaneCode[f][k] == code) {
                *form = f;
                *fused = k == 1;
            }
}

static bool plane_fits(const scft_ctx *c, int nx, int ny, int form)
{
    if (nx != ny || nx < 2 || form < 1 || form > PLANE_FORMS) return false;
    if (c->mixed) return scft_f32x::plane_form_supported<float>(form, nx);
    return c->fp32 ? plane_form_supported<float>(form, nx) : plane_form_supported<double>(form, nx);
}

// The pass launchers of the context's form: T = float selects the FP32
// arithmetic kernels or, for SCFT_FP32X, the FP64-arithmetic ones over the
// same FP32 storage (whose twiddle and k tables are FP64).
template <typename T>
static hipError_t run_pass(const scft_ctx *c, int axis, int n, int variant, const PassArgs<T> &a, bool fused)
{
    if constexpr (std::is_same<T, float>::value) {
        if (c->mixed)
            return fused ? scft_f32x::launch_fused<float>(axis, n, variant, a, c->stream)
                         : scft_f32x::launch_pass<float>(axis, n, variant, a, c->stream);
    }
    return fused ? launch_fused<T>(axis, n, variant, a, c->stream) : launch_pass<T>(axis, n, variant, a, c->stream);
}
template <typename T>
static hipError_t run_plane(const scft_ctx *c, int form, const PassArgs<T> &a, const typename cplx<T>::type *twy,
                            const typename kfac<T>::type *kfy, long np, bool fused)
{
    if constexpr (std::is_same<T, float>::value) {
        if (c->mixed)
            return fused ? scft_f32x::launch_plane_fused<float>(form, c->n[0], a, twy, kfy, np, c->stream)
                         : scft_f32x::launch_plane_form<float>(form, c->n[0], a, twy, kfy, np, c->stream);
    }
    return fused ? launch_plane_fused<T>(form, c->n[0], a, twy, kfy, np, c->stream)
                 : launch_plane_form<T>(form, c->n[0], a, twy, kfy, np, c->stream);
}

// Context constructor.  inst_L / inst_chiN / inst_fA (length p.batch, each may
// be null) give per-instance box lengths (cubic L, L, L), chiN and fA in place
// of the uniform ones in p; stream, when not null, is shared (not owned).
static int create_ctx(const scft_params *pin, const double *inst_L, const double *inst_chiN, const double *inst_fA,
                      int device, hipStream_t stream, scft_ctx **out)
{
    if (!out) return set_err(nullptr, SCFT_E_INVALID, "scft_create: out is NULL");
    *out = nullptr;
    if (!pin) return set_err(nullptr, SCFT_E_INVALID, "scft_create: params is NULL");
    scft_params p = *pin;
    if (p.pi == 0.0) p.pi = SCFT_REFERENCE_PI;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) return set_err(nullptr, SCFT_E_INVALID, "grid sizes must be >= 1");
    if (!(p.lx > 0 && p.ly > 0 && p.lz > 0) || !std::isfinite(p.lx) || !std::isfinite(p.ly) || !std::isfinite(p.lz))
        return set_err(nullptr, SCFT_E_INVALID, "box lengths must be finite and > 0");
    if (!(p.ds > 0.0) || !(p.ds < 0.5)) return set_err(nullptr, SCFT_E_INVALID, "ds must be in (0, 0.5)");
    if (!(p.fA > 0.0 && p.fA < 1.0)) return set_err(nullptr, SCFT_E_INVALID, "fA must be in (0, 1)");
    if (!std::isfinite(p.chiN)) return set_err(nullptr, SCFT_E_INVALID, "chiN must be finite");
    if (p.batch < 1) return set_err(nullptr, SCFT_E_INVALID, "batch must be >= 1");
    if (p.precision != SCFT_FP64 && p.precision != SCFT_FP32 && p.precision != SCFT_FP32X)
        return set_err(nullptr, SCFT_E_INVALID, "precision must be SCFT_FP64, SCFT_FP32 or SCFT_FP32X");
    if (p.mixing != SCFT_MIX_SIMPLE && p.mixing != SCFT_MIX_ANDERSON)
        return set_err(nullptr, SCFT_E_INVALID, "mixing must be SCFT_MIX_SIMPLE or SCFT_MIX_ANDERSON");
    if (p.mixing == SCFT_MIX_ANDERSON && (p.anderson_m < 0 || p.anderson_m > AND_MAX))
        return set_err(nullptr, SCFT_E_INVALID, "anderson_m must be in [0, 32]");
    int NsA, NsB;
    scft_contour_counts(p.fA, p.ds, &NsA, &NsB);
    if (NsA < 1 || NsB < 1)
        return set_err(nullptr, SCFT_E_INVALID, "both blocks need at least one contour step (fA/ds, (1-fA)/ds >= 1)");
    std::vector<int2> nsab(p.batch, int2{NsA, NsB});
    std::vector<double> chis(p.batch, p.chiN);
    int Ns_max = NsA + NsB;
    bool ns_uniform = true;
    for (int b = 0; b < p.batch; ++b) {
        if (inst_fA) {
            if (!(inst_fA[b] > 0.0 && inst_fA[b] < 1.0))
                return set_err(nullptr, SCFT_E_INVALID, "instance fA must be in (0, 1)");
            int a1, b1;
            scft_contour_counts(inst_fA[b], p.ds, &a1, &b1);
            if (a1 < 1 || b1 < 1)
                return set_err(nullptr, SCFT_E_INVALID, "instance " + std::to_string(b) + ": a block has no contour step");
            nsab[b] = int2{a1, b1};
            Ns_max = std::max(Ns_max, a1 + b1);
            ns_uniform = ns_uniform && (a1 + b1 == NsA + NsB);
        }
        if (inst_chiN) {
            if (!std::isfinite(inst_chiN[b])) return set_err(nullptr, SCFT_E_INVALID, "instance chiN must be finite");
            chis[b] = inst_chiN[b];
        }
        if (inst_L && !(inst_L[b] > 0.0 && std::isfinite(inst_L[b])))
            return set_err(nullptr, SCFT_E_INVALID, "instance box length must be finite and > 0");
    }
    if (inst_fA) {
        // the context-level counts describe instance 0
        NsA = nsab[0].x;
        NsB = nsab[0].y;
    }
    bool blocks_uniform = true;
    for (int b = 0; b < p.batch; ++b)
        blocks_uniform = blocks_uniform && nsab[b].x == nsab[0].x && nsab[b].y == nsab[0].y;
    if (p.checkpoint && !blocks_uniform)
        return set_err(nullptr, SCFT_E_UNSUPPORTED, "checkpointing needs the same NsA and NsB for every instance");
    (void)ns_uniform;
    const int n[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a)
        if (n[a] > 1 && (n[a] & 1))
            return set_err(nullptr, SCFT_E_UNSUPPORTED,
                           "odd grid length " + std::to_string(n[a]) + ": the reference's wavenumber table "
                           "(init_cuda.cu:408-418) is not even in k for odd lengths, so its D2Z/Z2D step is "
                           "not a real operator there; use even lengths (or 1)");
    for (int a = 0; a < 3; ++a)
        if (n[a] > 1 && !pass_supported(n[a]))
            return set_err(nullptr, SCFT_E_UNSUPPORTED,
                           "grid length " + std::to_string(n[a]) + " has no kernel (supported: 1 and every even "
                           "length up to " + std::to_string(GEN_MAX) + " whose prime factors are 2, 3 and 5)");
    if (p.nx == 1 && p.ny == 1 && p.nz == 1) return set_err(nullptr, SCFT_E_INVALID, "grid must have a length > 1");

    std::unique_ptr<scft_ctx> c(new scft_ctx());
    c->p = p;
    c->device = device;
    c->NsA = NsA;
    c->NsB = NsB;
    c->Ns = NsA + NsB;
    c->Ns_max = Ns_max;
    c->blocks_uniform = blocks_uniform;
    c->N = (long)p.nx * p.ny * p.nz;
    c->batch = p.batch;
    c->fp32 = p.precision != SCFT_FP64;
    c->mixed = p.precision == SCFT_FP32X;
    c->csize = c->fp32 ? 8 : 16;
    c->tsize = c->fp32 ? 4 : 8;
    // the pass kernels address the four W arrays ([4][batch][N]) with 32-bit byte offsets
    if ((double)4 * c->tsize * c->batch * c->N >= 4294967296.0)
        return set_err(nullptr, SCFT_E_UNSUPPORTED,
                       "batch * N too large for one context (4 * batch * N * sizeof(real) must stay below 2^32 "
                       "bytes): split the batch over several contexts");
    for (int a = 0; a < 3; ++a) {
        c->n[a] = n[a];
        c->active[a] = n[a] > 1;
        if (c->active[a]) {
            if (c->first_axis < 0) c->first_axis = a;
            c->last_axis = a;
        }
    }
    {
        if (p.xy_plane < -1 || p.xy_plane > PLANE_FORMS)
            return set_err(nullptr, SCFT_E_INVALID, "xy_plane must be -1.." + std::to_string(PLANE_FORMS));
        if (!(p.anderson_reg >= 0.0)) return set_err(nullptr, SCFT_E_INVALID, "anderson_reg must be >= 0");
        if (p.xy_plane > 0 && !plane_fits(c.get(), p.nx, p.ny, p.xy_plane))
            return set_err(nullptr, SCFT_E_UNSUPPORTED, "xy_plane needs nx == ny with a plane kernel that fits LDS "
                                                        "(16, 32, 48, 64; 128 except the FP64 xy-plane kernel)");
        c->use_plane = p.xy_plane > 0 ? p.xy_plane : 0;  // auto: decided by scft_autotune
    }
    if (p.tiling < -1 || p.tiling >= PASS_VARIANTS)
        return set_err(nullptr, SCFT_E_INVALID, "tiling must be -1 or 0.." + std::to_string(PASS_VARIANTS - 1));
    for (int a = 0; a < 3; ++a) {  // the three-level tiling only where the length has one
        c->pass_variant[a] = p.tiling < 0 ? 0 : p.tiling;
        if (c->pass_variant[a] == 2 && !pass3_supported(n[a])) c->pass_variant[a] = 0;
    }
    // blocks of the point-wise and reduction kernels: two points per thread,
    // at most 1024 blocks (eight points per thread at 128^3: fewer block
    // partials for the Anderson Gram row and F, 2048 blocks measured 1.4 %
    // slower per 128^3 iteration)
    c->nblk = (int)std::min<long>(1024, std::max<long>(1, (c->N + RT * 2 - 1) / (RT * 2)));

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, SCFT_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return set_err(nullptr, SCFT_E_INVALID, "device index out of range");
    {
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
        if (stream) {
            c->stream = stream;
            c->own_stream = false;
        } else {
            e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
            if (e != hipSuccess)
                return set_err(nullptr, SCFT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        e = hipEventCreate(&c->ev0);
        if (e == hipSuccess) e = hipEventCreate(&c->ev1);
        if (e != hipSuccess) {
            free_all(c.get());
            return set_err(nullptr, SCFT_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    const size_t BN = (size_t)c->batch * c->N;
    int rc = SCFT_OK;
    auto A = [&](void **ptr, size_t bytes) {
        if (rc == SCFT_OK) rc = dalloc(c.get(), ptr, bytes);
    };
    A((void **)&c->wA, 8 * BN);
    A((void **)&c->wB, 8 * BN);
    A((void **)&c->phiA, 8 * BN);
    A((void **)&c->phiB, 8 * BN);
    A(&c->wdz, 4 * c->tsize * BN);
    c->ckpt = p.checkpoint ? 1 : 0;
    const long nslots = slot_count(c->Ns_max, c->ckpt, &c->ck_K, &c->ck_n);
    A((void **)&c->nsab, sizeof(int2) * c->batch);
    A((void **)&c->chiN_b, sizeof(double) * c->batch);
    A(&c->slots, (size_t)nslots * c->csize * BN);
    A(&c->work, c->csize * BN);
    A((void **)&c->part, sizeof(double) * 2 * c->batch * c->nblk);
    A((void **)&c->part2, sizeof(double) * 2 * c->batch * c->nblk);
    A((void **)&c->fpart, sizeof(double) * 3 * c->batch * c->nblk);
    A((void **)&c->Q, sizeof(double) * c->batch);
    A((void **)&c->fe, sizeof(double) * 5 * c->batch);
    A((void **)&c->ffl, sizeof(double) * c->batch);
    if (p.mixing == SCFT_MIX_ANDERSON) {
        const int m1 = p.anderson_m + 1;
        c->and_maxm = p.anderson_m <= 8 ? 8 : (p.anderson_m <= 16 ? 16 : 32);
        A((void **)&c->and_cnt, sizeof(long));
        A((void **)&c->wh, sizeof(double) * m1 * 2 * BN);
        A((void **)&c->dh, sizeof(double) * m1 * 2 * BN);
        A((void **)&c->G, sizeof(double) * c->batch * m1 * m1);
        A((void **)&c->coef, sizeof(double) * c->batch * AND_MAX);
        A((void **)&c->aerr, sizeof(double) * c->batch);
        A((void **)&c->apart, sizeof(double) * c->batch * (c->and_maxm + 2) * c->nblk);
    }
    if (rc != SCFT_OK) {
        std::string m = c->err;
        free_all(c.get());
        return set_err(nullptr, rc, m);
    }
    if ((rc = upload_tables(c.get(), inst_L)) != SCFT_OK) {
        std::string m = c->err;
        free_all(c.get());
        return set_err(nullptr, rc, m);
    }
    {
        hipError_t e = hipMemcpy(c->nsab, nsab.data(), sizeof(int2) * c->batch, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->chiN_b, chis.data(), sizeof(double) * c->batch, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            free_all(c.get());
            return set_err(nullptr, SCFT_E_HIP, std::string("parameter upload: ") + hipGetErrorString(e));
        }
        c->nsab_h = nsab;
    }
    {
        hipError_t e = hipMemsetAsync(c->wA, 0, 8 * BN, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->wB, 0, 8 * BN, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->phiA, 0, 8 * BN, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->phiB, 0, 8 * BN, c->stream);
        if (e == hipSuccess && c->and_cnt) e = hipMemsetAsync(c->and_cnt, 0, sizeof(long), c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            free_all(c.get());
            return set_err(nullptr, SCFT_E_HIP, std::string("initial clear: ") + hipGetErrorString(e));
        }
    }
    *out = c.release();
    return SCFT_OK;
}

extern "C" int scft_create(const scft_params *p, int device, scft_ctx **out)
{
    return create_ctx(p, nullptr, nullptr, nullptr, device, nullptr, out);
}

extern "C" int scft_sizes(const scft_ctx *ctx, long *N, int *NsA, int *NsB, long *device_bytes)
{
    if (!ctx) return SCFT_E_INVALID;
    if (N) *N = ctx->N;
    if (NsA) *NsA = ctx->NsA;
    if (NsB) *NsB = ctx->NsB;
    if (device_bytes) *device_bytes = ctx->device_bytes;
    return SCFT_OK;
}

// Copy kernel of the bandwidth probe: four 16-byte vectors per thread and
// pass, nontemporal loads and stores (measured on MI355X: 6.1-6.2 TB/s for
// 1 GiB, against 5.2 for the runtime's device-to-device copy).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy_probe(const f32x4 *__restrict__ a, f32x4 *__restrict__ b, long n)
{
    constexpr int U = 4;
    const long stride = (long)gridDim.x * 256 * U;
    for (long i = (long)blockIdx.x * 256 * U + threadIdx.x; i < n; i += stride) {
        f32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i + 256L * u < n) v[u] = __builtin_nontemporal_load(a + i + 256L * u);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (i + 256L * u < n) __builtin_nontemporal_store(v[u], b + i + 256L * u);
    }
}

extern "C" int scft_copy_bandwidth(int device, size_t bytes, int reps, double *gbps)
{
    if (!gbps || reps < 1 || bytes < 16 || bytes % 16 != 0)
        return set_err(nullptr, SCFT_E_INVALID, "scft_copy_bandwidth: bytes must be a positive multiple of 16, reps >= 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return set_err(nullptr, SCFT_E_INVALID, "scft_copy_bandwidth: no such device");
    if (hipSetDevice(device) != hipSuccess) return set_err(nullptr, SCFT_E_HIP, "scft_copy_bandwidth: hipSetDevice");
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const long n = (long)(bytes / 16);
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    double best = 0.0;
    auto copy = [&]() {
        hipLaunchKernelGGL(k_copy_probe, dim3(16384), dim3(256), 0, nullptr, (const f32x4 *)a, (f32x4 *)b, n);
        return hipGetLastError();
    };
    if (e == hipSuccess) e = copy();  // untimed: code loading, first touch
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        e = hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = copy();
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.0f) best = std::max(best, 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9);
    }
    const hipError_t es = hipDeviceSynchronize();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("scft_copy_bandwidth: ") + hipGetErrorString(e));
    *gbps = best;
    return SCFT_OK;
}

extern "C" int scft_set_fields(scft_ctx *ctx, const double *wA, const double *wB)
{
    if (!ctx) return SCFT_E_INVALID;
    if (!wA || !wB) return set_err(ctx, SCFT_E_INVALID, "scft_set_fields: NULL field");
    const size_t BN = (size_t)ctx->batch * ctx->N;
    for (size_t i = 0; i < BN; ++i)
        if (!std::isfinite(wA[i]) || !std::isfinite(wB[i]))
            return set_err(ctx, SCFT_E_NONFINITE, "scft_set_fields: non-finite value at index " + std::to_string(i));
    HIPX(ctx, hipSetDevice(ctx->device));
    HIPX(ctx, hipMemcpyAsync(ctx->wA, wA, 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    HIPX(ctx, hipMemcpyAsync(ctx->wB, wB, 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_conc = 0;
    ctx->have_fe = 0;
    if (ctx->and_cnt) {  // a new field restarts the Anderson history
        const long zero = 0;
        HIPX(ctx, hipMemcpy(ctx->and_cnt, &zero, sizeof(long), hipMemcpyHostToDevice));
    }
    return SCFT_OK;
}

extern "C" int scft_get_fields(scft_ctx *ctx, double *phiA, double *phiB, double *wA, double *wB)
{
    if (!ctx) return SCFT_E_INVALID;
    const size_t BN = (size_t)ctx->batch * ctx->N;
    HIPX(ctx, hipSetDevice(ctx->device));
    if (phiA) HIPX(ctx, hipMemcpyAsync(phiA, ctx->phiA, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (phiB) HIPX(ctx, hipMemcpyAsync(phiB, ctx->phiB, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (wA) HIPX(ctx, hipMemcpyAsync(wA, ctx->wA, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (wB) HIPX(ctx, hipMemcpyAsync(wB, ctx->wB, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    return SCFT_OK;
}

// ---------------------------------------------------------------------------
// enqueue helpers (no host synchronisation)
// ---------------------------------------------------------------------------
static unsigned pt_grid(long n) { return (unsigned)std::min<long>(4096, std::max<long>(1, (n + RT - 1) / RT)); }

// points per thread of the contour kernels: 16 bytes of each plane of a slice
// (4 FP32 or 2 FP64 values) when N allows
template <typename T>
static int contour_points(long N)
{
    const int p = 16 / (int)sizeof(T);
    if (N % p == 0) return p;
    return N % 2 == 0 ? 2 : 1;
}

// Cost of a checkpoint evaluation beyond the march for spacing K
// (enqueue_march_contour_ck): J = (Ns/2)/K segments below the last one, each
// re-marching K - 1 steps from its checkpoint and adding two march chunks
// (one more launch each, about half a step): J (K - 1) + J = J K.
static long ck_remarch_steps(int Ns, int K)
{
    return (long)((Ns / 2) / K) * K;
}

// Slice store.  Full mode: slot t holds z(t), t = 0..Ns.  Checkpoint mode
// (sqrt(Ns) checkpointing): only z(kK) is kept (checkpoint k), plus 2K + 4
// working slots.  Slot indices:
//   full:  z(t) = t;                     ZB = Ns+1, Ns+2
//   ckpt:  CK[k] = k (k < nck); working slots nck .. nck + 2K + 3, used as
//          the march-and-contour's window, segment buffers, carry and
//          ping-pong pairs (enqueue_march_contour_ck), then by the adjoint's
//          pair walk as D[i] = nck + i, A[i] = nck + K + 2 + i (i = 1..K-1)
//          and ZB[j] = nck + 2K + 2 + j; then the march's opened states
//          O(kK) at nck + 2K + 4 + k (k < nck, kK < Ns) and the re-marches'
//          work field at nck + 2K + 4 + nck
// The spacing K is the one, within [K0/2, 2 K0] for K0 = ceil(sqrt(Ns+1)),
// of the lowest ck_remarch_steps with a store no larger than K0's (Ns = 200:
// K = 15, 84 re-marched steps; Ns = 100: K = 11, 40).
static long slot_count(int Ns, int ckpt, int *K_out, int *nck_out)
{
    int K0 = 1;
    while (K0 * K0 < Ns + 1) ++K0;
    auto slices = [Ns](int k) { return (long)(Ns / k + 1) + 2 * k + 4; };
    int K = K0;
    long best = ck_remarch_steps(Ns, K0);
    for (int k = std::max(2, K0 / 2); k <= 2 * K0; ++k) {
        if (slices(k) > slices(K0)) continue;
        const long r = ck_remarch_steps(Ns, k);
        if (r < best || (r == best && slices(k) < slices(K))) {
            best = r;
            K = k;
        }
    }
    const int nck = Ns / K + 1;
    if (K_out) *K_out = K;
    if (nck_out) *nck_out = nck;
    // + the march's opened states O(kK), k < nck, and the re-marches' work field
    return ckpt ? (long)nck + 2 * K + 4 + nck + 1 : (long)Ns + 3;
}

// block of the real (br) and imaginary (bi) part at packed step t (1..Ns):
// q(t) steps with block A for t <= NsA; qdag(Ns - t) with block B for t <= NsB
static inline void step_blocks(const scft_ctx *c, int t, int &br, int &bi)
{
    br = (t <= c->NsA) ? 0 : 1;
    bi = (t <= c->NsB) ? 1 : 0;
}

// one packed step dst = S(src): three axis passes (or the fused xy-plane
// kernel and the z pass).  t > 0: contour step t of every instance's packed
// chain (the kernels pick each instance's blocks, and leave instances whose
// chain is shorter than t alone); t == 0: the real part uses block br's W and
// the imaginary part block bi's for every instance.
// arguments of one pass along `axis` (3: the xy-plane kernel, x tables)
template <typename T>
static PassArgs<T> pass_args(scft_ctx *c, int axis, const typename cplx<T>::type *src, typename cplx<T>::type *dst,
                             int t, int br, int bi, bool pre, bool post)
{
    using C = typename cplx<T>::type;
    const size_t BN = (size_t)c->batch * c->N;
    const int a = axis == 3 ? 0 : axis;
    PassArgs<T> pa;
    pa.src = src;
    pa.dst = dst;
    pa.W = static_cast<const T *>(c->wdz);  // [pre A, pre B, post A, post B]
    pa.wo[0] = (unsigned)(br * BN);
    pa.wo[1] = (unsigned)(bi * BN);
    pa.wo[2] = (unsigned)((2 + br) * BN);
    pa.wo[3] = (unsigned)((2 + bi) * BN);
    pa.slot = nullptr;
    pa.wn[0] = pa.wn[1] = 0u;
    pa.post_scale = (T)(1.0 / (double)c->N);
    pa.post_from_pre = c->post_exact;
    pa.nsab = (const int2 *)c->nsab;
    pa.step = t;
    pa.BN = (unsigned)BN;
    pa.tw = (const C *)c->tw[a];
    pa.kf = (const typename kfac<T>::type *)c->kf[a];
    pa.kf_stride = c->kf_per_inst ? (unsigned)c->n[a] : 0u;
    pa.nlines = (unsigned)(c->batch * (c->N / c->n[a]));
    pa.N = (unsigned)c->N;
    pa.nx = c->n[0];
    pa.ny = c->n[1];
    pa.nz = c->n[2];
    pa.pre = pre;
    pa.post = post;
    return pa;
}

// launch one unit of a step: an axis pass (0..2) or the xy-plane kernel (3);
// `fused`: the step-boundary form
template <typename T>
static hipError_t launch_unit(scft_ctx *c, int unit, const PassArgs<T> &pa, bool fused)
{
    using C = typename cplx<T>::type;
    if (unit == 3) {
        const auto *kfy = (const typename kfac<T>::type *)c->kf[1];
        const long np = (long)c->batch * c->n[2];
        const C *twy = (const C *)c->tw[1];
        return run_plane<T>(c, c->use_plane, pa, twy, kfy, np, fused);
    }
    return run_pass<T>(c, unit, c->n[unit], c->pass_variant[unit], pa, fused);
}

// the units of a step in forward order (the xy-plane kernel counts as one)
static int step_units(const scft_ctx *c, int *units)
{
    int nu = 0;
    if (c->use_plane) {
        units[nu++] = 3;
        if (c->active[2]) units[nu++] = 2;
    } else {
        for (int a = 0; a < 3; ++a)
            if (c->active[a]) units[nu++] = a;
    }
    return nu;
}

// one packed step dst = S(src): three axis passes (or the fused xy-plane
// kernel and the z pass).  t > 0: contour step t of every instance's packed
// chain (the kernels pick each instance's blocks, and leave instances whose
// chain is shorter than t alone); t == 0: the real part uses block br's W and
// the imaginary part block bi's for every instance.
template <typename T>
static hipError_t enqueue_step(scft_ctx *c, const typename cplx<T>::type *src, typename cplx<T>::type *dst, int t,
                               int br, int bi)
{
    using C = typename cplx<T>::type;
    C *work = (C *)c->work;
    int units[3];
    const int nu = step_units(c, units);
    for (int k = 0; k < nu; ++k) {
        const bool first = k == 0, last = k == nu - 1;
        hipError_t e = launch_unit<T>(
            c, units[k], pass_args<T>(c, units[k], first ? src : work, last ? dst : work, t, br, bi, first, last),
            false);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
static typename cplx<T>::type *slot_ptr(scft_ctx *c, long k)
{
    return (typename cplx<T>::type *)c->slots + (size_t)k * c->batch * c->N;
}

// destination of z(t) in the slice store (checkpoint mode keeps z(kK) and
// two scratch slices)
template <typename T>
static typename cplx<T>::type *march_slot(scft_ctx *c, int t)
{
    if (!c->ckpt) return slot_ptr<T>(c, t);
    if (t % c->ck_K == 0) return slot_ptr<T>(c, t / c->ck_K);
    return slot_ptr<T>(c, c->ck_n + c->ck_K + (t & 1));
}

// The opened state O(t) of the step-boundary march is the field a boundary
// launch leaves after it closes step t: z(t) times step t+1's pre factor with
// step t+1's first unit applied (O(0): the first launch's output).  A chunk
// can start from O(t0) instead of z(t0) (src_open) and skip the plain launch
// that would recompute it, and can end with the boundary launch that closes
// its last step (end_open, its O(t_end) left in wk) instead of a plain one; a
// march started from a saved O(t0) then reproduces the uninterrupted march
// bit for bit.  open_of(t) (t0 <= t, may return nullptr) names a buffer where
// O(t) is kept (the next launch reads it there); wk is the in-place work
// buffer of the chunk.
template <typename T, typename DstOf, typename OpenOf>
static hipError_t enqueue_march_fused_ex(scft_ctx *c, const typename cplx<T>::type *src0, bool src_open, int t0,
                                         int nsteps, DstOf dst_of, OpenOf open_of, typename cplx<T>::type *wk,
                                         bool end_open, typename cplx<T>::type **last)
{
    using C = typename cplx<T>::type;
    int units[3];
    const int nu = step_units(c, units);
    C *dst = nullptr;
    const C *cur = src0;  // the field the next launch reads
    const int t_end = t0 + nsteps;
    for (int t = t0 + 1; t <= t_end; ++t) {
        dst = dst_of(t);
        for (int k = 0; k < nu; ++k) {
            const int u = (t & 1) ? units[k] : units[nu - 1 - k];
            const bool first = k == 0, lastu = k == nu - 1;
            if (first && (t > t0 + 1 || src_open)) continue;  // opened by the launch that closed t-1
            hipError_t e;
            if (lastu && (t < t_end || end_open)) {
                C *o = open_of(t);
                C *out = o ? o : wk;
                PassArgs<T> pa = pass_args<T>(c, u, cur, out, t, 0, 0, first, false);
                pa.slot = dst;
                e = launch_unit<T>(c, u, pa, true);
                cur = out;
            } else {
                C *o = (first && t == t0 + 1) ? open_of(t0) : nullptr;  // O(t0) of a chunk opened from z(t0)
                C *out = lastu ? dst : (o ? o : wk);
                e = launch_unit<T>(c, u, pass_args<T>(c, u, cur, out, t, 0, 0, first, lastu), false);
                cur = out;
            }
            if (e != hipSuccess) return e;
        }
    }
    if (last) *last = dst;
    return hipSuccess;
}

// The packed march with step-boundary kernels.  S = W C_x C_y C_z W and the
// axis convolutions commute, so odd steps run the units in forward order and
// even steps in reverse: the last unit of step t is the first of step t+1, and
// one kernel closes t (z(t) to its slot) and opens t+1 on the data in
// registers.  Steps then cost nu - 1 launches instead of nu (nu >= 2 units).
// z(t0) = src marched to z(t0 + nsteps); z(t) goes to dst_of(t).  The order
// follows t itself, not t - t0, so a march split into chunks, or re-marched
// from a checkpoint, runs its steps' units in the order the march from z(0)
// does.  A chunk started from a slice has one launch the march from z(0) does
// not: a plain kernel runs the first unit where that march has a boundary
// kernel (an xy-plane boundary kernel opens its step y then x, a plain plane
// kernel runs x then y; FP32X rounds the carried slice to FP32), so its
// slices can differ in the last bits; enqueue_march_fused_ex started from a
// saved opened state has no such launch.
template <typename T, typename DstOf>
static hipError_t enqueue_march_fused_from(scft_ctx *c, const typename cplx<T>::type *src0, int t0, int nsteps,
                                           DstOf dst_of, typename cplx<T>::type **last)
{
    using C = typename cplx<T>::type;
    return enqueue_march_fused_ex<T>(c, src0, false, t0, nsteps, dst_of, [](int) -> C * { return nullptr; },
                                     (C *)c->work, false, last);
}

template <typename T>
static hipError_t enqueue_march_fused(scft_ctx *c, int nsteps, typename cplx<T>::type **last)
{
    return enqueue_march_fused_from<T>(c, slot_ptr<T>(c, 0), 0, nsteps, [c](int t) { return march_slot<T>(c, t); },
                                       last);
}

// march `nsteps` steps from slot 0.  block -1: the packed chain; 0/1: one
// block's W on both parts (sovDifFft's march, full mode only).  In checkpoint
// mode only z(kK) and the last slice are kept; *last receives z(nsteps).
template <typename T>
static hipError_t enqueue_march(scft_ctx *c, int nsteps, int block, typename cplx<T>::type **last = nullptr)
{
    using C = typename cplx<T>::type;
    int units[3];
    if (block < 0 && c->p.fuse_steps && nsteps > 1 && step_units(c, units) >= 2)
        return enqueue_march_fused<T>(c, nsteps, last);
    C *cur = slot_ptr<T>(c, 0);
    for (int t = 1; t <= nsteps; ++t) {
        C *dst = march_slot<T>(c, t);
        hipError_t e = block < 0 ? enqueue_step<T>(c, cur, dst, t, 0, 0) : enqueue_step<T>(c, cur, dst, 0, block, block);
        if (e != hipSuccess) return e;
        cur = dst;
    }
    if (last) *last = cur;
    return hipSuccess;
}

// z(t0) = src marched n steps of the packed chain, z(t) to dst_of(t): the
// step-boundary march where it applies (for n = 1 too, so that the step runs
// its units in the order of t), else one step at a time
template <typename T, typename DstOf>
static hipError_t enqueue_march_from(scft_ctx *c, const typename cplx<T>::type *src, int t0, int n, DstOf dst_of)
{
    using C = typename cplx<T>::type;
    int units[3];
    if (n <= 0) return hipSuccess;
    if (c->p.fuse_steps && step_units(c, units) >= 2)
        return enqueue_march_fused_from<T>(c, src, t0, n, dst_of, nullptr);
    const C *cur = src;
    for (int t = t0 + 1; t <= t0 + n; ++t) {
        C *dst = dst_of(t);
        hipError_t e = enqueue_step<T>(c, cur, dst, t, 0, 0);
        if (e != hipSuccess) return e;
        cur = dst;
    }
    return hipSuccess;
}

// Visits tau = 0, 1, ... and yields z(tau) (ascending) and z(Ns - tau)
// (descending).  Full mode reads the store.  Checkpoint mode keeps two
// K-slice segment buffers: when tau reaches checkpoint j the ascending one is
// re-marched to z(jK + 1 .. jK + K - 1), and when Ns - tau enters segment k
// (s = kK .. kK + K - 1) the descending one is re-marched from checkpoint k --
// about 2 Ns extra steps for a full traversal, each segment one step-boundary
// march, with O(sqrt Ns) slices held.  With step-boundary kernels a segment
// is re-marched from the forward march's opened state O(kK), kept beside
// checkpoint k, and ends open in the work field (a plain last launch where it
// reaches Ns, as the forward march's does), so its slices equal the forward
// march's bit for bit (from z(kK) a plain first launch would round
// differently).
template <typename T>
struct PairWalk {
    using C = typename cplx<T>::type;
    scft_ctx *c;
    int seg = -1;
    explicit PairWalk(scft_ctx *cc) : c(cc) {}

    template <typename DstOf>
    hipError_t remarch(int k, int len, DstOf dst_of)
    {
        const int K = c->ck_K, nck = c->ck_n;
        int units[3];
        if (len <= 0) return hipSuccess;
        if (c->p.fuse_steps && step_units(c, units) >= 2)
            return enqueue_march_fused_ex<T>(c, slot_ptr<T>(c, (long)nck + 2L * K + 4 + k), true, k * K, len, dst_of,
                                             [](int) -> C * { return nullptr; }, (C *)c->work, k * K + len < c->Ns,
                                             nullptr);
        return enqueue_march_from<T>(c, slot_ptr<T>(c, k), k * K, len, dst_of);
    }

    hipError_t next(int tau, const C **zt, const C **zu)
    {
        const int Ns = c->Ns, K = c->ck_K, nck = c->ck_n;
        if (!c->ckpt) {
            *zt = slot_ptr<T>(c, tau);
            *zu = slot_ptr<T>(c, Ns - tau);
            return hipSuccess;
        }
        // ascending: A[i] = slot nck + K + 2 + i holds z(jK + i), i = 1..K-1
        const long a0 = nck + K + 2;
        const int j = tau / K, ia = tau - j * K;
        if (ia == 0) {
            const int len = std::min(K - 1, Ns - j * K);
            hipError_t e = remarch(j, len, [&](int t) { return slot_ptr<T>(c, a0 + (t - j * K)); });
            if (e != hipSuccess) return e;
        }
        *zt = (ia == 0) ? slot_ptr<T>(c, j) : slot_ptr<T>(c, a0 + ia);
        // descending: D[i] = slot nck + i holds z(kK + i)
        const int s = Ns - tau, k = s / K;
        if (k != seg) {
            const int len = std::min(K - 1, Ns - k * K);
            hipError_t e = remarch(k, len, [&](int t) { return slot_ptr<T>(c, nck + (t - k * K)); });
            if (e != hipSuccess) return e;
            seg = k;
        }
        const int i = s - k * K;
        *zu = (i == 0) ? slot_ptr<T>(c, k) : slot_ptr<T>(c, nck + i);
        return hipSuccess;
    }
};

// Checkpoint-mode march with the contour summed on the way.  Pair
// (tau, Ns - tau), tau <= Ns/2, needs z(tau) from the first half of the march
// and z(Ns - tau) from the second.  The march runs once, keeping every
// checkpoint z(kK), and the pairs are taken in segments of K (tau = jK ..
// jK + K - 1) from the last one, j = J = (Ns/2)/K, down:
//   * segment J's slices z(JK .. Ns - JK) all lie on the march's own path
//     (at most 2K of them): the first march chunk, z(0) -> z(Ns - JK), keeps
//     them in a window;
//   * each earlier segment j takes z(Ns - tau) from the next K steps of the
//     march and z(tau) from a re-march of K - 1 steps from checkpoint j.
// That is J (K - 1), about Ns/2, re-marched steps per evaluation.  Q comes
// from z(Ns) as soon as the march reaches it, before the last segment's
// pairs, whose launch applies the ds/Q factor to the summed phi.
// With step-boundary kernels the march keeps its opened state in the work
// field between its chunks (the re-marches have a work field of their own)
// and keeps O(kK) beside every checkpoint: each re-march (here and in the
// adjoint's pair walk) starts from O(kK) and ends open, so no chunk has a
// plain first or last launch and every re-marched slice equals the march's
// own bit for bit.
// Slots: checkpoint k at k; window z(JK + 1 + i) and the re-marched z(jK + i)
// at nck + 1 + i and nck + i; the march's z(Ns - jK - K + 1 + i) at nck + K + i;
// each chunk's last slice in the carry pair nck + 2K, nck + 2K + 1; the steps
// before the window in the ping-pong pair nck + 2K + 2, nck + 2K + 3; O(kK)
// at nck + 2K + 4 + k and the re-marches' work field at nck + 2K + 4 + nck.
template <typename T>
static hipError_t enqueue_march_contour_ck(scft_ctx *c, double *phiA, double *phiB)
{
    using C = typename cplx<T>::type;
    const int Ns = c->Ns, K = c->ck_K, nck = c->ck_n, half = Ns / 2, J = half / K;
    const long carry0 = (long)nck + 2L * K, pp0 = carry0 + 2, open0 = carry0 + 4;
    const dim3 gc((unsigned)((c->N + RT - 1) / RT), c->batch);
    auto ck = [&](int t) { return slot_ptr<T>(c, t / K); };
    int units[3];
    const bool fused = c->p.fuse_steps && step_units(c, units) >= 2;
    C *const wk1 = (C *)c->work, *const wk2 = slot_ptr<T>(c, open0 + nck);
    auto open_ck = [&](int t) -> C * { return (t % K == 0 && t < Ns) ? slot_ptr<T>(c, open0 + t / K) : nullptr; };
    // where a chunk that ends open at t leaves O(t): its checkpoint's slot, else the work field
    auto open_at = [&](int t) -> C * { return open_ck(t) ? open_ck(t) : wk1; };
    auto none = [](int) -> C * { return nullptr; };
    int launched = 0;
    const int P = contour_points<T>(c->N);  // points per thread (16-byte loads of each plane)
    const dim3 gcp((unsigned)((c->N / P + RT - 1) / RT), c->batch);
    const long BN = (long)c->batch * c->N;
    // pairs tau0..tau1; last: the contour's final launch applies ds/Q (Q already known)
    auto pairs = [&](int tau0, int tau1, auto zt_of, auto zu_of, bool last) -> hipError_t {
        for (int t0 = tau0; t0 <= tau1; t0 += CK_PAIRS) {
            PairSlices<C> ps{};
            const int np = std::min(CK_PAIRS, tau1 - t0 + 1);
            for (int k = 0; k < np; ++k) {
                ps.zt[k] = zt_of(t0 + k);
                ps.zu[k] = zu_of(Ns - t0 - k);
            }
            const double *f = (last && t0 + np > tau1) ? (const double *)c->ffl : nullptr;
            const int init = launched++ == 0 ? 1 : 0;
            if (P == 4)
                hipLaunchKernelGGL((k_contour_pairs<T, 4>), gcp, dim3(RT), 0, c->stream, ps, t0, np, c->NsA, Ns, c->N,
                                   BN, phiA, phiB, init, f);
            else if (P == 2)
                hipLaunchKernelGGL((k_contour_pairs<T, 2>), gcp, dim3(RT), 0, c->stream, ps, t0, np, c->NsA, Ns, c->N,
                                   BN, phiA, phiB, init, f);
            else
                hipLaunchKernelGGL((k_contour_pairs<T, 1>), gcp, dim3(RT), 0, c->stream, ps, t0, np, c->NsA, Ns, c->N,
                                   BN, phiA, phiB, init, f);
        }
        return hipGetLastError();
    };
    // Q from z(Ns), once the march has reached it
    auto q_of = [&](const C *zNs) -> hipError_t {
        const dim3 g(c->nblk, c->batch);
        hipLaunchKernelGGL(k_qsum<T>, g, dim3(RT), 0, c->stream, zNs, (const C *)c->slots, (const int2 *)nullptr, c->N,
                           c->batch, c->nblk, c->part2);
        hipLaunchKernelGGL(k_q_final, dim3(c->batch), dim3(RT), 0, c->stream, c->part2, c->nblk, c->N, c->p.ds, c->Q,
                           c->ffl);
        return hipGetLastError();
    };
    // segment J: z(0) -> z(Ns - JK) in one chunk
    const int wlo = J * K, whi = Ns - J * K;
    auto first_dst = [&](int t) -> C * {
        if (t % K == 0) return ck(t);
        if (t == whi) return slot_ptr<T>(c, carry0);
        if (t >= wlo) return slot_ptr<T>(c, nck + 1 + (t - wlo));
        return slot_ptr<T>(c, pp0 + (t & 1));
    };
    hipError_t e = fused ? enqueue_march_fused_ex<T>(c, slot_ptr<T>(c, 0), false, 0, whi, first_dst, open_ck, wk1,
                                                     J > 0, nullptr)
                         : enqueue_march_from<T>(c, slot_ptr<T>(c, 0), 0, whi, first_dst);
    if (e != hipSuccess) return e;
    auto first_at = [&](int t) -> const C * { return first_dst(t); };
    if (J == 0 && (e = q_of(first_at(Ns))) != hipSuccess) return e;  // the first chunk reached z(Ns)
    if ((e = pairs(wlo, half, first_at, first_at, J == 0)) != hipSuccess) return e;
    int tcur = whi;
    C *ocur = open_at(whi);  // O(tcur), with step-boundary kernels
    long carry = carry0;
    for (int j = J - 1; j >= 0; --j) {
        const int tau0 = j * K, tau1 = tau0 + K - 1, dlo = Ns - tau1, dhi = Ns - tau0;
        const long next = carry == carry0 ? carry0 + 1 : carry0;
        // z(dlo .. dhi): the march's next K steps (tcur = dlo - 1)
        auto ddst = [&](int t) -> C * {
            if (t % K == 0) return ck(t);
            if (t == dhi) return slot_ptr<T>(c, next);
            return slot_ptr<T>(c, nck + K + (t - dlo));
        };
        if (fused) {  // on from the march's opened state O(tcur)
            e = enqueue_march_fused_ex<T>(c, ocur, true, tcur, dhi - tcur, ddst, open_ck, wk1, j > 0, nullptr);
            ocur = open_at(dhi);
        } else {
            const C *src = (tcur % K == 0) ? ck(tcur) : slot_ptr<T>(c, carry);
            e = enqueue_march_from<T>(c, src, tcur, dhi - tcur, ddst);
        }
        if (e != hipSuccess) return e;
        if (j == 0 && (e = q_of(ddst(Ns))) != hipSuccess) return e;  // dhi == Ns
        // z(tau0 + 1 .. tau1) re-marched from checkpoint j
        auto adst = [&](int t) -> C * { return slot_ptr<T>(c, nck + (t - tau0)); };
        e = fused ? enqueue_march_fused_ex<T>(c, open_ck(tau0), true, tau0, K - 1, adst, none, wk2, true, nullptr)
                  : enqueue_march_from<T>(c, ck(tau0), tau0, K - 1, adst);
        if (e != hipSuccess) return e;
        auto at = [&](int t) -> const C * { return t == tau0 ? ck(t) : adst(t); };
        auto dt = [&](int t) -> const C * { return ddst(t); };
        if ((e = pairs(tau0, tau1, at, dt, j == 0)) != hipSuccess) return e;
        tcur = dhi;
        carry = next;
    }
    return hipGetLastError();
}

// The trapezoid contour over the full slice store (k_contour): P consecutive
// points per thread, so that each plane of a slice moves in 16-byte loads (4
// points in FP32, 2 in FP64, fewer when N is not a multiple); 4 pairs of
// slices in flight per thread.  It also finishes Q from k_qsum's partials
// (part2) into Qout and ds/Q into fflout.
template <typename T>
static hipError_t launch_contour(scft_ctx *c, double *phiA, double *phiB, double *Qout, double *fflout)
{
    using C = typename cplx<T>::type;
    const int P = contour_points<T>(c->N);
    const dim3 gc((unsigned)((c->N / P + RT - 1) / RT), c->batch);
    if (P == 4)
        hipLaunchKernelGGL((k_contour<T, 4>), gc, dim3(RT), 0, c->stream, (const C *)c->slots, c->N, c->batch,
                           (const int2 *)c->nsab, (const double *)c->part2, c->nblk, c->p.ds, Qout, fflout, phiA,
                           phiB);
    else if (P == 2)
        hipLaunchKernelGGL((k_contour<T, 2>), gc, dim3(RT), 0, c->stream, (const C *)c->slots, c->N, c->batch,
                           (const int2 *)c->nsab, (const double *)c->part2, c->nblk, c->p.ds, Qout, fflout, phiA,
                           phiB);
    else
        hipLaunchKernelGGL((k_contour<T, 1>), gc, dim3(RT), 0, c->stream, (const C *)c->slots, c->N, c->batch,
                           (const int2 *)c->nsab, (const double *)c->part2, c->nblk, c->p.ds, Qout, fflout, phiA,
                           phiB);
    return hipGetLastError();
}

// k_prep with the context's W factor form (SCFT_FP32X: pre factors as W - 1)
template <typename T>
static void launch_prep(scft_ctx *c, double *fA, const double *fB, const double *mean_part)
{
    const dim3 g(c->nblk, c->batch);
    if (c->mixed)
        hipLaunchKernelGGL((k_prep<T, true>), g, dim3(RT), 0, c->stream, fA, fB, c->N, c->nblk, mean_part,
                           c->p.ds / 2, (const double *)c->post_b, (T *)c->wdz, slot_ptr<T>(c, 0), c->batch);
    else
        hipLaunchKernelGGL((k_prep<T, false>), g, dim3(RT), 0, c->stream, fA, fB, c->N, c->nblk, mean_part,
                           c->p.ds / 2, (const double *)c->post_b, (T *)c->wdz, slot_ptr<T>(c, 0), c->batch);
}

// getConc on the device for fields (fA, fB); mean_part: partial sums of fA to
// subtract first (or null: fields used as given)
template <typename T>
static hipError_t enqueue_conc_fields(scft_ctx *c, double *fA, const double *fB, const double *mean_part,
                                      double *phiA, double *phiB)
{
    using C = typename cplx<T>::type;
    const dim3 g(c->nblk, c->batch);
    launch_prep<T>(c, fA, fB, mean_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (c->ckpt) return enqueue_march_contour_ck<T>(c, phiA, phiB);  // one Ns for all instances
    C *zNs = nullptr;
    if ((e = enqueue_march<T>(c, c->Ns_max, -1, &zNs)) != hipSuccess) return e;
    // full store: each instance's q(Ns_b) is slice Ns_b
    hipLaunchKernelGGL(k_qsum<T>, g, dim3(RT), 0, c->stream, (const C *)zNs, (const C *)c->slots,
                       (const int2 *)c->nsab, c->N, c->batch, c->nblk, c->part2);
    return launch_contour<T>(c, phiA, phiB, c->Q, c->ffl);  // Q and ds/Q too
}

template <typename T>
static hipError_t enqueue_conc(scft_ctx *c, const double *mean_part)
{
    return enqueue_conc_fields<T>(c, c->wA, c->wB, mean_part, c->phiA, c->phiB);
}

static hipError_t enqueue_conc_any(scft_ctx *c, const double *mean_part)
{
    return c->fp32 ? enqueue_conc<float>(c, mean_part) : enqueue_conc<double>(c, mean_part);
}

static hipError_t enqueue_update(scft_ctx *c)
{
    const dim3 g(c->nblk, c->batch);
    if (c->p.mixing == SCFT_MIX_SIMPLE) {
        hipLaunchKernelGGL(k_mix_simple, g, dim3(RT), 0, c->stream, c->wA, c->wB, c->phiA, c->phiB, c->N, c->nblk,
                           (const double *)c->chiN_b, c->p.wopt, c->p.wcmp, c->fpart);
    } else {
        const int m = c->p.anderson_m;
#define SCFT_AND_RESID(MM)                                                                                     \
    hipLaunchKernelGGL(k_and_resid<MM>, g, dim3(RT), 0, c->stream, c->wA, c->wB, c->phiA, c->phiB, c->N, c->batch, \
                       c->nblk, (const double *)c->chiN_b, c->p.wcmp, c->p.anderson_mix, c->wh, c->dh, c->and_cnt, m,  \
                       c->apart)
        switch (c->and_maxm) {
        case 8: SCFT_AND_RESID(8); break;
        case 16: SCFT_AND_RESID(16); break;
        default: SCFT_AND_RESID(32); break;
        }
#undef SCFT_AND_RESID
        hipLaunchKernelGGL(k_and_solve, dim3(c->batch), dim3(RT), 0, c->stream, c->apart, c->nblk, c->and_maxm, m,
                           c->and_cnt, c->G, c->coef, c->aerr, c->p.anderson_reg);
        hipLaunchKernelGGL(k_and_update, g, dim3(RT), 0, c->stream, c->wA, c->wB, c->phiA, c->phiB, c->N, c->batch,
                           c->nblk, m, c->and_cnt, (const double *)c->chiN_b, c->wh, c->coef, c->fpart);
    }
    hipLaunchKernelGGL(k_fe_final, dim3(c->batch), dim3(RT), 0, c->stream, c->fpart, c->nblk, c->Q, c->N, c->fe,
                       c->p.mixing == SCFT_MIX_ANDERSON ? c->and_cnt : nullptr);
    return hipGetLastError();
}

// one full iteration of Free's loop
static hipError_t enqueue_iteration(scft_ctx *c)
{
    const dim3 g(c->nblk, c->batch);
    hipLaunchKernelGGL(k_sums, g, dim3(RT), 0, c->stream, c->wA, c->wB, c->N, c->nblk, c->part);
    hipLaunchKernelGGL(k_center, g, dim3(RT), 0, c->stream, c->wA, c->wB, c->N, c->nblk, c->part, c->part2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = enqueue_conc_any(c, c->part2)) != hipSuccess) return e;
    return enqueue_update(c);
}

// one iteration: replay of the captured graph (captured on first use), or
// the eager launch sequence when the params ask for no graph
static hipError_t capture_iterations(scft_ctx *c, int n, hipGraph_t *graph, hipGraphExec_t *exec)
{
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t e1 = hipSuccess;
    for (int i = 0; i < n && e1 == hipSuccess; ++i) e1 = enqueue_iteration(c);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(c->stream, &g);
    if (e1 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e1;
    }
    if (e != hipSuccess) return e;
    *graph = g;
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) *exec = nullptr;
    return e;
}

static hipError_t ensure_graph(scft_ctx *c)
{
    if (c->p.use_graph && !c->graph_exec) return capture_iterations(c, 1, &c->graph, &c->graph_exec);
    return hipSuccess;
}

static hipError_t launch_iteration(scft_ctx *c)
{
    if (!c->p.use_graph) return enqueue_iteration(c);
    hipError_t e = ensure_graph(c);
    if (e != hipSuccess) return e;
    return hipGraphLaunch(c->graph_exec, c->stream);
}

// n iterations: the pair graph for each pair, the single one for an odd last
static hipError_t launch_iterations(scft_ctx *c, int n)
{
    if (!c->p.use_graph || n < 2) {
        for (int i = 0; i < n; ++i) {
            hipError_t e = launch_iteration(c);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    if (!c->graph2_exec) {
        hipError_t e = capture_iterations(c, 2, &c->graph2, &c->graph2_exec);
        if (e != hipSuccess) return e;
    }
    int i = 0;
    for (; i + 2 <= n; i += 2) {
        hipError_t e = hipGraphLaunch(c->graph2_exec, c->stream);
        if (e != hipSuccess) return e;
    }
    return i < n ? launch_iteration(c) : hipSuccess;
}

static int check_fe(scft_ctx *c, const std::vector<double> &fe)
{
    for (int b = 0; b < c->batch; ++b)
        if (!std::isfinite(fe[b * 5 + 0]) || !std::isfinite(fe[b * 5 + 3]))
            return set_err(c, SCFT_E_NONFINITE,
                           "non-finite free energy in instance " + std::to_string(b) + " (Q <= 0 or overflow)");
    return SCFT_OK;
}

extern "C" int scft_propagate(scft_ctx *ctx, int block, int sign, const double *q_init, double *q_out)
{
    if (!ctx) return SCFT_E_INVALID;
    if (block != 0 && block != 1) return set_err(ctx, SCFT_E_INVALID, "block must be 0 (A) or 1 (B)");
    if (sign != 1 && sign != -1) return set_err(ctx, SCFT_E_INVALID, "sign must be +1 or -1");
    if (!q_out) return set_err(ctx, SCFT_E_INVALID, "q_out is NULL");
    if (ctx->ckpt) return set_err(ctx, SCFT_E_STATE, "scft_propagate returns every slice: needs checkpoint = 0");
    HIPX(ctx, hipSetDevice(ctx->device));
    const int ns = block == 0 ? ctx->NsA : ctx->NsB;
    const size_t BN = (size_t)ctx->batch * ctx->N;
    const double *w = block == 0 ? ctx->wA : ctx->wB;
    double *stage = (double *)ctx->work;  // BN doubles fit in the complex work buffer
    char *wdz_b = (char *)ctx->wdz + (size_t)block * BN * ctx->tsize;
    char *wpost_b = (char *)ctx->wdz + (size_t)(2 + block) * BN * ctx->tsize;
    if (ctx->mixed)
        hipLaunchKernelGGL((k_wdz_one<float, true>), dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, w, (long)BN,
                           ctx->N, ctx->p.ds / 2, (const double *)ctx->post_b, (float *)wdz_b, (float *)wpost_b);
    else if (ctx->fp32)
        hipLaunchKernelGGL((k_wdz_one<float, false>), dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, w, (long)BN,
                           ctx->N, ctx->p.ds / 2, (const double *)ctx->post_b, (float *)wdz_b, (float *)wpost_b);
    else
        hipLaunchKernelGGL((k_wdz_one<double, false>), dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, w, (long)BN,
                           ctx->N, ctx->p.ds / 2, (const double *)ctx->post_b, (double *)wdz_b, (double *)wpost_b);
    if (q_init) HIPX(ctx, hipMemcpyAsync(stage, q_init, 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    if (ctx->fp32)
        hipLaunchKernelGGL(k_seed_real<float>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (float2 *)ctx->slots,
                           q_init ? stage : nullptr, (long)BN);
    else
        hipLaunchKernelGGL(k_seed_real<double>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (double2 *)ctx->slots,
                           q_init ? stage : nullptr, (long)BN);
    HIPX(ctx, hipGetLastError());
    hipError_t e = ctx->fp32 ? enqueue_march<float>(ctx, ns, block) : enqueue_march<double>(ctx, ns, block);
    HIPX(ctx, e);
    std::vector<double> host(BN);
    for (int t = 0; t <= ns; ++t) {
        const char *sl = (const char *)ctx->slots + (size_t)t * BN * ctx->csize;
        if (ctx->fp32)
            hipLaunchKernelGGL(k_take_real<float>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (const float2 *)sl,
                               stage, (long)BN);
        else
            hipLaunchKernelGGL(k_take_real<double>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream,
                               (const double2 *)sl, stage, (long)BN);
        HIPX(ctx, hipGetLastError());
        HIPX(ctx, hipMemcpyAsync(host.data(), stage, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
        HIPX(ctx, hipStreamSynchronize(ctx->stream));
        const int s = sign == 1 ? t : ns - t;
        for (int b = 0; b < ctx->batch; ++b)
            std::memcpy(q_out + ((size_t)b * (ns + 1) + s) * ctx->N, host.data() + (size_t)b * ctx->N,
                        8 * ctx->N);
    }
    ctx->have_conc = 0;  // the slice store no longer holds the chain of the current fields
    return SCFT_OK;
}

extern "C" int scft_concentrations(scft_ctx *ctx, double *Q)
{
    if (!ctx) return SCFT_E_INVALID;
    HIPX(ctx, hipSetDevice(ctx->device));
    const dim3 g(ctx->nblk, ctx->batch);
    // getConc's average_value(wa) (cuda_scft.cu:71)
    hipLaunchKernelGGL(k_sums, g, dim3(RT), 0, ctx->stream, ctx->wA, (const double *)nullptr, ctx->N, ctx->nblk,
                       ctx->part);
    HIPX(ctx, hipGetLastError());
    HIPX(ctx, enqueue_conc_any(ctx, ctx->part));
    std::vector<double> q(ctx->batch);
    HIPX(ctx, hipMemcpyAsync(q.data(), ctx->Q, 8 * ctx->batch, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < ctx->batch; ++b)
        if (!std::isfinite(q[b]) || !(q[b] > 0.0))
            return set_err(ctx, SCFT_E_NONFINITE, "Q of instance " + std::to_string(b) + " is not finite and > 0");
    if (Q) std::memcpy(Q, q.data(), 8 * ctx->batch);
    ctx->have_conc = 1;
    return SCFT_OK;
}

extern "C" int scft_update_fields(scft_ctx *ctx)
{
    if (!ctx) return SCFT_E_INVALID;
    if (!ctx->have_conc) return set_err(ctx, SCFT_E_STATE, "scft_update_fields needs scft_concentrations first");
    HIPX(ctx, hipSetDevice(ctx->device));
    HIPX(ctx, enqueue_update(ctx));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_conc = 0;
    ctx->have_fe = 1;
    return SCFT_OK;
}

extern "C" int scft_free_energy(scft_ctx *ctx, double *out)
{
    if (!ctx || !out) return SCFT_E_INVALID;
    if (!ctx->have_fe) return set_err(ctx, SCFT_E_STATE, "no free energy yet: run an update or an iteration");
    HIPX(ctx, hipSetDevice(ctx->device));
    std::vector<double> fe(5 * ctx->batch);
    HIPX(ctx, hipMemcpyAsync(fe.data(), ctx->fe, 8 * fe.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(out, fe.data(), 8 * fe.size());
    return check_fe(ctx, fe);
}

// Size in doubles of one iteration-state snapshot: the fields, phi, Q, F and,
// with Anderson mixing, the Gram matrix, coefficients, residual norms, the
// update counter and the one ring slot (of g and of d) the next update writes.
static size_t state_doubles(const scft_ctx *c)
{
    const size_t BN = (size_t)c->batch * c->N;
    size_t n = 4 * BN + 6 * (size_t)c->batch;
    if (c->and_cnt) {
        const size_t m1 = (size_t)c->p.anderson_m + 1;
        n += (size_t)c->batch * m1 * m1 + (size_t)c->batch * AND_MAX + c->batch + 1 + 4 * BN;
    }
    return n;
}

// Segments between the live iteration state and snapshot bk (to_bk: save,
// else restore); `slot` is the Anderson ring slot the next update overwrites.
static CopySegs state_segs(scft_ctx *c, double *bk, long slot, bool to_bk)
{
    const size_t BN = (size_t)c->batch * c->N;
    CopySegs sg{};
    double *p = bk;
    auto add = [&](double *live, size_t n) {
        sg.src[sg.count] = to_bk ? live : p;
        sg.dst[sg.count] = to_bk ? p : live;
        sg.n[sg.count] = (long)n;
        ++sg.count;
        p += n;
    };
    add(c->wA, BN);
    add(c->wB, BN);
    add(c->phiA, BN);
    add(c->phiB, BN);
    add(c->Q, c->batch);
    add(c->fe, 5 * (size_t)c->batch);
    if (c->and_cnt) {
        const size_t m1 = (size_t)c->p.anderson_m + 1;
        add(c->G, (size_t)c->batch * m1 * m1);
        add(c->coef, (size_t)c->batch * AND_MAX);
        add(c->aerr, c->batch);
        add(reinterpret_cast<double *>(c->and_cnt), 1);  // bits copied as they are
        add(c->wh + (size_t)slot * 2 * BN, 2 * BN);
        add(c->dh + (size_t)slot * 2 * BN, 2 * BN);
    }
    return sg;
}

// scft_iterate with a tolerance.  The host checks iteration i's inCompMax
// while iteration i+1 already runs, so the GPU never waits for the host;
// after every iteration one kernel saves its state to one of two snapshots,
// and when iteration i is the first below tol, the state after i is restored
// from its snapshot, undoing iteration i+1 entirely (Anderson history included).
static int iterate_tol(scft_ctx *c, int n_iter, double tol, int *done, std::vector<double> &fe)
{
    const size_t nb = state_doubles(c);
    for (int k = 0; k < 2; ++k) {
        if (!c->tol_bk[k]) {
            int rc = dalloc(c, (void **)&c->tol_bk[k], nb * sizeof(double));
            if (rc) return rc;
        }
        if (!c->tol_ev[k]) HIPX(c, hipEventCreateWithFlags(&c->tol_ev[k], hipEventDisableTiming));
    }
    const size_t nfe = 5 * (size_t)c->batch;
    if (!c->tol_fe) HIPX(c, hipHostMalloc((void **)&c->tol_fe, 2 * nfe * sizeof(double), hipHostMallocDefault));
    long cnt0 = 0;
    const long m1 = c->and_cnt ? c->p.anderson_m + 1 : 1;
    if (c->and_cnt) {
        HIPX(c, hipMemcpyAsync(&cnt0, c->and_cnt, sizeof(long), hipMemcpyDeviceToHost, c->stream));
        HIPX(c, hipStreamSynchronize(c->stream));
    }
    const unsigned grid = pt_grid((long)nb);
    auto check = [&](int i, bool &below) -> int {
        HIPX(c, hipEventSynchronize(c->tol_ev[i & 1]));
        std::memcpy(fe.data(), c->tol_fe + (i & 1) * nfe, nfe * sizeof(double));
        double worst = 0.0;
        for (int b = 0; b < c->batch; ++b) worst = std::max(worst, fe[b * 5 + 4]);
        below = worst < tol;
        return check_fe(c, fe);
    };
    int it = 0;
    for (; it < n_iter; ++it) {
        HIPX(c, launch_iteration(c));
        c->have_fe = 1;
        const long slot = (cnt0 + it + 1) % m1;  // ring slot the next update writes
        hipLaunchKernelGGL(k_copy_segs, dim3(grid), dim3(RT), 0, c->stream, state_segs(c, c->tol_bk[it & 1], slot, true));
        HIPX(c, hipGetLastError());
        HIPX(c, hipMemcpyAsync(c->tol_fe + (it & 1) * nfe, c->fe, nfe * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        HIPX(c, hipEventRecord(c->tol_ev[it & 1], c->stream));
        if (it >= 1) {
            bool below = false;
            const int rc = check(it - 1, below);
            if (rc || below) {  // stop after iteration it-1: undo iteration it
                const long s1 = (cnt0 + it) % m1;
                hipLaunchKernelGGL(k_copy_segs, dim3(grid), dim3(RT), 0, c->stream,
                                   state_segs(c, c->tol_bk[(it - 1) & 1], s1, false));
                HIPX(c, hipGetLastError());
                HIPX(c, hipStreamSynchronize(c->stream));
                if (done) *done = it;
                return rc ? rc : SCFT_OK;  // converged
            }
        }
    }
    if (it > 0) {
        bool below = false;
        const int rc = check(it - 1, below);
        if (done) *done = it;
        if (rc) return rc;
    }
    return SCFT_OK;
}

extern "C" int scft_iterate(scft_ctx *ctx, int n_iter, double tol, int *done, double *fe_out)
{
    if (!ctx) return SCFT_E_INVALID;
    if (done) *done = 0;
    if (n_iter < 0) return set_err(ctx, SCFT_E_INVALID, "n_iter must be >= 0");
    HIPX(ctx, hipSetDevice(ctx->device));
    std::vector<double> fe(5 * ctx->batch, 0.0);
    int it = 0;
    if (tol > 0.0) {
        const int rc = iterate_tol(ctx, n_iter, tol, &it, fe);
        if (done) *done = it;
        ctx->have_conc = 0;
        if (rc) return rc;
    } else {
        HIPX(ctx, launch_iterations(ctx, n_iter));
        it = n_iter;
        if (it > 0) ctx->have_fe = 1;
    }
    if (done) *done = it;
    ctx->have_conc = 0;
    if (it == 0) return SCFT_OK;
    HIPX(ctx, hipMemcpyAsync(fe.data(), ctx->fe, 8 * fe.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    if (fe_out) std::memcpy(fe_out, fe.data(), 8 * fe.size());
    return check_fe(ctx, fe);
}

extern "C" int scft_time_iterations(scft_ctx *ctx, int n_iter, float *ms)
{
    if (!ctx || !ms || n_iter < 1) return SCFT_E_INVALID;
    HIPX(ctx, hipSetDevice(ctx->device));
    HIPX(ctx, ensure_graph(ctx));  // capture + instantiate outside the timed region
    if (ctx->p.use_graph && n_iter >= 2 && !ctx->graph2_exec)
        HIPX(ctx, capture_iterations(ctx, 2, &ctx->graph2, &ctx->graph2_exec));
    HIPX(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPX(ctx, launch_iterations(ctx, n_iter));
    HIPX(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPX(ctx, hipEventSynchronize(ctx->ev1));
    HIPX(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    ctx->have_fe = 1;
    ctx->have_conc = 0;
    std::vector<double> fe(5 * ctx->batch);
    HIPX(ctx, hipMemcpy(fe.data(), ctx->fe, 8 * fe.size(), hipMemcpyDeviceToHost));
    return check_fe(ctx, fe);
}

template <typename T>
static hipError_t enqueue_pass_only(scft_ctx *c, int axis, int reps, int variant)
{
    using C = typename cplx<T>::type;
    PassArgs<T> pa;
    pa.src = (const C *)c->work;
    pa.dst = (C *)c->work;
    pa.W = static_cast<const T *>(c->wdz);
    pa.wo[0] = pa.wo[1] = pa.wo[2] = pa.wo[3] = 0u;
    // codes: see kPlaneCode.  A boundary kernel runs in place on the work
    // field, as in the march, with the arguments of consecutive contour steps.
    int form = 0;
    bool fused = false;
    code_form(axis, &form, &fused);
    pa.slot = fused ? slot_ptr<T>(c, 1) : nullptr;
    pa.wn[0] = pa.wn[1] = 0u;
    pa.post_scale = (T)(1.0 / (double)c->N);
    pa.post_from_pre = c->post_exact;
    pa.nsab = (const int2 *)c->nsab;
    pa.step = 0;
    pa.BN = (unsigned)(c->batch * c->N);
    const int ax = (axis >= 5 && axis <= 7) ? axis - 5 : (form ? 0 : axis);
    pa.kf_stride = c->kf_per_inst ? (unsigned)c->n[ax] : 0u;
    pa.tw = (const C *)c->tw[ax];
    pa.kf = (const typename kfac<T>::type *)c->kf[ax];
    pa.nlines = (unsigned)(c->batch * (c->N / c->n[ax]));
    pa.N = (unsigned)c->N;
    pa.nx = c->n[0];
    pa.ny = c->n[1];
    pa.nz = c->n[2];
    pa.pre = 0;
    pa.post = 0;
    const auto *kfy = (const typename kfac<T>::type *)c->kf[1];
    const long np = (long)c->batch * c->n[2];
    const int tmax = std::max(1, c->Ns_max - 1);
    for (int r = 0; r < reps; ++r) {
        const C *twy = (const C *)c->tw[1];
        if (fused) {  // as in the march: contour step t's W blocks, z(t) to its own slot
            const int t = 1 + r % tmax;
            pa.step = t;
            pa.slot = march_slot<T>(c, t);
        }
        hipError_t e = form    ? run_plane<T>(c, form, pa, twy, kfy, np, fused)
                       : fused ? run_pass<T>(c, ax, c->n[ax], variant, pa, true)
                               : run_pass<T>(c, axis, c->n[axis], variant, pa, false);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
__global__ void __launch_bounds__(RT) k_fill(T *__restrict__ p, T v, long n)
{
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) p[i] = v;
}

// mean duration (ms) of `reps` packed marches of Ns_max steps (the propagator
// part of an iteration, in the current plane form and tilings) on stand-in
// factors W = 1, W/N = 1/N and z(0) = 1.  Overwrites the W factors and the
// slice store, which the next iteration or concentrations call recomputes.
static int time_march(scft_ctx *ctx, int reps, float *ms)
{
    const size_t BN = (size_t)ctx->batch * ctx->N;
    const double invN = 1.0 / (double)ctx->N;
    if (ctx->fp32) {  // pre factors 1 (stored as 0 in FP32X, where they are W - 1)
        hipLaunchKernelGGL(k_fill<float>, dim3(pt_grid(2 * BN)), dim3(RT), 0, ctx->stream, (float *)ctx->wdz,
                           ctx->mixed ? 0.0f : 1.0f, (long)(2 * BN));
        hipLaunchKernelGGL(k_fill<float>, dim3(pt_grid(2 * BN)), dim3(RT), 0, ctx->stream, (float *)ctx->wdz + 2 * BN,
                           (float)invN, (long)(2 * BN));
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(pt_grid(2 * BN)), dim3(RT), 0, ctx->stream, (double *)ctx->wdz, 1.0,
                           (long)(2 * BN));
        hipLaunchKernelGGL(k_fill<double>, dim3(pt_grid(2 * BN)), dim3(RT), 0, ctx->stream,
                           (double *)ctx->wdz + 2 * BN, invN, (long)(2 * BN));
    }
    if (ctx->fp32)
        hipLaunchKernelGGL(k_seed_real<float>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (float2 *)ctx->slots,
                           (const double *)nullptr, (long)BN);
    else
        hipLaunchKernelGGL(k_seed_real<double>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (double2 *)ctx->slots,
                           (const double *)nullptr, (long)BN);
    HIPX(ctx, hipGetLastError());
    auto march = [&]() {
        return ctx->fp32 ? enqueue_march<float>(ctx, ctx->Ns_max, -1) : enqueue_march<double>(ctx, ctx->Ns_max, -1);
    };
    // the march as the iterations run it: one captured graph, replayed (an
    // eager march at small grids would time the host's launch rate)
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPX(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t em = march();
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (em != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_err(ctx, SCFT_E_HIP, std::string("march capture: ") +
                                            hipGetErrorString(em != hipSuccess ? em : ec));
    }
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(ge, ctx->stream);  // untimed: code loading, first touch
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = hipGraphLaunch(ge, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    HIPX(ctx, e);
    *ms = t / reps;
    return SCFT_OK;
}

// mean duration (ms) of `reps` back-to-back launches of one pass kernel on a
// field of ones in the work buffer
static int time_pass_variant(scft_ctx *ctx, int axis, int reps, int variant, float *ms)
{
    const size_t BN = (size_t)ctx->batch * ctx->N;
    if (ctx->fp32)
        hipLaunchKernelGGL(k_seed_real<float>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (float2 *)ctx->work,
                           (const double *)nullptr, (long)BN);
    else
        hipLaunchKernelGGL(k_seed_real<double>, dim3(pt_grid(BN)), dim3(RT), 0, ctx->stream, (double2 *)ctx->work,
                           (const double *)nullptr, (long)BN);
    HIPX(ctx, hipGetLastError());
    // one untimed launch so code loading and first-touch costs stay outside
    HIPX(ctx, ctx->fp32 ? enqueue_pass_only<float>(ctx, axis, 1, variant)
                        : enqueue_pass_only<double>(ctx, axis, 1, variant));
    HIPX(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPX(ctx, ctx->fp32 ? enqueue_pass_only<float>(ctx, axis, reps, variant)
                        : enqueue_pass_only<double>(ctx, axis, reps, variant));
    HIPX(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPX(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    HIPX(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t / reps;
    return SCFT_OK;
}

static int time_contour(scft_ctx *ctx, int reps, float *ms);

extern "C" int scft_time_pass(scft_ctx *ctx, int axis, int reps, float *ms)
{
    if (!ctx || !ms || reps < 1 || axis < 0 || axis >= PASS_CODES) return SCFT_E_INVALID;
    if (axis == CONTOUR_CODE) return time_contour(ctx, reps, ms);
    int form = 0;
    bool fused = false;
    code_form(axis, &form, &fused);
    if (form && !plane_fits(ctx, ctx->n[0], ctx->n[1], form))
        return set_err(ctx, SCFT_E_INVALID, "no xy-plane kernel of this form for this grid");
    const int line_axis = axis < 3 ? axis : (axis >= 5 && axis <= 7 ? axis - 5 : -1);
    if (line_axis >= 0 && !ctx->active[line_axis]) return set_err(ctx, SCFT_E_INVALID, "axis has length 1 (no pass)");
    HIPX(ctx, hipSetDevice(ctx->device));
    return time_pass_variant(ctx, axis, reps, line_axis >= 0 ? ctx->pass_variant[line_axis] : 0, ms);
}

// Autotuner: time every tiling variant of every active axis pass (and the
// fused xy-plane kernel where it fits) on this context's grid, keep the
// fastest combination, and drop the captured graph so the next iteration
// re-captures with the new choice.
extern "C" int scft_autotune(scft_ctx *ctx, int reps, int *choice, float *ms_out)
{
    if (!ctx || reps < 1) return SCFT_E_INVALID;
    HIPX(ctx, hipSetDevice(ctx->device));
    float best_sep = 0.0f;
    float per_axis[3] = {0.0f, 0.0f, 0.0f};
    for (int a = 0; a < 3; ++a) {
        if (!ctx->active[a]) continue;
        int bv = 0;
        float bt = 1e30f;
        for (int v = 0; v < PASS_VARIANTS; ++v) {
            float t = 0.0f;
            int rc = time_pass_variant(ctx, a, reps, v, &t);
            if (rc == SCFT_E_HIP && v > 0) {  // tiling does not fit this size: skip it
                (void)hipGetLastError();
                continue;
            }
            if (rc) return rc;
            if (t < bt) {
                bt = t;
                bv = v;
            }
        }
        ctx->pass_variant[a] = bv;
        per_axis[a] = bt;
        best_sep += bt;
    }
    // xy-plane form: forced by p.xy_plane > 0, off at 0; at -1 every form that
    // fits is timed over whole marches (with step-boundary fusion the plane and
    // line kernels trade places differently than their single-pass times say)
    (void)best_sep;
    int plane = ctx->p.xy_plane > 0 ? ctx->p.xy_plane : 0;
    if (ctx->p.xy_plane < 0) {
        float best = 0.0f;
        for (int form = 0; form <= PLANE_FORMS; ++form) {
            if (form > 0 && !plane_fits(ctx, ctx->n[0], ctx->n[1], form)) continue;
            ctx->use_plane = form;
            float t = 0.0f;
            int rc = time_march(ctx, 3, &t);
            if (rc) return rc;
            if (form == 0 || t < best) {
                best = t;
                plane = form;
            }
        }
    }
    ctx->use_plane = plane;
    // refine the axis variants on whole marches: the boundary kernels follow
    // their axis's variant, and a pass that wins alone need not win there
    if (ctx->p.fuse_steps) {
        float best = 0.0f;
        int rc = time_march(ctx, 3, &best);
        if (rc) return rc;
        for (int a = plane ? 2 : 0; a < 3; ++a) {
            if (!ctx->active[a]) continue;
            const int keep = ctx->pass_variant[a];
            for (int v = 0; v < PASS_VARIANTS; ++v) {
                if (v == keep) continue;
                float t = 0.0f;
                if (time_pass_variant(ctx, a, 1, v, &t) != SCFT_OK) {  // no such tiling for this length
                    (void)hipGetLastError();
                    continue;
                }
                const int prev = ctx->pass_variant[a];
                ctx->pass_variant[a] = v;
                if ((rc = time_march(ctx, 3, &t))) return rc;
                if (t < best) best = t;
                else ctx->pass_variant[a] = prev;
            }
            per_axis[a] = 0.0f;
            (void)time_pass_variant(ctx, a, reps, ctx->pass_variant[a], &per_axis[a]);
        }
    }
    float tplane = 0.0f;
    if (plane) {
        int rc = time_pass_variant(ctx, kPlaneCode[plane][0], reps, 0, &tplane);
        if (rc) return rc;
    }
    drop_graphs(ctx);
    if (choice) {
        choice[0] = ctx->pass_variant[0];
        choice[1] = ctx->pass_variant[1];
        choice[2] = ctx->pass_variant[2];
        choice[3] = plane;
    }
    if (ms_out) {
        ms_out[0] = per_axis[0];
        ms_out[1] = per_axis[1];
        ms_out[2] = per_axis[2];
        ms_out[3] = tplane;
    }
    return SCFT_OK;
}

extern "C" int scft_set_tiling(scft_ctx *ctx, const int *choice)
{
    if (!ctx || !choice) return SCFT_E_INVALID;
    HIPX(ctx, hipSetDevice(ctx->device));
    for (int a = 0; a < 3; ++a)
        if (choice[a] < 0 || choice[a] >= PASS_VARIANTS)
            return set_err(ctx, SCFT_E_INVALID, "tiling variant must be 0.." + std::to_string(PASS_VARIANTS - 1));
    if (choice[3] < 0 || choice[3] > PLANE_FORMS)
        return set_err(ctx, SCFT_E_INVALID, "plane form must be 0.." + std::to_string(PLANE_FORMS));
    if (choice[3] != 0 && !plane_fits(ctx, ctx->n[0], ctx->n[1], choice[3]))
        return set_err(ctx, SCFT_E_UNSUPPORTED, "xy-plane kernel does not fit this grid");
    for (int a = 0; a < 3; ++a) {  // one trial launch: a variant whose LDS does not fit fails here
        if (!ctx->active[a] || choice[a] == 0) continue;
        float t = 0.0f;
        if (time_pass_variant(ctx, a, 1, choice[a], &t) != SCFT_OK) {
            (void)hipGetLastError();
            return set_err(ctx, SCFT_E_UNSUPPORTED, "tiling variant " + std::to_string(choice[a]) +
                                                        " does not fit axis " + std::to_string(a));
        }
    }
    for (int a = 0; a < 3; ++a) ctx->pass_variant[a] = choice[a];
    ctx->use_plane = choice[3];
    drop_graphs(ctx);
    return SCFT_OK;
}

extern "C" int scft_set_precision(scft_ctx *ctx, int precision)
{
    if (!ctx) return SCFT_E_INVALID;
    if (precision != SCFT_FP32 && precision != SCFT_FP32X)
        return set_err(ctx, SCFT_E_INVALID, "scft_set_precision switches between SCFT_FP32 and SCFT_FP32X");
    if (!ctx->fp32)
        return set_err(ctx, SCFT_E_STATE, "scft_set_precision needs a context created with FP32 storage "
                                          "(SCFT_FP32 or SCFT_FP32X)");
    if (precision == ctx->p.precision) return SCFT_OK;
    HIPX(ctx, hipSetDevice(ctx->device));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < 3; ++a) {
        std::swap(ctx->tw[a], ctx->tw_alt[a]);
        std::swap(ctx->kf[a], ctx->kf_alt[a]);
    }
    ctx->mixed = precision == SCFT_FP32X;
    ctx->p.precision = precision;
    // the tiling carries over where the new form has it (its registers and LDS differ); else the
    // default variant and separate x and y passes
    if (ctx->use_plane && !plane_fits(ctx, ctx->n[0], ctx->n[1], ctx->use_plane)) ctx->use_plane = 0;
    for (int a = 0; a < 3; ++a) {
        if (!ctx->active[a] || ctx->pass_variant[a] == 0) continue;
        float t = 0.0f;
        if (time_pass_variant(ctx, a, 1, ctx->pass_variant[a], &t) != SCFT_OK) {
            (void)hipGetLastError();
            ctx->pass_variant[a] = 0;
        }
    }
    drop_graphs(ctx);
    return SCFT_OK;
}

// ===========================================================================
// Discrete adjoint (no reference counterpart)
// ===========================================================================
// Reverse mode of one chain evaluation: for fields u = (uA, uB) used as given,
// L = sum_r (lamA phiA + lamB phiB) + mu ln Q, dL/du.  The step operator
// S = W K W is symmetric, so the adjoint recursions
//   qbar(Ns) = e(Ns),  qbar(j-1) = e(j-1) + S qbar(j)
//   qbar_dag(0) = ed(0), qbar_dag(j) = ed(j) + S qbar_dag(j-1)
// (e, ed: explicit derivatives of L in q, qdag) are the forward packed march
// with point sources: zb(tau) = qbar_dag(tau) + i qbar(Ns - tau).  Summing the
// W-derivatives of every step gives
//   dL/du_b = -ds trap_b[ qbar q + qbar_dag qdag - g q qdag / Q ],
// whose terms at contour points tau and Ns - tau need only z(tau) and
// z(Ns - tau) at adjoint step tau (k_adj_point).
static int ensure_adjoint_scratch(scft_ctx *c)
{
    if (c->adj) return SCFT_OK;
    const size_t BN = (size_t)c->batch * c->N;
    int rc = dalloc(c, (void **)&c->adj, 8 * BN * sizeof(double) + 2 * sizeof(double) * c->batch);
    return rc;
}

// `reps` contour launches over the current slice store into the adjoint
// scratch (phi, Q and ds/Q of the context are not touched)
static int time_contour(scft_ctx *ctx, int reps, float *ms)
{
    if (ctx->ckpt)
        return set_err(ctx, SCFT_E_INVALID, "checkpoint mode sums the contour during the march (no single contour kernel)");
    HIPX(ctx, hipSetDevice(ctx->device));
    int rc = ensure_adjoint_scratch(ctx);
    if (rc) return rc;
    const size_t BN = (size_t)ctx->batch * ctx->N;
    double *pA = ctx->adj + 6 * BN, *pB = pA + BN, *q = pB + BN, *f = q + ctx->batch;
    auto one = [&]() {
        return ctx->fp32 ? launch_contour<float>(ctx, pA, pB, q, f) : launch_contour<double>(ctx, pA, pB, q, f);
    };
    HIPX(ctx, one());  // untimed
    HIPX(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps; ++r) HIPX(ctx, one());
    HIPX(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPX(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    HIPX(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t / reps;
    return SCFT_OK;
}

// W factors of the scratch fields u (no mean subtraction) and slice 0
template <typename T>
static hipError_t enqueue_prep_scratch(scft_ctx *c)
{
    double *uA = c->adj, *uB = uA + (size_t)c->batch * c->N;
    launch_prep<T>(c, uA, (const double *)uB, (const double *)nullptr);
    return hipGetLastError();
}

// forward = false: the caller has restored what the forward evaluation of u
// leaves behind (W factors, the slice store or the checkpoints, Q, phi at pA/pB)
template <typename T>
static hipError_t enqueue_conc_adjoint(scft_ctx *c, double mu, bool forward = true)
{
    using C = typename cplx<T>::type;
    const size_t BN = (size_t)c->batch * c->N;
    double *uA = c->adj, *uB = uA + BN, *lA = uB + BN, *lB = lA + BN, *gA = lB + BN, *gB = gA + BN, *pA = gB + BN,
           *pB = pA + BN, *sc = pB + BN;
    hipError_t e = forward ? enqueue_conc_fields<T>(c, uA, uB, nullptr, pA, pB) : hipSuccess;
    if (e != hipSuccess) return e;
    const dim3 g(c->nblk, c->batch);
    hipLaunchKernelGGL(k_dot_lamphi, g, dim3(RT), 0, c->stream, (const double *)lA, (const double *)lB,
                       (const double *)pA, (const double *)pB, c->N, c->nblk, c->part2);
    hipLaunchKernelGGL(k_adj_scalars, dim3(c->batch), dim3(RT), 0, c->stream, (const double *)c->part2, c->nblk,
                       (const double *)c->Q, mu, sc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const long zb0 = c->ckpt ? (long)c->ck_n + 2 * c->ck_K + 2 : (long)c->Ns + 1;
    const dim3 gc((unsigned)((c->N + RT - 1) / RT), c->batch);
    PairWalk<T> walk(c);
    for (int tau = 0; tau <= c->Ns; ++tau) {
        const C *zt, *zu;
        if ((e = walk.next(tau, &zt, &zu)) != hipSuccess) return e;
        C *zb = slot_ptr<T>(c, zb0 + (tau & 1));
        if (tau > 0 && (e = enqueue_step<T>(c, slot_ptr<T>(c, zb0 + ((tau - 1) & 1)), zb, tau, 0, 0)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_adj_point<T>, gc, dim3(RT), 0, c->stream, zb, zt, zu, (const double *)lA,
                           (const double *)lB, tau, c->NsA, c->Ns, c->p.ds, c->N, c->batch, (const double *)c->Q,
                           (const double *)sc, gA, gB, tau == 0 ? 1 : 0);
    }
    return hipGetLastError();
}

// device forward of one chain evaluation on the adjoint scratch fields (uA, uB -> phiA, phiB, Q)
static hipError_t enqueue_conc_scratch(scft_ctx *c)
{
    const size_t BN = (size_t)c->batch * c->N;
    double *uA = c->adj, *uB = uA + BN, *pA = uB + 5 * BN, *pB = pA + BN;
    return c->fp32 ? enqueue_conc_fields<float>(c, uA, uB, nullptr, pA, pB)
                   : enqueue_conc_fields<double>(c, uA, uB, nullptr, pA, pB);
}

extern "C" int scft_conc_adjoint(scft_ctx *ctx, const double *uA, const double *uB, const double *lamA,
                                 const double *lamB, double mu, double *L, double *phiA, double *phiB, double *Q,
                                 double *gA, double *gB)
{
    if (!ctx || !uA || !uB) return ctx ? set_err(ctx, SCFT_E_INVALID, "uA/uB are NULL") : SCFT_E_INVALID;
    if (!ctx->blocks_uniform) return set_err(ctx, SCFT_E_UNSUPPORTED, "the adjoint needs one (NsA, NsB) for all instances");
    HIPX(ctx, hipSetDevice(ctx->device));
    int rc = ensure_adjoint_scratch(ctx);
    if (rc) return rc;
    const size_t BN = (size_t)ctx->batch * ctx->N;
    double *d = ctx->adj;
    HIPX(ctx, hipMemcpyAsync(d, uA, 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    HIPX(ctx, hipMemcpyAsync(d + BN, uB, 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    std::vector<double> zero;
    if (!lamA || !lamB) zero.assign(BN, 0.0);
    HIPX(ctx, hipMemcpyAsync(d + 2 * BN, lamA ? lamA : zero.data(), 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    HIPX(ctx, hipMemcpyAsync(d + 3 * BN, lamB ? lamB : zero.data(), 8 * BN, hipMemcpyHostToDevice, ctx->stream));
    HIPX(ctx, ctx->fp32 ? enqueue_conc_adjoint<float>(ctx, mu) : enqueue_conc_adjoint<double>(ctx, mu));
    std::vector<double> sc(2 * ctx->batch), q(ctx->batch);
    HIPX(ctx, hipMemcpyAsync(sc.data(), d + 8 * BN, 16 * ctx->batch, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipMemcpyAsync(q.data(), ctx->Q, 8 * ctx->batch, hipMemcpyDeviceToHost, ctx->stream));
    if (gA) HIPX(ctx, hipMemcpyAsync(gA, d + 4 * BN, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (gB) HIPX(ctx, hipMemcpyAsync(gB, d + 5 * BN, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (phiA) HIPX(ctx, hipMemcpyAsync(phiA, d + 6 * BN, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    if (phiB) HIPX(ctx, hipMemcpyAsync(phiB, d + 7 * BN, 8 * BN, hipMemcpyDeviceToHost, ctx->stream));
    HIPX(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_conc = 0;
    for (int b = 0; b < ctx->batch; ++b) {
        if (!std::isfinite(q[b]) || !(q[b] > 0.0))
            return set_err(ctx, SCFT_E_NONFINITE, "Q of instance " + std::to_string(b) + " is not finite and > 0");
        if (L) L[b] = sc[2 * b];
        if (Q) Q[b] = q[b];
    }
    return SCFT_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of the unrolled SCFT loop with an external field h on the A blocks.
// Forward, k = 0..n-1:
//   wt = P w (mean removed per instance and field), u = (wtA + h, wtB),
//   (phi, Q) = conc(u); if k < n-1: d = R(wt, phi), w = mix(wt, d, history)
// Loss = sum_r (phiA(n-1) - target)^2 per instance.  R is the residual that
// simple mixing steps along (d_A = chiN phiB + eta - wA - wcmp psi, ...);
// mix is simple mixing (w = wt + wopt d) or Anderson (the library's
// algorithm, with U and V formed directly from the differences).  The reverse
// sweep runs the Anderson least-squares adjoint (cbar -> Vbar = U^-1 cbar,
// Ubar = -Vbar c^T), the residual's linear adjoint, the chain adjoint, and P
// (symmetric) for the mean removal.
//
// Everything runs on the device in FP64 (the chain in the context's
// precision): the trajectory (wt, phi, d per iteration) and its adjoint
// accumulators stay in device memory, the passes over grid points are the
// kernels below, and only the m x m least-squares solves (forward and
// reverse) run on the host, on sums the device reduces in a fixed order.
// ---------------------------------------------------------------------------
namespace {

// Gaussian elimination with partial pivoting on an n x n system (row major)
bool host_solve(int n, std::vector<double> A, std::vector<double> &x)
{
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > best) {
                best = std::fabs(A[i * n + k]);
                piv = i;
            }
        if (best == 0.0) return false;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(x[k], x[piv]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            x[i] -= f * x[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double s = x[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * x[j];
        x[k] = s / A[k * n + k];
    }
    return true;
}

// A trajectory slice holds [2][batch][N] doubles (field, instance, point);
// point q in [0, 2N) of instance b is element (q / N) * BN + b * N + q % N.
constexpr int LCH = RT;  // points per LDS chunk of k_loop_dot (one per thread)
constexpr int LMAXE = 3 * RT;  // entries per reduction (three per thread)
static_assert(AND_MAX + AND_MAX * (AND_MAX + 1) / 2 <= LMAXE, "Gram entries exceed the reduction");

enum LoopDotMode { LD_GRAM = 0, LD_CBAR = 1, LD_LOSS = 2, LD_MEAN = 3 };

struct LoopDotArgs {
    const double *x0;   // GRAM: d_k; CBAR: wbar; LOSS: phi; MEAN: the field
    const double *dd;   // residual slices (GRAM, CBAR)
    const double *wt;   // centered-field slices (CBAR)
    const double *tgt;  // LOSS: target [batch][N]
    long k;             // current slice (GRAM, CBAR)
    int mp;             // history length (GRAM, CBAR)
    double lam;         // mixing fraction (CBAR)
    long N, BN;
    int nblk;  // blocks per instance
    int ne;    // entries
};

// Per-block partial sums part[(b * nblk + blk) * ne + e] of the entries:
//   GRAM: e < mp: V_i = <D_i, d_k>; then U_ij = <D_i, D_j>, 1 <= i <= j <= mp
//         in row order, D_i = d_k - d_{k-i}
//   CBAR: <wbar, (wt_{k-i} - wt_k) + lam (d_{k-i} - d_k)>, i = 1..mp
//   LOSS: sum over the A field of (phi - target)^2
//   MEAN: the sums of field A and of field B
// over both fields of instance b.  Each block stages LCH points at a time in
// LDS, every thread accumulates its entries over the chunk, and the partials
// combine in block order (k_loop_dot_final): the same bits on every run.
template <int MODE>
__global__ void __launch_bounds__(RT) k_loop_dot(LoopDotArgs a, double *__restrict__ part)
{
    constexpr int XR = (MODE == LD_GRAM || MODE == LD_CBAR) ? AND_MAX + 1 : (MODE == LD_MEAN ? 3 : 1);
    __shared__ double X[XR][LCH + 1];
    const int b = blockIdx.y, tid = threadIdx.x;
    const int nx = (MODE == LD_GRAM || MODE == LD_CBAR) ? a.mp + 1 : XR;
    int ia[3], ib[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int e = tid + s * RT;
        ia[s] = -1;
        ib[s] = 0;
        if (e >= a.ne) continue;
        if (MODE == LD_GRAM) {
            if (e < a.mp) {
                ia[s] = 0;
                ib[s] = e + 1;
            } else {
                int t = e - a.mp, i = 1;
                while (t >= a.mp - i + 1) {
                    t -= a.mp - i + 1;
                    ++i;
                }
                ia[s] = i;
                ib[s] = i + t;
            }
        } else if (MODE == LD_CBAR) {
            ia[s] = 0;
            ib[s] = e + 1;
        } else if (MODE == LD_LOSS) {
            ia[s] = 0;
            ib[s] = 0;
        } else {
            ia[s] = e;
            ib[s] = 2;
        }
    }
    double acc[3] = {0.0, 0.0, 0.0};
    const long n2 = 2 * a.N;
    const size_t S = 2 * (size_t)a.BN;
    for (long c0 = (long)blockIdx.x * LCH; c0 < n2; c0 += (long)a.nblk * LCH) {
        const long q = c0 + tid;
        if (q < n2) {
            const long f = q / a.N, p = q - f * a.N;
            const size_t idx = (size_t)f * a.BN + (size_t)b * a.N + p;
            if (MODE == LD_GRAM) {
                const double d0 = a.x0[idx];
                X[0][tid] = d0;
                for (int i = 1; i <= a.mp; ++i) X[i][tid] = d0 - a.dd[(size_t)(a.k - i) * S + idx];
            } else if (MODE == LD_CBAR) {
                X[0][tid] = a.x0[idx];
                const double w0 = a.wt[(size_t)a.k * S + idx], d0 = a.dd[(size_t)a.k * S + idx];
                for (int i = 1; i <= a.mp; ++i)
                    X[i][tid] = (a.wt[(size_t)(a.k - i) * S + idx] - w0) + a.lam * (a.dd[(size_t)(a.k - i) * S + idx] - d0);
            } else if (MODE == LD_LOSS) {
                X[0][tid] = f == 0 ? a.x0[idx] - a.tgt[(size_t)b * a.N + p] : 0.0;
            } else {
                const double v = a.x0[idx];
                X[0][tid] = f == 0 ? v : 0.0;
                X[1][tid] = f == 1 ? v : 0.0;
                X[2][tid] = 1.0;
            }
        } else {
            for (int i = 0; i < nx; ++i) X[i][tid] = 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            if (ia[s] < 0) continue;
            double t = 0.0;
            for (int j = 0; j < LCH; ++j) t += X[ia[s]][j] * X[ib[s]][j];
            acc[s] += t;
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 3; ++s)
        if (ia[s] >= 0) part[((size_t)b * a.nblk + blockIdx.x) * a.ne + tid + s * RT] = acc[s];
}

// out[b * ne + e] = sum over blocks of the partials, in block order
__global__ void __launch_bounds__(RT) k_loop_dot_final(const double *__restrict__ part, int nblk, int ne,
                                                       double *__restrict__ out)
{
    const int b = blockIdx.x;
    for (int e = threadIdx.x; e < ne; e += RT) {
        double s = 0.0;
        for (int k = 0; k < nblk; ++k) s += part[((size_t)b * nblk + k) * ne + e];
        out[(size_t)b * ne + e] = s;
    }
}

// element i of a slice: field f = i / BN, instance (i % BN) / N
__device__ __forceinline__ long loop_inst(long i, long N, long BN) { return (i - (i / BN) * BN) / N; }

// dst = src minus the mean of its (field, instance); sums [batch][2] from LD_MEAN
__global__ void __launch_bounds__(RT) k_loop_center(const double *__restrict__ src, double *__restrict__ dst,
                                                    const double *__restrict__ sums, long N, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    const long f = i / BN;
    dst[i] = src[i] - sums[2 * loop_inst(i, N, BN) + f] / (double)N;
}

// chain input u = (wtA + h, wtB)
__global__ void __launch_bounds__(RT) k_loop_u(const double *__restrict__ wt, const double *__restrict__ h,
                                               double *__restrict__ u, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    u[i] = i < BN ? wt[i] + h[i] : wt[i];
}

// residual d = R(wt, phi)
__global__ void __launch_bounds__(RT) k_loop_resid(const double *__restrict__ wt, const double *__restrict__ ph,
                                                   double *__restrict__ d, double chiN, double wcmp, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= BN) return;
    const double a = wt[i], c = wt[BN + i], pa = ph[i], pb = ph[BN + i];
    const double eta = (a + c - chiN) / 2, psum = 1.0 - pa - pb;
    d[i] = chiN * pb + eta - a - wcmp * psum;
    d[BN + i] = chiN * pa + eta - c - wcmp * psum;
}

// w = wt_k + sum_i c_i (wt_{k-i} - wt_k) + lam (d_k + sum_i c_i (d_{k-i} - d_k));
// coef [batch][AND_MAX]
__global__ void __launch_bounds__(RT) k_loop_mix(const double *__restrict__ WT, const double *__restrict__ DD, long k,
                                                 int mp, const double *__restrict__ coef, double lam,
                                                 double *__restrict__ w, long N, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    const size_t S = 2 * (size_t)BN;
    const double *c = coef + loop_inst(i, N, BN) * AND_MAX;
    const double w0 = WT[(size_t)k * S + i], d0 = DD[(size_t)k * S + i];
    double wn = w0, dn = d0;
    for (int j = 1; j <= mp; ++j) {
        wn += c[j - 1] * (WT[(size_t)(k - j) * S + i] - w0);
        dn += c[j - 1] * (DD[(size_t)(k - j) * S + i] - d0);
    }
    w[i] = wn + lam * dn;
}

// seed of the reverse sweep: pbar = (2 (phiA - target), 0)
__global__ void __launch_bounds__(RT) k_loop_pbar0(const double *__restrict__ ph, const double *__restrict__ tgt,
                                                   double *__restrict__ pbar, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    pbar[i] = i < BN ? 2.0 * (ph[i] - tgt[i]) : 0.0;
}

// adjoint of k_loop_mix with c held fixed: wbar into Wtbar and Dbar of k and its history
__global__ void __launch_bounds__(RT) k_loop_mix_adj(const double *__restrict__ wbar, double *__restrict__ WTB,
                                                     double *__restrict__ DB, long k, int mp,
                                                     const double *__restrict__ coef, double lam, long N, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    const size_t S = 2 * (size_t)BN;
    const double *c = coef + loop_inst(i, N, BN) * AND_MAX;
    double sc = 0.0;
    for (int j = 0; j < mp; ++j) sc += c[j];
    const double wb = wbar[i];
    WTB[(size_t)k * S + i] += (1.0 - sc) * wb;
    DB[(size_t)k * S + i] += lam * (1.0 - sc) * wb;
    for (int j = 1; j <= mp; ++j) {
        WTB[(size_t)(k - j) * S + i] += c[j - 1] * wb;
        DB[(size_t)(k - j) * S + i] += lam * c[j - 1] * wb;
    }
}

// least-squares part of the Anderson adjoint: Ubar_ij = -Vbar_i c_j and the
// ridge term give Dbar_i = Vbar_i (d_k - S1) - c_i S2 + 2 rdiag Delta_i with
// S1 = sum_j c_j Delta_j, S2 = sum_j Vbar_j Delta_j; rv [batch][AND_MAX + 1]
// holds Vbar and, last, rdiag (zeros for an instance whose solve failed)
__global__ void __launch_bounds__(RT) k_loop_dbar(const double *__restrict__ DD, double *__restrict__ DB, long k,
                                                  int mp, const double *__restrict__ coef,
                                                  const double *__restrict__ rv, long N, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    const size_t S = 2 * (size_t)BN;
    const long b = loop_inst(i, N, BN);
    const double *c = coef + b * AND_MAX, *vb = rv + b * (AND_MAX + 1);
    const double rdiag = vb[AND_MAX];
    const double d0 = DD[(size_t)k * S + i];
    double S1 = 0.0, S2 = 0.0;
    for (int j = 1; j <= mp; ++j) {
        const double Dj = d0 - DD[(size_t)(k - j) * S + i];
        S1 += c[j - 1] * Dj;
        S2 += vb[j - 1] * Dj;
    }
    double acc0 = 0.0;
    for (int j = 1; j <= mp; ++j) {
        const double Dj = d0 - DD[(size_t)(k - j) * S + i];
        const double Dbj = vb[j - 1] * (d0 - S1) - c[j - 1] * S2 + 2.0 * rdiag * Dj;
        acc0 += vb[j - 1] * Dj + Dbj;
        DB[(size_t)(k - j) * S + i] -= Dbj;
    }
    DB[(size_t)k * S + i] += acc0;
}

// residual's linear adjoint: Dbar_k into Wtbar_k and pbar (pbar assigned)
__global__ void __launch_bounds__(RT) k_loop_resid_adj(const double *__restrict__ DBk, double *__restrict__ WTBk,
                                                       double *__restrict__ pbar, double chiN, double wcmp, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= BN) return;
    const double da = DBk[i], db = DBk[BN + i];
    WTBk[i] += -0.5 * da + 0.5 * db;
    WTBk[BN + i] += 0.5 * da - 0.5 * db;
    pbar[i] = wcmp * da + (chiN + wcmp) * db;
    pbar[BN + i] = (chiN + wcmp) * da + wcmp * db;
}

// ubar = dL/du of the chain: into hbar (A field) and Wtbar_k
__global__ void __launch_bounds__(RT) k_loop_ubar(const double *__restrict__ ubar, double *__restrict__ hbar,
                                                  double *__restrict__ WTBk, long BN)
{
    const long i = (long)blockIdx.x * RT + threadIdx.x;
    if (i >= 2 * BN) return;
    WTBk[i] += ubar[i];
    if (i < BN) hbar[i] += ubar[i];
}

__global__ void __launch_bounds__(RT) k_loop_zero(double *__restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * RT + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// device memory of one scft_adjoint_loop call, kept in the context (grown on demand)
int ensure_loop_buffer(scft_ctx *c, size_t bytes)
{
    if (c->loop_bytes >= bytes) return SCFT_OK;
    if (c->loop) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->loop);
        c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)c->loop), c->allocs.end());
        c->device_bytes -= (long)c->loop_bytes;
        c->loop = nullptr;
        c->loop_bytes = 0;
    }
    int rc = dalloc(c, (void **)&c->loop, bytes);
    if (rc) return rc;
    c->loop_bytes = bytes;
    return SCFT_OK;
}

}  // namespace

extern "C" int scft_adjoint_loop(scft_ctx *ctx, const double *h, const double *w0A, const double *w0B, int n_iter,
                                 const double *target, double *loss, double *grad_h, double *phiA_last)
{
    if (!ctx) return SCFT_E_INVALID;
    if (!h || !w0A || !w0B || !target || n_iter < 1)
        return set_err(ctx, SCFT_E_INVALID, "scft_adjoint_loop: NULL input or n_iter < 1");
    if (!ctx->blocks_uniform) return set_err(ctx, SCFT_E_UNSUPPORTED, "the adjoint needs one (NsA, NsB) for all instances");
    HIPX(ctx, hipSetDevice(ctx->device));
    int rc = ensure_adjoint_scratch(ctx);
    if (rc) return rc;
    const long N = ctx->N;
    const int nb = ctx->batch;
    const long BN = (long)nb * N;
    const size_t S = 2 * (size_t)BN;  // doubles per trajectory slice
    const double chiN = ctx->p.chiN, wcmp = ctx->p.wcmp;
    const bool anderson = ctx->p.mixing == SCFT_MIX_ANDERSON;
    const double ridge = ctx->p.anderson_reg;
    const int m = anderson ? ctx->p.anderson_m : 0;
    const double lam = anderson ? ctx->p.anderson_mix : ctx->p.wopt;
    const int K = n_iter;
    const bool grad = grad_h != nullptr;
    const int nbl = (int)std::min<long>(256, (2 * N + LCH - 1) / LCH);  // reduction blocks per instance
    hipStream_t s = ctx->stream;
    // device layout (doubles): WT, PH, DD [K][S]; WTB, DB [K][S] (gradient only);
    // w, wbar [S]; hbar, h, target [BN]; partials; sums; coefficients; Vbar/rdiag; Q
    const size_t n_traj = (size_t)K * S;
    const size_t n_part = (size_t)nbl * nb * LMAXE, n_red = (size_t)nb * LMAXE;
    const size_t n_coef = (size_t)K * nb * AND_MAX, n_rv = (size_t)K * nb * (AND_MAX + 1), n_q = (size_t)K * nb;
    // checkpoint mode: each iteration's checkpoints z(kK) (the slice store's first ck_n slots) and, with
    // step-boundary kernels, the march's opened states O(kK) (ck_n slots from nck + 2K + 4), so that the
    // reverse sweep restores the chain's forward state instead of recomputing it (march and contour)
    const bool keep_ck = grad && ctx->ckpt;
    int units3[3];
    const bool keep_open = keep_ck && ctx->p.fuse_steps && step_units(ctx, units3) >= 2;
    const size_t ck_bytes = (size_t)ctx->ck_n * BN * ctx->csize, ck_rec = (keep_open ? 2 : 1) * ck_bytes;
    const size_t n_ck = keep_ck ? (size_t)K * ck_rec / 8 : 0;
    char *const open_slots = (char *)ctx->slots + ((size_t)ctx->ck_n + 2 * ctx->ck_K + 4) * BN * ctx->csize;
    const size_t total =
        (grad ? 5 : 3) * n_traj + 2 * S + 3 * (size_t)BN + n_part + n_red + n_coef + n_rv + n_q + n_ck;
    if ((rc = ensure_loop_buffer(ctx, total * sizeof(double)))) return rc;
    double *p = ctx->loop;
    auto take = [&p](size_t n) {
        double *r = p;
        p += n;
        return r;
    };
    double *WT = take(n_traj), *PH = take(n_traj), *DD = take(n_traj);
    double *WTB = grad ? take(n_traj) : nullptr, *DB = grad ? take(n_traj) : nullptr;
    double *wd = take(S), *wbar = take(S), *hbar = take(BN), *hd = take(BN), *tg = take(BN);
    double *part = take(n_part), *red = take(n_red), *coefd = take(n_coef), *rvd = take(n_rv), *Qk = take(n_q);
    char *CK = reinterpret_cast<char *>(take(n_ck));
    // pinned host side of the small transfers: sums, coefficients, Vbar/rdiag, Q
    struct Pinned {
        double *p = nullptr;
        ~Pinned()
        {
            if (p) (void)hipHostFree(p);
        }
    } hp;
    HIPX(ctx, hipHostMalloc((void **)&hp.p, sizeof(double) * (n_red + n_coef + n_rv + n_q), hipHostMallocDefault));
    double *h_red = hp.p, *h_coef = h_red + n_red, *h_rv = h_coef + n_coef, *h_q = h_rv + n_rv;
    std::memset(h_coef, 0, sizeof(double) * (n_coef + n_rv));
    double *dev = ctx->adj;  // chain scratch: u [0, 2BN), lam = pbar [2BN, 4BN), ubar [4BN, 6BN), phi [6BN, 8BN)
    const unsigned g1 = (unsigned)((BN + RT - 1) / RT), g2 = (unsigned)((2 * BN + RT - 1) / RT);
    auto zero = [&](double *x, size_t n) {
        hipLaunchKernelGGL(k_loop_zero, dim3((unsigned)((n + RT - 1) / RT)), dim3(RT), 0, s, x, n);
    };
    LoopDotArgs la{};
    la.N = N;
    la.BN = BN;
    la.nblk = nbl;
    // sums into red[b * ne + e]
    auto dot = [&](int mode, int ne) {
        la.ne = ne;
        const dim3 g((unsigned)nbl, (unsigned)nb);
        switch (mode) {
        case LD_GRAM: hipLaunchKernelGGL(k_loop_dot<LD_GRAM>, g, dim3(RT), 0, s, la, part); break;
        case LD_CBAR: hipLaunchKernelGGL(k_loop_dot<LD_CBAR>, g, dim3(RT), 0, s, la, part); break;
        case LD_LOSS: hipLaunchKernelGGL(k_loop_dot<LD_LOSS>, g, dim3(RT), 0, s, la, part); break;
        default: hipLaunchKernelGGL(k_loop_dot<LD_MEAN>, g, dim3(RT), 0, s, la, part); break;
        }
        hipLaunchKernelGGL(k_loop_dot_final, dim3((unsigned)nb), dim3(RT), 0, s, (const double *)part, nbl, ne, red);
    };
    // dst = P src (mean removal per field and instance)
    auto center_dev = [&](const double *src, double *dst) {
        la.x0 = src;
        dot(LD_MEAN, 2);
        hipLaunchKernelGGL(k_loop_center, dim3(g2), dim3(RT), 0, s, src, dst, (const double *)red, N, BN);
    };
    // sums of the last dot() on the host
    auto fetch = [&](int ne) -> int {
        HIPX(ctx, hipMemcpyAsync(h_red, red, sizeof(double) * nb * ne, hipMemcpyDeviceToHost, s));
        HIPX(ctx, hipStreamSynchronize(s));
        return SCFT_OK;
    };
    HIPX(ctx, hipMemcpyAsync(wd, w0A, 8 * BN, hipMemcpyHostToDevice, s));
    HIPX(ctx, hipMemcpyAsync(wd + BN, w0B, 8 * BN, hipMemcpyHostToDevice, s));
    HIPX(ctx, hipMemcpyAsync(hd, h, 8 * BN, hipMemcpyHostToDevice, s));
    HIPX(ctx, hipMemcpyAsync(tg, target, 8 * BN, hipMemcpyHostToDevice, s));
    // Anderson per update k, per instance: mp, U, solve status (coefficients in h_coef)
    std::vector<int> mpk(K, 0);
    std::vector<std::vector<double>> Uk(K);
    std::vector<std::vector<char>> okk(K);
    // ---- forward ----
    for (int k = 0; k < K; ++k) {
        double *wtk = WT + (size_t)k * S, *phk = PH + (size_t)k * S;
        center_dev(wd, wtk);
        hipLaunchKernelGGL(k_loop_u, dim3(g2), dim3(RT), 0, s, (const double *)wtk, (const double *)hd, dev, BN);
        HIPX(ctx, enqueue_conc_scratch(ctx));
        HIPX(ctx, hipMemcpyAsync(phk, dev + 6 * BN, 8 * S, hipMemcpyDeviceToDevice, s));
        HIPX(ctx, hipMemcpyAsync(Qk + (size_t)k * nb, ctx->Q, 8 * nb, hipMemcpyDeviceToDevice, s));
        if (keep_ck) HIPX(ctx, hipMemcpyAsync(CK + (size_t)k * ck_rec, ctx->slots, ck_bytes, hipMemcpyDeviceToDevice, s));
        if (keep_open)
            HIPX(ctx, hipMemcpyAsync(CK + (size_t)k * ck_rec + ck_bytes, open_slots, ck_bytes, hipMemcpyDeviceToDevice, s));
        if (k == K - 1) break;
        double *ddk = DD + (size_t)k * S;
        hipLaunchKernelGGL(k_loop_resid, dim3(g1), dim3(RT), 0, s, (const double *)wtk, (const double *)phk, ddk, chiN,
                           wcmp, BN);
        const int mp = anderson ? std::min(k, m) : 0;
        mpk[k] = mp;
        okk[k].assign(nb, 0);
        double *hc = h_coef + (size_t)k * nb * AND_MAX;
        if (mp > 0) {
            // U_ij = <d_k - d_{k-i}, d_k - d_{k-j}>, V_i = <d_k - d_{k-i}, d_k>
            const int ne = mp + mp * (mp + 1) / 2;
            la.x0 = ddk;
            la.dd = DD;
            la.k = k;
            la.mp = mp;
            dot(LD_GRAM, ne);
            if ((rc = fetch(ne))) return rc;
            Uk[k].assign((size_t)nb * mp * mp, 0.0);
            for (int b = 0; b < nb; ++b) {
                const double *t = h_red + (size_t)b * ne;
                std::vector<double> U(mp * mp), V(t, t + mp);
                for (int i = 0, e = mp; i < mp; ++i)
                    for (int j = i; j < mp; ++j, ++e) U[i * mp + j] = U[j * mp + i] = t[e];
                if (ridge > 0.0) {  // U += ridge * tr(U)/mp * I, as k_and_solve
                    double tr = 0.0;
                    for (int i = 0; i < mp; ++i) tr += U[i * mp + i];
                    for (int i = 0; i < mp; ++i) U[i * mp + i] += ridge * tr / mp;
                }
                std::vector<double> x = V;
                const bool ok = host_solve(mp, U, x);
                okk[k][b] = ok;
                for (int i = 0; i < mp; ++i) hc[(size_t)b * AND_MAX + i] = ok ? x[i] : 0.0;
                std::copy(U.begin(), U.end(), Uk[k].begin() + (size_t)b * mp * mp);
            }
            HIPX(ctx, hipMemcpyAsync(coefd + (size_t)k * nb * AND_MAX, hc, sizeof(double) * nb * AND_MAX,
                                     hipMemcpyHostToDevice, s));
        }
        hipLaunchKernelGGL(k_loop_mix, dim3(g2), dim3(RT), 0, s, (const double *)WT, (const double *)DD, (long)k, mp,
                           (const double *)(coefd + (size_t)k * nb * AND_MAX), lam, wd, N, BN);
    }
    HIPX(ctx, hipGetLastError());
    la.x0 = PH + (size_t)(K - 1) * S;
    la.tgt = tg;
    dot(LD_LOSS, 1);
    if ((rc = fetch(1))) return rc;
    HIPX(ctx, hipMemcpyAsync(h_q, Qk, 8 * n_q, hipMemcpyDeviceToHost, s));
    HIPX(ctx, hipStreamSynchronize(s));
    for (size_t i = 0; i < n_q; ++i)
        if (!std::isfinite(h_q[i]) || !(h_q[i] > 0.0)) {
            ctx->have_conc = 0;
            return set_err(ctx, SCFT_E_NONFINITE, "Q not finite and > 0 in the adjoint's forward loop (iteration " +
                                                      std::to_string(i / nb) + ")");
        }
    if (loss)
        for (int b = 0; b < nb; ++b) loss[b] = h_red[b];
    if (phiA_last) {
        HIPX(ctx, hipMemcpyAsync(phiA_last, PH + (size_t)(K - 1) * S, 8 * BN, hipMemcpyDeviceToHost, s));
        HIPX(ctx, hipStreamSynchronize(s));
    }
    if (!grad) {
        ctx->have_conc = 0;
        return SCFT_OK;
    }
    // ---- reverse ----
    zero(WTB, n_traj);
    zero(DB, n_traj);
    zero(hbar, BN);
    zero(wbar, S);
    double *pbar = dev + 2 * BN, *ubar = dev + 4 * BN;
    for (int k = K - 1; k >= 0; --k) {
        double *wtbk = WTB + (size_t)k * S, *dbk = DB + (size_t)k * S;
        if (k == K - 1) {
            hipLaunchKernelGGL(k_loop_pbar0, dim3(g2), dim3(RT), 0, s, (const double *)(PH + (size_t)k * S),
                               (const double *)tg, pbar, BN);
        } else {
            const int mp = mpk[k];
            const double *cd = coefd + (size_t)k * nb * AND_MAX;
            hipLaunchKernelGGL(k_loop_mix_adj, dim3(g2), dim3(RT), 0, s, (const double *)wbar, WTB, DB, (long)k, mp, cd,
                               lam, N, BN);
            if (mp > 0) {
                // cbar_i = <wbar, (wt_{k-i} - wt_k) + lam (d_{k-i} - d_k)>
                la.x0 = wbar;
                la.wt = WT;
                la.dd = DD;
                la.k = k;
                la.mp = mp;
                la.lam = lam;
                dot(LD_CBAR, mp);
                if ((rc = fetch(mp))) return rc;
                const double *hc = h_coef + (size_t)k * nb * AND_MAX;
                double *hr = h_rv + (size_t)k * nb * (AND_MAX + 1);
                for (int b = 0; b < nb; ++b) {
                    if (!okk[k][b]) continue;  // c = 0: no least-squares term
                    std::vector<double> U(Uk[k].begin() + (size_t)b * mp * mp, Uk[k].begin() + (size_t)(b + 1) * mp * mp);
                    std::vector<double> Vb(h_red + (size_t)b * mp, h_red + (size_t)(b + 1) * mp);
                    host_solve(mp, U, Vb);  // U symmetric: Vbar = U^-1 cbar
                    // the ridge ridge*tr(U)/mp*I adds -(ridge/mp)(Vbar.c) I to Ubar
                    double rdiag = 0.0;
                    if (ridge > 0.0)
                        for (int i = 0; i < mp; ++i) rdiag -= ridge / mp * Vb[i] * hc[(size_t)b * AND_MAX + i];
                    for (int i = 0; i < mp; ++i) hr[(size_t)b * (AND_MAX + 1) + i] = Vb[i];
                    hr[(size_t)b * (AND_MAX + 1) + AND_MAX] = rdiag;
                }
                double *rd = rvd + (size_t)k * nb * (AND_MAX + 1);
                HIPX(ctx, hipMemcpyAsync(rd, hr, sizeof(double) * nb * (AND_MAX + 1), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_loop_dbar, dim3(g2), dim3(RT), 0, s, (const double *)DD, DB, (long)k, mp, cd,
                                   (const double *)rd, N, BN);
            }
            // residual d_k = R(wt_k, phi_k): linear adjoint
            hipLaunchKernelGGL(k_loop_resid_adj, dim3(g1), dim3(RT), 0, s, (const double *)dbk, wtbk, pbar, chiN, wcmp,
                               BN);
        }
        // chain adjoint: lam = pbar at u_k = (wt_k,A + h, wt_k,B)
        hipLaunchKernelGGL(k_loop_u, dim3(g2), dim3(RT), 0, s, (const double *)(WT + (size_t)k * S), (const double *)hd,
                           dev, BN);
        if (keep_ck) {  // the forward's W factors, checkpoints (and opened states), Q and phi of u_k
            HIPX(ctx, ctx->fp32 ? enqueue_prep_scratch<float>(ctx) : enqueue_prep_scratch<double>(ctx));
            HIPX(ctx, hipMemcpyAsync(ctx->slots, CK + (size_t)k * ck_rec, ck_bytes, hipMemcpyDeviceToDevice, s));
            if (keep_open)
                HIPX(ctx, hipMemcpyAsync(open_slots, CK + (size_t)k * ck_rec + ck_bytes, ck_bytes, hipMemcpyDeviceToDevice,
                                         s));
            HIPX(ctx, hipMemcpyAsync(ctx->Q, Qk + (size_t)k * nb, 8 * nb, hipMemcpyDeviceToDevice, s));
            HIPX(ctx, hipMemcpyAsync(dev + 6 * BN, PH + (size_t)k * S, 8 * S, hipMemcpyDeviceToDevice, s));
        }
        HIPX(ctx, ctx->fp32 ? enqueue_conc_adjoint<float>(ctx, 0.0, !keep_ck)
                            : enqueue_conc_adjoint<double>(ctx, 0.0, !keep_ck));
        hipLaunchKernelGGL(k_loop_ubar, dim3(g2), dim3(RT), 0, s, (const double *)ubar, hbar, wtbk, BN);
        // mean removal P is symmetric: wbar(k-1) = P Wtbar(k)
        if (k > 0) center_dev(wtbk, wbar);
    }
    HIPX(ctx, hipGetLastError());
    HIPX(ctx, hipMemcpyAsync(grad_h, hbar, 8 * BN, hipMemcpyDeviceToHost, s));
    HIPX(ctx, hipStreamSynchronize(s));
    ctx->have_conc = 0;
    return SCFT_OK;
}

// ===========================================================================
// Heterogeneous parameter sweep (no reference counterpart: the reference's
// batch shares chiN, f, box and grid, struct.h:42-68).  Instances are grouped
// by grid size; each group is one context with per-instance chiN, fA (hence
// NsA, NsB, Ns) and box length, all groups on one stream, and one SCFT
// iteration of every group is captured as a single HIP graph.
// ===========================================================================
struct scft_sweep {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<scft_ctx *> groups;
    std::vector<int> grp_of, idx_in;  // per input instance
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

static int sweep_err(scft_sweep *s, int code, const std::string &m)
{
    if (s) s->err = m;
    else g_create_error = m;
    return code;
}

static void sweep_free(scft_sweep *s)
{
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
    for (scft_ctx *c : s->groups) scft_destroy(c);
    s->groups.clear();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

extern "C" int scft_sweep_create(const scft_params *base, const scft_instance *inst, int count, int device,
                                 scft_sweep **out)
{
    if (!out) return sweep_err(nullptr, SCFT_E_INVALID, "scft_sweep_create: out is NULL");
    *out = nullptr;
    if (!base || !inst || count < 1) return sweep_err(nullptr, SCFT_E_INVALID, "scft_sweep_create: empty sweep");
    std::unique_ptr<scft_sweep> s(new scft_sweep());
    s->device = device;
    {
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&s->ev0);
        if (e == hipSuccess) e = hipEventCreate(&s->ev1);
        if (e != hipSuccess) {
            sweep_free(s.get());
            return sweep_err(nullptr, SCFT_E_HIP, std::string("sweep stream/events: ") + hipGetErrorString(e));
        }
    }
    std::vector<int> sizes;
    for (int i = 0; i < count; ++i)
        if (std::find(sizes.begin(), sizes.end(), inst[i].n) == sizes.end()) sizes.push_back(inst[i].n);
    std::sort(sizes.begin(), sizes.end());
    s->grp_of.assign(count, 0);
    s->idx_in.assign(count, 0);
    for (size_t g = 0; g < sizes.size(); ++g) {
        std::vector<double> L, chi, f;
        for (int i = 0; i < count; ++i)
            if (inst[i].n == sizes[g]) {
                s->grp_of[i] = (int)g;
                s->idx_in[i] = (int)L.size();
                L.push_back(inst[i].L);
                chi.push_back(inst[i].chiN);
                f.push_back(inst[i].fA);
            }
        scft_params p = *base;
        p.nx = p.ny = p.nz = sizes[g];
        p.lx = p.ly = p.lz = L[0];
        p.chiN = chi[0];
        p.fA = f[0];
        p.batch = (int)L.size();
        p.checkpoint = 0;
        scft_ctx *c = nullptr;
        int rc = create_ctx(&p, L.data(), chi.data(), f.data(), device, s->stream, &c);
        if (rc) {
            const std::string m = "size group " + std::to_string(sizes[g]) + ": " + g_create_error;
            sweep_free(s.get());
            return sweep_err(nullptr, rc, m);
        }
        s->groups.push_back(c);
    }
    *out = s.release();
    return SCFT_OK;
}

extern "C" int scft_sweep_destroy(scft_sweep *s)
{
    if (!s) return SCFT_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    sweep_free(s);
    delete s;
    return SCFT_OK;
}

extern "C" const char *scft_sweep_last_error(const scft_sweep *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

extern "C" int scft_sweep_set_fields(scft_sweep *s, int i, const double *wA, const double *wB)
{
    if (!s || i < 0 || i >= (int)s->grp_of.size() || !wA || !wB)
        return s ? sweep_err(s, SCFT_E_INVALID, "scft_sweep_set_fields: bad instance or NULL field") : SCFT_E_INVALID;
    scft_ctx *c = s->groups[s->grp_of[i]];
    for (long r = 0; r < c->N; ++r)
        if (!std::isfinite(wA[r]) || !std::isfinite(wB[r]))
            return sweep_err(s, SCFT_E_NONFINITE, "non-finite field value in instance " + std::to_string(i));
    const size_t off = (size_t)s->idx_in[i] * c->N;
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess) e = hipMemcpyAsync(c->wA + off, wA, 8 * c->N, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->wB + off, wB, 8 * c->N, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && c->and_cnt) {
        const long zero = 0;  // new fields restart the group's Anderson history
        e = hipMemcpyAsync(c->and_cnt, &zero, sizeof(long), hipMemcpyHostToDevice, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("scft_sweep_set_fields: ") + hipGetErrorString(e));
    return SCFT_OK;
}

extern "C" int scft_sweep_get_fields(scft_sweep *s, int i, double *phiA, double *phiB, double *wA, double *wB)
{
    if (!s || i < 0 || i >= (int)s->grp_of.size())
        return s ? sweep_err(s, SCFT_E_INVALID, "scft_sweep_get_fields: bad instance") : SCFT_E_INVALID;
    scft_ctx *c = s->groups[s->grp_of[i]];
    const size_t off = (size_t)s->idx_in[i] * c->N;
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess && phiA) e = hipMemcpyAsync(phiA, c->phiA + off, 8 * c->N, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && phiB) e = hipMemcpyAsync(phiB, c->phiB + off, 8 * c->N, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && wA) e = hipMemcpyAsync(wA, c->wA + off, 8 * c->N, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && wB) e = hipMemcpyAsync(wB, c->wB + off, 8 * c->N, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("scft_sweep_get_fields: ") + hipGetErrorString(e));
    return SCFT_OK;
}

static hipError_t sweep_launch(scft_sweep *s)
{
    if (!s->exec) {
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t e1 = hipSuccess;
        for (scft_ctx *c : s->groups)
            if (e1 == hipSuccess) e1 = enqueue_iteration(c);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s->stream, &g);
        if (e1 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return e1;
        }
        if (e != hipSuccess) return e;
        s->graph = g;
        e = hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            s->exec = nullptr;
            return e;
        }
    }
    return hipGraphLaunch(s->exec, s->stream);
}

static int sweep_collect_fe(scft_sweep *s, double *fe)
{
    const int count = (int)s->grp_of.size();
    std::vector<std::vector<double>> gfe(s->groups.size());
    for (size_t g = 0; g < s->groups.size(); ++g) {
        scft_ctx *c = s->groups[g];
        gfe[g].resize(5 * c->batch);
        hipError_t e = hipMemcpyAsync(gfe[g].data(), c->fe, 8 * gfe[g].size(), hipMemcpyDeviceToHost, s->stream);
        if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("free energy copy: ") + hipGetErrorString(e));
        c->have_fe = 1;
        c->have_conc = 0;
    }
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("sweep: ") + hipGetErrorString(e));
    int rc = SCFT_OK;
    for (int i = 0; i < count; ++i) {
        const double *f = gfe[s->grp_of[i]].data() + 5 * s->idx_in[i];
        if (fe) std::memcpy(fe + 5 * i, f, 5 * sizeof(double));
        if (!std::isfinite(f[0]) && rc == SCFT_OK)
            rc = sweep_err(s, SCFT_E_NONFINITE, "non-finite free energy in instance " + std::to_string(i));
    }
    return rc;
}

extern "C" int scft_sweep_iterate(scft_sweep *s, int n_iter, double *fe)
{
    if (!s || n_iter < 0) return SCFT_E_INVALID;
    hipError_t e = hipSetDevice(s->device);
    for (int it = 0; it < n_iter && e == hipSuccess; ++it) e = sweep_launch(s);
    if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("sweep iteration: ") + hipGetErrorString(e));
    if (n_iter == 0) return SCFT_OK;
    return sweep_collect_fe(s, fe);
}

extern "C" int scft_sweep_time_iterations(scft_sweep *s, int n_iter, float *ms)
{
    if (!s || !ms || n_iter < 1) return SCFT_E_INVALID;
    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess && !s->exec) {  // capture outside the timed region
        e = sweep_launch(s);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(s->ev0, s->stream);
    for (int it = 0; it < n_iter && e == hipSuccess; ++it) e = sweep_launch(s);
    if (e == hipSuccess) e = hipEventRecord(s->ev1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(s->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, s->ev0, s->ev1);
    if (e != hipSuccess) return sweep_err(s, SCFT_E_HIP, std::string("sweep timing: ") + hipGetErrorString(e));
    return sweep_collect_fe(s, nullptr);
}

extern "C" int scft_sweep_autotune(scft_sweep *s, int reps)
{
    if (!s || reps < 1) return SCFT_E_INVALID;
    for (scft_ctx *c : s->groups) {
        int rc = scft_autotune(c, reps, nullptr, nullptr);
        if (rc) return sweep_err(s, rc, std::string("autotune: ") + c->err);
    }
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
    return SCFT_OK;
}

extern "C" int scft_sweep_size(const scft_sweep *s, int i, long *N, int *NsA, int *NsB)
{
    if (!s || i < 0 || i >= (int)s->grp_of.size()) return SCFT_E_INVALID;
    const scft_ctx *c = s->groups[s->grp_of[i]];
    const int2 ns = c->nsab_h[s->idx_in[i]];
    if (N) *N = c->N;
    if (NsA) *NsA = ns.x;
    if (NsB) *NsB = ns.y;
    return SCFT_OK;
}

// ===========================================================================
// scft_propagate_noalloc: scft_propagate on caller-owned device arrays with
// every work buffer in one file-scope device array, so that it allocates
// nothing (no reference counterpart; for callers that cannot allocate).
// ===========================================================================
namespace {
constexpr size_t NA_MAX_POINTS = (size_t)1 << 18;  // batch * N limit (64^3)
// tw + kf of 3 axes (with their 256-byte alignment), then one post factor per instance
constexpr size_t NA_TABLE_BYTES = 3 * (size_t)GEN_MAX * (16 + 8) + 8192;
// W [4][batch][N] + two march slices + the work field: 80 bytes per point in
// FP64, plus the 256-byte alignment of the four regions
constexpr size_t NA_SCRATCH_BYTES = NA_TABLE_BYTES + 80 * NA_MAX_POINTS + 1024;
constexpr int NA_MAX_DEVICES = 64;

// Every GPU holds its own copy of this array from the library's load until
// the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (scratch_on).
static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[NA_MAX_DEVICES] = {};

// device address of `scratch` on the current device (looked up once per device)
int scratch_on(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_err(nullptr, SCFT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[device] = (unsigned char *)ptr;
    }
    *out = g_scratch_dev[device];
    return SCFT_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

// real parts of a slice into out[b * inst_stride + r] (q_out's [batch][ns+1][N])
template <typename T>
__global__ void __launch_bounds__(RT) k_take_real_to(const typename cplx<T>::type *__restrict__ slot,
                                                     double *__restrict__ out, long n, long N, long inst_stride)
{
    const T *z = reinterpret_cast<const T *>(slot);  // planar: the real parts come first
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)gridDim.x * RT)
        out[(i / N) * inst_stride + i % N] = (double)z[i];
}

template <typename T>
static int propagate_noalloc(scft_ctx &c, unsigned char *scr, int block, int sign, const double *w,
                             const double *q_init, double *q_out)
{
    using C = typename cplx<T>::type;
    const size_t BN = (size_t)c.batch * c.N;
    const bool dbl_tab = !c.fp32 || c.mixed;  // table entries in the arithmetic precision
    const size_t ts = dbl_tab ? 8 : 4;
    // tables: tw (n complex) and kf (n reals) per active axis, then the post factors
    std::vector<unsigned char> host(NA_TABLE_BYTES, 0);
    size_t off = 0;
    std::vector<double> k2;
    double fold = 1.0;
    for (int a = 0; a < 3; ++a) {
        ksq_axis(c.n[a], a == 0 ? c.p.lx : (a == 1 ? c.p.ly : c.p.lz), c.p.pi, k2);
        if (!c.active[a]) {
            fold *= std::exp(-c.p.ds * k2[0]);
            continue;
        }
        const int n = c.n[a];
        c.tw[a] = scr + off;
        for (int k = 0; k < n; ++k) {
            const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
            const double re = (double)cosl(ang), im = (double)-sinl(ang);
            if (dbl_tab) {
                std::memcpy(&host[off + 16 * (size_t)k], &re, 8);
                std::memcpy(&host[off + 16 * (size_t)k + 8], &im, 8);
            } else {
                const float fr = (float)re, fi = (float)im;
                std::memcpy(&host[off + 8 * (size_t)k], &fr, 4);
                std::memcpy(&host[off + 8 * (size_t)k + 4], &fi, 4);
            }
        }
        off = align256(off + 2 * ts * n);
        c.kf[a] = scr + off;
        for (int k = 0; k < n; ++k) {
            const double v = std::exp(-c.p.ds * k2[k]);
            if (dbl_tab) std::memcpy(&host[off + 8 * (size_t)k], &v, 8);
            else {
                const float f = (float)v;
                std::memcpy(&host[off + 4 * (size_t)k], &f, 4);
            }
        }
        off = align256(off + ts * n);
    }
    c.post = fold / (double)c.N;
    c.post_b = (double *)(scr + off);
    if (off + 8 * (size_t)c.batch > NA_TABLE_BYTES)
        return set_err(nullptr, SCFT_E_INVALID, "scft_propagate_noalloc: batch above " +
                                                    std::to_string((NA_TABLE_BYTES - off) / 8) + " instances");
    for (int b = 0; b < c.batch; ++b) std::memcpy(&host[off + 8 * (size_t)b], &c.post, 8);
    off = align256(off + 8 * (size_t)c.batch);
    // the tables reach the device before any kernel of this call reads them
    hipError_t e = hipMemcpy(scr, host.data(), off, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("table upload: ") + hipGetErrorString(e));
    size_t p = NA_TABLE_BYTES;
    c.wdz = scr + p;
    p = align256(p + 4 * BN * sizeof(T));
    C *sl[2] = {(C *)(scr + p), nullptr};
    p = align256(p + BN * sizeof(C));
    sl[1] = (C *)(scr + p);
    p = align256(p + BN * sizeof(C));
    c.work = scr + p;
    p = align256(p + BN * sizeof(C));
    if (p > NA_SCRATCH_BYTES) return set_err(nullptr, SCFT_E_INVALID, "no-allocation propagate: scratch overflow");
    c.slots = sl[0];
    const int ns = block == 0 ? c.NsA : c.NsB;
    const long stride = (long)(ns + 1) * c.N;
    T *wpre = (T *)c.wdz + (size_t)block * BN, *wpost = (T *)c.wdz + (size_t)(2 + block) * BN;
    if (c.mixed)
        hipLaunchKernelGGL((k_wdz_one<T, true>), dim3(pt_grid(BN)), dim3(RT), 0, c.stream, w, (long)BN, c.N,
                           c.p.ds / 2, (const double *)c.post_b, wpre, wpost);
    else
        hipLaunchKernelGGL((k_wdz_one<T, false>), dim3(pt_grid(BN)), dim3(RT), 0, c.stream, w, (long)BN, c.N,
                           c.p.ds / 2, (const double *)c.post_b, wpre, wpost);
    hipLaunchKernelGGL(k_seed_real<T>, dim3(pt_grid(BN)), dim3(RT), 0, c.stream, sl[0], q_init, (long)BN);
    hipLaunchKernelGGL(k_take_real_to<T>, dim3(pt_grid(BN)), dim3(RT), 0, c.stream, sl[0],
                       q_out + (sign == 1 ? 0 : (long)ns * c.N), (long)BN, c.N, stride);
    e = hipGetLastError();
    for (int t = 1; t <= ns && e == hipSuccess; ++t) {
        e = enqueue_step<T>(&c, sl[(t - 1) & 1], sl[t & 1], 0, block, block);
        if (e != hipSuccess) break;
        const int s = sign == 1 ? t : ns - t;
        hipLaunchKernelGGL(k_take_real_to<T>, dim3(pt_grid(BN)), dim3(RT), 0, c.stream, sl[t & 1],
                           q_out + (long)s * c.N, (long)BN, c.N, stride);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("no-allocation propagate: ") + hipGetErrorString(e));
    return SCFT_OK;
}

extern "C" int scft_propagate_noalloc(const scft_params *pin, int device, int block, int sign, const double *w,
                                      const double *q_init, double *q_out, void *stream)
{
    if (!pin) return set_err(nullptr, SCFT_E_INVALID, "scft_propagate_noalloc: params is NULL");
    if (!w || !q_out) return set_err(nullptr, SCFT_E_INVALID, "scft_propagate_noalloc: w or q_out is NULL");
    if (block != 0 && block != 1) return set_err(nullptr, SCFT_E_INVALID, "block must be 0 (A) or 1 (B)");
    if (sign != 1 && sign != -1) return set_err(nullptr, SCFT_E_INVALID, "sign must be +1 or -1");
    scft_params p = *pin;
    if (p.pi == 0.0) p.pi = SCFT_REFERENCE_PI;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1 || (p.nx == 1 && p.ny == 1 && p.nz == 1))
        return set_err(nullptr, SCFT_E_INVALID, "grid sizes must be >= 1 with a length > 1");
    if (!(p.lx > 0 && p.ly > 0 && p.lz > 0) || !std::isfinite(p.lx) || !std::isfinite(p.ly) || !std::isfinite(p.lz))
        return set_err(nullptr, SCFT_E_INVALID, "box lengths must be finite and > 0");
    if (!(p.ds > 0.0) || !(p.ds < 0.5)) return set_err(nullptr, SCFT_E_INVALID, "ds must be in (0, 0.5)");
    if (!(p.fA > 0.0 && p.fA < 1.0)) return set_err(nullptr, SCFT_E_INVALID, "fA must be in (0, 1)");
    if (p.batch < 1) return set_err(nullptr, SCFT_E_INVALID, "batch must be >= 1");
    if (p.precision != SCFT_FP64 && p.precision != SCFT_FP32 && p.precision != SCFT_FP32X)
        return set_err(nullptr, SCFT_E_INVALID, "precision must be SCFT_FP64, SCFT_FP32 or SCFT_FP32X");
    const int n[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; ++a)
        if (n[a] > 1 && ((n[a] & 1) || !pass_supported(n[a])))
            return set_err(nullptr, SCFT_E_UNSUPPORTED, "grid length " + std::to_string(n[a]) + " has no kernel");
    const size_t BN = (size_t)p.batch * p.nx * p.ny * p.nz;
    if (BN > NA_MAX_POINTS)
        return set_err(nullptr, SCFT_E_INVALID,
                       "scft_propagate_noalloc: batch * N = " + std::to_string(BN) + " exceeds its scratch limit of " +
                           std::to_string(NA_MAX_POINTS) + " points; use scft_propagate");
    int NsA, NsB;
    scft_contour_counts(p.fA, p.ds, &NsA, &NsB);
    if (NsA < 1 || NsB < 1) return set_err(nullptr, SCFT_E_INVALID, "both blocks need at least one contour step");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(nullptr, SCFT_E_HIP, "no HIP device");
    if (device < 0 || device >= ndev || device >= NA_MAX_DEVICES)
        return set_err(nullptr, SCFT_E_INVALID, "device index out of range");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    unsigned char *scr = nullptr;
    int rc = scratch_on(device, &scr);
    if (rc) return rc;
    scft_ctx c;  // a view over the scratch array: owns no device memory
    c.p = p;
    c.device = device;
    c.stream = (hipStream_t)stream;
    c.own_stream = false;
    c.NsA = NsA;
    c.NsB = NsB;
    c.Ns = NsA + NsB;
    c.Ns_max = c.Ns;
    c.N = (long)p.nx * p.ny * p.nz;
    c.batch = p.batch;
    c.fp32 = p.precision != SCFT_FP64;
    c.mixed = p.precision == SCFT_FP32X;
    c.csize = c.fp32 ? 8 : 16;
    c.tsize = c.fp32 ? 4 : 8;
    for (int a = 0; a < 3; ++a) {
        c.n[a] = n[a];
        c.active[a] = n[a] > 1;
        if (c.active[a]) {
            if (c.first_axis < 0) c.first_axis = a;
            c.last_axis = a;
        }
    }
    rc = c.fp32 ? propagate_noalloc<float>(c, scr, block, sign, w, q_init, q_out)
                : propagate_noalloc<double>(c, scr, block, sign, w, q_init, q_out);
    if (rc) return rc;
    e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return set_err(nullptr, SCFT_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return SCFT_OK;
}
