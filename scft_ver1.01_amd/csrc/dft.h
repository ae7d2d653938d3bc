// dft.h -- in-register discrete Fourier transforms for gfx950 kernels.
//
// Dft<R, S>::run(v) transforms R complex values held in registers, in place,
// into natural order:  X[k] = sum_j v[j] * exp(S * 2*pi*i*j*k/R)  (S = -1
// forward, +1 inverse, unnormalised).  Radices 2, 3, 4, 5 are written out;
// composite R is split Cooley-Tukey style R = P*Q with compile-time twiddles,
// so every index is a constant and the arrays stay in VGPRs.
#pragma once

#include <hip/hip_runtime.h>

namespace scft {

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };

template <typename C> __device__ __forceinline__ C c_add(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C c_sub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C c_mul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C c_mulc(C a, C b)
{
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C, typename T> __device__ __forceinline__ C c_scale(C a, T s) { return C{a.x * s, a.y * s}; }
// multiply by S*i  (S = +-1)
template <int S, typename C> __device__ __forceinline__ C c_muli(C a)
{
    return S > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}

// ---------------------------------------------------------------------------
// compile-time cos/sin of 2*pi*k/n: integer octant reduction, then a Taylor
// series on [0, pi/4] (1-2 ulp).  Multiples of pi/4 come out exact.
// ---------------------------------------------------------------------------
constexpr double kQuarterPi = 0.78539816339744830961566084581987572;
constexpr double kSqrtHalf = 0.70710678118654752440084436210484904;

constexpr double poly_sin(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 12; ++i) {
        term *= -x2 / double((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double poly_cos(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 12; ++i) {
        term *= -x2 / double((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}

struct CS { double c, s; };

// angle = 2*pi*k/n
constexpr CS cossin2pi(long k, long n)
{
    k %= n;
    if (k < 0) k += n;
    const long u = 8 * k;
    const long o = u / n;     // octant 0..7
    const long rr = u - o * n;
    if (rr == 0) {
        switch (o) {
        case 0: return CS{1.0, 0.0};
        case 1: return CS{kSqrtHalf, kSqrtHalf};
        case 2: return CS{0.0, 1.0};
        case 3: return CS{-kSqrtHalf, kSqrtHalf};
        case 4: return CS{-1.0, 0.0};
        case 5: return CS{-kSqrtHalf, -kSqrtHalf};
        case 6: return CS{0.0, -1.0};
        default: return CS{kSqrtHalf, -kSqrtHalf};
        }
    }
    if ((o & 1) == 0) {
        const double ph = kQuarterPi * double(rr) / double(n);
        const double c = poly_cos(ph), s = poly_sin(ph);
        switch (o) {
        case 0: return CS{c, s};
        case 2: return CS{-s, c};
        case 4: return CS{-c, -s};
        default: return CS{s, -c};
        }
    }
    const double ph = kQuarterPi * double(n - rr) / double(n);
    const double c = poly_cos(ph), s = poly_sin(ph);
    switch (o) {
    case 1: return CS{s, c};
    case 3: return CS{-c, s};
    case 5: return CS{-s, -c};
    default: return CS{c, -s};
    }
}

// v *= exp(S * 2*pi*i*E/R), specialised for the trivial angles
template <int E, int R, int S, typename C>
__device__ __forceinline__ C tw_const(C v)
{
    using T = decltype(v.x);
    constexpr int e = ((E % R) + R) % R;
    if constexpr (e == 0) {
        return v;
    } else if constexpr ((4 * e) % R == 0) {
        constexpr int q = (4 * e) / R;  // quarter turns
        if constexpr (q == 1) return c_muli<S>(v);
        else if constexpr (q == 2) return C{-v.x, -v.y};
        else return c_muli<-S>(v);
    } else if constexpr ((8 * e) % R == 0) {
        constexpr int o = (8 * e) / R;  // odd octant 1,3,5,7
        constexpr CS cs = cossin2pi(o, 8);
        constexpr T h = T(kSqrtHalf);
        constexpr int sc = cs.c > 0 ? 1 : -1;
        constexpr int ss = (cs.s > 0 ? 1 : -1) * S;
        // (x + i y)(sc + i ss) * h
        const T xr = sc * v.x - ss * v.y;
        const T xi = ss * v.x + sc * v.y;
        return C{xr * h, xi * h};
    } else {
        constexpr CS cs = cossin2pi(e, R);
        constexpr T c = T(cs.c);
        constexpr T s = T(S * cs.s);
        return C{v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

template <int R> struct first_factor {
    static constexpr int value = (R % 4 == 0 && R != 4) ? 4
                                 : (R % 2 == 0) ? 2
                                 : (R % 3 == 0) ? 3
                                 : (R % 5 == 0) ? 5
                                 : R;
};

template <int R, int S> struct Dft;

template <int S> struct Dft<1, S> {
    template <typename C> __device__ __forceinline__ static void run(C *) {}
};

template <int S> struct Dft<2, S> {
    template <typename C> __device__ __forceinline__ static void run(C *v)
    {
        const C a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    }
};

template <int S> struct Dft<3, S> {
    template <typename C> __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        constexpr T s3 = T(0.86602540378443864676372317075293618);  // sin(2 pi/3)
        const C t1 = c_add(v[1], v[2]);
        const C t2 = C{v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
        const C d = c_sub(v[1], v[2]);
        const C t3 = c_muli<S>(C{d.x * s3, d.y * s3});
        v[0] = c_add(v[0], t1);
        v[1] = c_add(t2, t3);
        v[2] = c_sub(t2, t3);
    }
};

template <int S> struct Dft<4, S> {
    template <typename C> __device__ __forceinline__ static void run(C *v)
    {
        const C s0 = c_add(v[0], v[2]), d0 = c_sub(v[0], v[2]);
        const C s1 = c_add(v[1], v[3]), d1 = c_sub(v[1], v[3]);
        const C j = c_muli<S>(d1);
        v[0] = c_add(s0, s1);
        v[2] = c_sub(s0, s1);
        v[1] = c_add(d0, j);
        v[3] = c_sub(d0, j);
    }
};

template <int S> struct Dft<5, S> {
    template <typename C> __device__ __forceinline__ static void run(C *v)
    {
        using T = decltype(v[0].x);
        constexpr T c1 = T(0.30901699437494742410229341718281906);   // cos(2 pi/5)
        constexpr T c2 = T(-0.80901699437494742410229341718281906);  // cos(4 pi/5)
        constexpr T s1 = T(0.95105651629515357211643933337938214);   // sin(2 pi/5)
        constexpr T s2 = T(0.58778525229247312916870595463907277);   // sin(4 pi/5)
        const C a1 = c_add(v[1], v[4]), b1 = c_sub(v[1], v[4]);
        const C a2 = c_add(v[2], v[3]), b2 = c_sub(v[2], v[3]);
        const C r1 = C{v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        const C r2 = C{v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        const C i1 = c_muli<S>(C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
        const C i2 = c_muli<S>(C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
        v[0] = C{v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
        v[1] = c_add(r1, i1);
        v[4] = c_sub(r1, i1);
        v[2] = c_add(r2, i2);
        v[3] = c_sub(r2, i2);
    }
};

// composite R = P*Q:  j = p + P*q,  k = k2 + Q*k1
//   X[k2 + Q k1] = sum_p w_P^{p k1} [ w_R^{p k2} * sum_q v[p + P q] w_Q^{q k2} ]
template <int R, int S> struct Dft {
    static constexpr int P = first_factor<R>::value;
    static constexpr int Q = R / P;
    static_assert(P > 1 && P <= 5 && R % P == 0, "radix must factor into 2, 3, 4, 5");

    template <int p, typename C> __device__ __forceinline__ static void inner(C *v, C (&y)[R])
    {
        C t[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) t[q] = v[p + P * q];
        Dft<Q, S>::run(t);
        twid<p, 0>(t, y);
    }
    template <int p, int k2, typename C> __device__ __forceinline__ static void twid(C (&t)[Q], C (&y)[R])
    {
        y[p * Q + k2] = tw_const<p * k2, R, S>(t[k2]);
        if constexpr (k2 + 1 < Q) twid<p, k2 + 1>(t, y);
    }
    template <int p, typename C> __device__ __forceinline__ static void all_inner(C *v, C (&y)[R])
    {
        inner<p>(v, y);
        if constexpr (p + 1 < P) all_inner<p + 1>(v, y);
    }

    template <typename C> __device__ __forceinline__ static void run(C *v)
    {
        C y[R];  // y[p*Q + k2]
        all_inner<0>(v, y);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) {
            C t[P];
#pragma unroll
            for (int p = 0; p < P; ++p) t[p] = y[p * Q + k2];
            Dft<P, S>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < P; ++k1) v[k2 + Q * k1] = t[k1];
        }
    }
};

}  // namespace scft
