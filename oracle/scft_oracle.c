/*
 * scft_oracle.c -- CPU FP64 restatement of the reference SCFT method.
 *
 * TEST INFRASTRUCTURE ONLY: the parity reference of the tests and the timed CPU baseline.
 * The product never links or calls this file.
 *
 * Function-by-function correspondence with the reference (ren5885a/SCFT_Ver1.01):
 *   orc_contour_counts  init_cuda.cu:330-331   (truncating casts, +1e-8)
 *   orc_ksq_axis        init_cuda.cu:408-418   (k tables with the truncated Pi of struct.h:11)
 *   orc_ktable          init_cuda.cu:421-428   (kxyzdz = exp(-ds0*ksq))
 *   orc_step            cuda_aid.cu:227-241 (in = q*wdz), cuFFT D2Z, cuda_aid.cu:306-328
 *                       (spectrum *= kxyzdz), cuFFT Z2D, cuda_aid.cu:346-361 (q = in*wdz/N)
 *   orc_march           cuda_scft.cu:435-563   (sovDifFft: wdz = exp(-w*ds2), seed slice,
 *                                               forward iz=1..ns / backward iz=ns-1..0)
 *   orc_minus_average   cuda_scft.cu:14-59, cuda_aid.cuh:36-78, cuda_aid.cu:365-380
 *   orc_conc            cuda_scft.cu:61-138    (getConc: 4 marches, junction hand-offs
 *                                               cuda_aid.cu:170-208, Q cuda_aid.cu:119-168
 *                                               and cuda_scft.cu:123-126, phi cuda_aid.cu:61-118)
 *   orc_phi_w           cuda_aid.cu:382-416
 *   orc_free_energy     cuda_scft.cu:199-220
 *   orc_iterate         cuda_scft.cu:167-182
 * The FFT (orc_fft1d/orc_fft3d) is this file's own mixed-radix Stockham
 * transform standing in for cuFFT's D2Z/Z2D: a complex transform of the real
 * input gives the same spectrum, and the k-factor is real and even in k, so
 * multiplying the full spectrum and taking the real part of the inverse is the
 * same operator as the reference's half-spectrum D2Z/Z2D pair.
 */
#include "scft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFT: self-sorting Stockham, mixed radix (4, 2, 3, 5, any other prime).    */
/* Stage with radix R and span Ns (product of the radices already applied):  */
/*   for j in [0, n/R): v[r] = x[j + r*n/R] * w_{Ns*R}^{(j mod Ns)*r},        */
/*   v = DFT_R(v), y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns] = v[r].                */
/* Twiddles come from one table w_n^k = exp(sign*2*pi*i*k/n), computed with   */
/* libm cos/sin of the exact angle.                                           */
/* ------------------------------------------------------------------------ */

#define ORC_MAXF 32

typedef struct {
    int n;
    int nf;
    int f[ORC_MAXF];
    double *wr;  /* cos(2 pi k/n) */
    double *wi;  /* sin(2 pi k/n) */
} orc_plan;

static void plan_init(orc_plan *pl, int n)
{
    int m = n, k;
    pl->n = n;
    pl->nf = 0;
    while (m % 4 == 0) { pl->f[pl->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->f[pl->nf++] = 2; m /= 2; }
    for (k = 3; m > 1; k += 2)
        while (m % k == 0) { pl->f[pl->nf++] = k; m /= k; }
    pl->wr = (double *)malloc(sizeof(double) * (size_t)n);
    pl->wi = (double *)malloc(sizeof(double) * (size_t)n);
    for (k = 0; k < n; k++) {
        /* reduce the angle so that symmetric entries are exactly symmetric */
        double a = 2.0 * M_PI * (double)k / (double)n;
        pl->wr[k] = cos(a);
        pl->wi[k] = sin(a);
    }
}

static void plan_free(orc_plan *pl)
{
    free(pl->wr);
    free(pl->wi);
}

/* x,y: work arrays of length n (re/im split). Result lands in x. */
static void fft_run(const orc_plan *pl, double *xr, double *xi, double *yr, double *yi, int sign)
{
    const int n = pl->n;
    int Ns = 1, s;
    double *ar = xr, *ai = xi, *br = yr, *bi = yi;
    double vr[64], vi[64], tr[64], ti[64];
    for (s = 0; s < pl->nf; s++) {
        const int R = pl->f[s];
        const int nR = n / R;
        int j;
        for (j = 0; j < nR; j++) {
            const int jm = j % Ns;
            int r;
            for (r = 0; r < R; r++) {
                /* twiddle w_{Ns R}^{jm r} = w_n^{jm r n/(Ns R)} */
                long e = (long)jm * r * (n / (Ns * R));
                e %= n;
                double c = pl->wr[e], sn = sign * pl->wi[e];
                double a = ar[j + r * nR], b = ai[j + r * nR];
                vr[r] = a * c - b * sn;
                vi[r] = a * sn + b * c;
            }
            if (R == 2) {
                tr[0] = vr[0] + vr[1]; ti[0] = vi[0] + vi[1];
                tr[1] = vr[0] - vr[1]; ti[1] = vi[0] - vi[1];
            } else if (R == 4) {
                double s0r = vr[0] + vr[2], s0i = vi[0] + vi[2];
                double d0r = vr[0] - vr[2], d0i = vi[0] - vi[2];
                double s1r = vr[1] + vr[3], s1i = vi[1] + vi[3];
                double d1r = vr[1] - vr[3], d1i = vi[1] - vi[3];
                /* multiply d1 by (sign * i) */
                double jr = -sign * d1i, ji = sign * d1r;
                tr[0] = s0r + s1r; ti[0] = s0i + s1i;
                tr[2] = s0r - s1r; ti[2] = s0i - s1i;
                tr[1] = d0r + jr;  ti[1] = d0i + ji;
                tr[3] = d0r - jr;  ti[3] = d0i - ji;
            } else {
                /* generic prime radix: direct DFT with table twiddles */
                int k, q;
                for (k = 0; k < R; k++) {
                    double sr = 0.0, si = 0.0;
                    for (q = 0; q < R; q++) {
                        long e = ((long)(k * q) % R) * (n / R);
                        double c = pl->wr[e], sn = sign * pl->wi[e];
                        sr += vr[q] * c - vi[q] * sn;
                        si += vr[q] * sn + vi[q] * c;
                    }
                    tr[k] = sr; ti[k] = si;
                }
            }
            {
                const int idxD = (j / Ns) * Ns * R + jm;
                for (r = 0; r < R; r++) {
                    br[idxD + r * Ns] = tr[r];
                    bi[idxD + r * Ns] = ti[r];
                }
            }
        }
        Ns *= R;
        { double *t = ar; ar = br; br = t; t = ai; ai = bi; bi = t; }
    }
    if (ar != xr) {
        memcpy(xr, ar, sizeof(double) * (size_t)n);
        memcpy(xi, ai, sizeof(double) * (size_t)n);
    }
}

static int radix_ok(int n)
{
    int m = n, k;
    if (n < 1) return 0;
    for (k = 2; k <= 61 && m > 1; k++)
        while (m % k == 0) m /= k;
    return m == 1;
}

int orc_fft1d(int n, double *re, double *im, int sign)
{
    orc_plan pl;
    double *yr, *yi;
    if (!radix_ok(n)) return -1;
    if (n == 1) return 0;
    plan_init(&pl, n);
    yr = (double *)malloc(sizeof(double) * (size_t)n);
    yi = (double *)malloc(sizeof(double) * (size_t)n);
    fft_run(&pl, re, im, yr, yi, sign);
    free(yr); free(yi);
    plan_free(&pl);
    return 0;
}

/* FFT along one axis of a [nz][ny][nx] array: `len` points with stride
   `stride`, `nlines` lines whose start offsets are given by (outer, inner). */
static void fft_axis(const orc_plan *pl, double *re, double *im, int len, long stride,
                     long n_outer, long outer_stride, long n_inner, long inner_stride, int sign)
{
    long nl = n_outer * n_inner;
    long L;
    if (len == 1) return;
#pragma omp parallel
    {
        double *xr = (double *)malloc(sizeof(double) * (size_t)len * 4);
        double *xi = xr + len, *yr = xr + 2 * len, *yi = xr + 3 * len;
#pragma omp for schedule(static)
        for (L = 0; L < nl; L++) {
            long o = L / n_inner, in = L % n_inner;
            long base = o * outer_stride + in * inner_stride;
            int t;
            for (t = 0; t < len; t++) { xr[t] = re[base + t * stride]; xi[t] = im[base + t * stride]; }
            fft_run(pl, xr, xi, yr, yi, sign);
            for (t = 0; t < len; t++) { re[base + t * stride] = xr[t]; im[base + t * stride] = xi[t]; }
        }
        free(xr);
    }
}

typedef struct { orc_plan px, py, pz; } orc_plan3;

static void plan3_init(orc_plan3 *p3, int nx, int ny, int nz)
{
    plan_init(&p3->px, nx);
    plan_init(&p3->py, ny);
    plan_init(&p3->pz, nz);
}

static void plan3_free(orc_plan3 *p3)
{
    plan_free(&p3->px);
    plan_free(&p3->py);
    plan_free(&p3->pz);
}

static void fft3_run(const orc_plan3 *p3, int nx, int ny, int nz, double *re, double *im, int sign)
{
    long nxy = (long)nx * ny;
    /* x lines: for each (z, y) */
    fft_axis(&p3->px, re, im, nx, 1, nz, nxy, ny, nx, sign);
    /* y lines: for each (z, x) */
    fft_axis(&p3->py, re, im, ny, nx, nz, nxy, nx, 1, sign);
    /* z lines: for each (y, x) */
    fft_axis(&p3->pz, re, im, nz, nxy, ny, nx, nx, 1, sign);
}

int orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign)
{
    orc_plan3 p3;
    if (!radix_ok(nx) || !radix_ok(ny) || !radix_ok(nz)) return -1;
    plan3_init(&p3, nx, ny, nz);
    fft3_run(&p3, nx, ny, nz, re, im, sign);
    plan3_free(&p3);
    return 0;
}

void orc_dft1d_brute(int n, const double *re, const double *im, double *ore, double *oim, int sign)
{
    int k, j;
    for (k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (j = 0; j < n; j++) {
            long e = ((long)j * k) % n;
            long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
            long double c = cosl(a), s = sign * sinl(a);
            sr += (long double)re[j] * c - (long double)im[j] * s;
            si += (long double)re[j] * s + (long double)im[j] * c;
        }
        ore[k] = (double)sr;
        oim[k] = (double)si;
    }
}

/* ------------------------------------------------------------------------ */
/* constants                                                                 */
/* ------------------------------------------------------------------------ */

/* init_cuda.cu:330-331 (dfB = 1 - fA at init_cuda.cu:328) */
void orc_contour_counts(double fA, double ds, int *NsA, int *NsB)
{
    double dfB = 1 - fA;
    *NsA = ((int)(fA / ds + 1.0e-8));
    *NsB = ((int)(dfB / ds + 1.0e-8));
}

/* init_cuda.cu:408-418, with dx = lx/(double)Nx (init_cuda.cu:394-396).
   For n == 1 the second loop runs once with i - n = -1 (a nonzero wavenumber);
   that is what the reference computes and is kept. */
void orc_ksq_axis(int n, double l, double pi, double *ksq)
{
    double dx = l / (double)n;
    int i;
    for (i = 0; i <= n / 2 - 1; i++) ksq[i] = 2 * pi * i * 1.0 / n / dx;
    for (i = n / 2; i < n; i++) ksq[i] = 2 * pi * (i - n) * 1.0 / dx / n;
    for (i = 0; i < n; i++) ksq[i] *= ksq[i];
}

/* init_cuda.cu:421-428 */
void orc_ktable(const orc_params *p, double *kxyzdz)
{
    double *kx = (double *)malloc(sizeof(double) * (size_t)p->nx);
    double *ky = (double *)malloc(sizeof(double) * (size_t)p->ny);
    double *kz = (double *)malloc(sizeof(double) * (size_t)p->nz);
    int i, j, k;
    orc_ksq_axis(p->nx, p->lx, p->pi, kx);
    orc_ksq_axis(p->ny, p->ly, p->pi, ky);
    orc_ksq_axis(p->nz, p->lz, p->pi, kz);
    for (k = 0; k < p->nz; k++)
        for (j = 0; j < p->ny; j++)
            for (i = 0; i < p->nx; i++) {
                long ijk = (long)((k * p->ny + j) * p->nx + i);
                double ksq = kx[i] + ky[j] + kz[k];
                kxyzdz[ijk] = exp(-p->ds * ksq);
            }
    free(kx); free(ky); free(kz);
}

/* ------------------------------------------------------------------------ */
/* propagator                                                                */
/* ------------------------------------------------------------------------ */

static long npts(const orc_params *p) { return (long)p->nx * p->ny * p->nz; }

static void step_with_plan(const orc_params *p, const orc_plan3 *p3, const double *kxyzdz,
                           const double *wdz, const double *qin, double *qout,
                           double *re, double *im)
{
    const long N = npts(p);
    long r;
    /* initilize_in_go: in = q * wdz (cuda_aid.cu:237) */
    for (r = 0; r < N; r++) { re[r] = qin[r] * wdz[r]; im[r] = 0.0; }
    fft3_run(p3, p->nx, p->ny, p->nz, re, im, -1);          /* cufftExecD2Z */
    for (r = 0; r < N; r++) { re[r] *= kxyzdz[r]; im[r] *= kxyzdz[r]; } /* sufaceField_go */
    fft3_run(p3, p->nx, p->ny, p->nz, re, im, +1);          /* cufftExecZ2D */
    /* in_to_g_go: q = in*wdz/NxNyNz (cuda_aid.cu:358) */
    for (r = 0; r < N; r++) qout[r] = re[r] * wdz[r] / N;
}

void orc_step(const orc_params *p, const double *kxyzdz, const double *wdz,
              const double *qin, double *qout)
{
    const long N = npts(p);
    orc_plan3 p3;
    double *re = (double *)malloc(sizeof(double) * (size_t)N * 2);
    plan3_init(&p3, p->nx, p->ny, p->nz);
    step_with_plan(p, &p3, kxyzdz, wdz, qin, qout, re, re + N);
    plan3_free(&p3);
    free(re);
}

/* sovDifFft (cuda_scft.cu:435-563) for one instance */
static void march_with_plan(const orc_params *p, const orc_plan3 *p3, const double *kxyzdz,
                            const double *w, const double *qinit, int ns, int sign, double *q,
                            double *wdz, double *re, double *im)
{
    const long N = npts(p);
    const double ds2 = p->ds / 2;
    long r;
    int iz;
    /* initilize_wdz (cuda_aid.cu:254) */
    for (r = 0; r < N; r++) wdz[r] = exp(-w[r] * ds2);
    if (sign == 1) {
        memcpy(q, qinit, sizeof(double) * (size_t)N);                 /* initilize_q */
        for (iz = 1; iz <= ns; iz++)
            step_with_plan(p, p3, kxyzdz, wdz, q + (long)(iz - 1) * N, q + (long)iz * N, re, im);
    } else {
        memcpy(q + (long)ns * N, qinit, sizeof(double) * (size_t)N);  /* initilize_q_inverse */
        for (iz = ns - 1; iz >= 0; iz--)                               /* reads slice iz+1 (:528) */
            step_with_plan(p, p3, kxyzdz, wdz, q + (long)(iz + 1) * N, q + (long)iz * N, re, im);
    }
}

void orc_march(const orc_params *p, const double *kxyzdz, const double *w,
               const double *qinit, int ns, int sign, double *q)
{
    const long N = npts(p);
    orc_plan3 p3;
    double *buf = (double *)malloc(sizeof(double) * (size_t)N * 3);
    plan3_init(&p3, p->nx, p->ny, p->nz);
    march_with_plan(p, &p3, kxyzdz, w, qinit, ns, sign, q, buf, buf + N, buf + 2 * N);
    plan3_free(&p3);
    free(buf);
}

/* average_value: reduce3 sum, then data -= sum/(double)N (cuda_aid.cu:376) */
void orc_minus_average(long N, double *w)
{
    double sum = 0.0;
    long r;
    for (r = 0; r < N; r++) sum += w[r];
    for (r = 0; r < N; r++) w[r] = w[r] - sum / (double)N;
}

/* getConc (cuda_scft.cu:61-138) */
double orc_conc(const orc_params *p, const double *kxyzdz, double *wA, const double *wB,
                double *phiA, double *phiB, double *qa, double *qca, double *qb, double *qcb)
{
    const long N = npts(p);
    int NsA, NsB, own = 0, iz;
    double ql, ffl;
    long r;
    orc_plan3 p3;
    double *qInt, *buf;
    orc_contour_counts(p->fA, p->ds, &NsA, &NsB);
    if (!qa || !qca || !qb || !qcb) {
        own = 1;
        qa = (double *)malloc(sizeof(double) * (size_t)N * (NsA + 1));
        qca = (double *)malloc(sizeof(double) * (size_t)N * (NsA + 1));
        qb = (double *)malloc(sizeof(double) * (size_t)N * (NsB + 1));
        qcb = (double *)malloc(sizeof(double) * (size_t)N * (NsB + 1));
    }
    qInt = (double *)malloc(sizeof(double) * (size_t)N);
    buf = (double *)malloc(sizeof(double) * (size_t)N * 3);
    plan3_init(&p3, p->nx, p->ny, p->nz);

    orc_minus_average(N, wA);                                 /* :71 */
    for (r = 0; r < N; r++) qInt[r] = 1.0;                    /* qInt_init :77 */
    march_with_plan(p, &p3, kxyzdz, wA, qInt, NsA, 1, qa, buf, buf + N, buf + 2 * N);   /* :83 */
    march_with_plan(p, &p3, kxyzdz, wB, qInt, NsB, -1, qcb, buf, buf + N, buf + 2 * N); /* :86 */
    memcpy(qInt, qa + (long)NsA * N, sizeof(double) * (size_t)N);                     /* qa_to_qInt :93 */
    march_with_plan(p, &p3, kxyzdz, wB, qInt, NsB, 1, qb, buf, buf + N, buf + 2 * N);   /* :99 */
    memcpy(qInt, qcb, sizeof(double) * (size_t)N);                                    /* qa_to_qInt2 :106 */
    march_with_plan(p, &p3, kxyzdz, wA, qInt, NsA, -1, qca, buf, buf + N, buf + 2 * N); /* :112 */

    /* cal_ql + host divide (cuda_scft.cu:118-126) */
    ql = 0.0;
    for (r = 0; r < N; r++) ql = ql + qb[(long)NsB * N + r];
    ql /= N;
    ffl = p->ds / ql;

    /* w_to_phi_go (cuda_aid.cu:61-118): trapezoid, left to right in s */
    for (r = 0; r < N; r++) {
        double a = 0.0, b = 0.0;
        for (iz = 0; iz <= NsA; iz++) {
            long o = (long)iz * N + r;
            if (iz == 0 || (iz == NsA)) a += (0.50 * qa[o] * qca[o]);
            else a += qa[o] * qca[o];
        }
        for (iz = 0; iz <= NsB; iz++) {
            long o = (long)iz * N + r;
            if (iz == 0 || (iz == NsB)) b += (0.50 * qb[o] * qcb[o]);
            else b += qb[o] * qcb[o];
        }
        a *= ffl;
        b *= ffl;
        phiA[r] = a;
        phiB[r] = b;
    }

    plan3_free(&p3);
    free(buf);
    free(qInt);
    if (own) { free(qa); free(qca); free(qb); free(qcb); }
    return ql;
}

/* phi_w (cuda_aid.cu:382-416) with wopt, wcmp as parameters */
void orc_phi_w(const orc_params *p, long N, double *wA, double *wB,
               const double *phiA, const double *phiB)
{
    const double hAB = p->chiN, wopt = p->wopt, wcmp = p->wcmp;
    long i;
    for (i = 0; i < N; i++) {
        double eta = (wA[i] + wB[i] - hAB) / 2;
        double psum = 1.0 - phiA[i] - phiB[i];
        double waDiff = hAB * phiB[i] + eta - wA[i];
        double wbDiff = hAB * phiA[i] + eta - wB[i];
        waDiff -= wcmp * psum;
        wbDiff -= wcmp * psum;
        wA[i] += wopt * waDiff;
        wB[i] += wopt * wbDiff;
    }
}

/* Free(): host loop (cuda_scft.cu:199-220) */
void orc_free_energy(const orc_params *p, long N, const double *wA, const double *wB,
                     const double *phiA, const double *phiB, double Q, double *out)
{
    double freeAB = 0.0, freeW = 0.0, inCompMax = 0.0, freeS;
    long ijk;
    for (ijk = 0; ijk < N; ijk++) {
        double psum = 1 - phiA[ijk] - phiB[ijk];
        double fpsum = fabs(psum);
        if (fpsum > inCompMax) inCompMax = fpsum;
        freeAB = freeAB + p->chiN * phiA[ijk] * phiB[ijk];
        freeW = freeW - (wA[ijk] * phiA[ijk] + wB[ijk] * phiB[ijk]);
    }
    freeAB /= N;
    freeW /= N;
    freeS = -log(Q);
    out[0] = freeAB + freeW + freeS;
    out[1] = freeAB;
    out[2] = freeW;
    out[3] = freeS;
    out[4] = inCompMax;
}

/* loop body of Free (cuda_scft.cu:167-182) plus the F evaluation (:199-220) */
double orc_iterate(const orc_params *p, const double *kxyzdz, double *wA, double *wB,
                   double *phiA, double *phiB, double *fe)
{
    const long N = npts(p);
    double Q;
    orc_minus_average(N, wA);                   /* :170 */
    orc_minus_average(N, wB);                   /* :172 */
    Q = orc_conc(p, kxyzdz, wA, wB, phiA, phiB, NULL, NULL, NULL, NULL);  /* :174 */
    orc_phi_w(p, N, wA, wB, phiA, phiB);        /* :182 */
    if (fe) orc_free_energy(p, N, wA, wB, phiA, phiB, Q, fe);
    return Q;
}

void orc_run(const orc_params *p, double *wA, double *wB, double *phiA, double *phiB,
             int iters, double *fe_hist)
{
    const long N = npts(p);
    double *kt = (double *)malloc(sizeof(double) * (size_t)N);
    double fe[5];
    int it;
    orc_ktable(p, kt);
    for (it = 0; it < iters; it++) {
        orc_iterate(p, kt, wA, wB, phiA, phiB, fe);
        if (fe_hist) memcpy(fe_hist + 5 * (long)it, fe, sizeof(fe));
    }
    free(kt);
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint of one chain evaluation (no reference counterpart).       */
/* Forward: q(0) = 1, q(j) = S_b(j) q(j-1); qd(Ns) = 1, qd(j-1) = S_b(j) qd(j), */
/* S_b = W_b K W_b with W_b = exp(-u_b ds/2), b(j) = A for j <= NsA;          */
/* Q = mean q(Ns); phi_b = (ds/Q) trap_b[q qd].                              */
/* L = sum(lamA phiA + lamB phiB) + mu ln Q.  With e(s) = dL/dq(s) and        */
/* ed(s) = dL/dqd(s) (explicit parts), S symmetric gives                     */
/*   qb(Ns) = e(Ns), qb(j-1) = e(j-1) + S_b(j) qb(j)                          */
/*   qdb(0) = ed(0), qdb(j) = ed(j) + S_b(j) qdb(j-1)                         */
/* and, per step j of block b (dW = -(ds/2) W du):                            */
/*   du_b += -(ds/2)[qb(j) q(j) + q(j-1)(qb(j-1) - e(j-1))                    */
/*                  + qdb(j-1) qd(j-1) + qd(j)(qdb(j) - ed(j))]               */
/* ------------------------------------------------------------------------ */

double orc_conc_adjoint(const orc_params *p, const double *kxyzdz, const double *uA, const double *uB,
                        const double *lamA, const double *lamB, double mu, double *phiA, double *phiB,
                        double *Qout, double *gA, double *gB)
{
    const long N = npts(p);
    int NsA, NsB, Ns, j;
    long r;
    double Q, L, dLdQ, lphi = 0.0;
    orc_plan3 p3;
    double *q, *qd, *qb, *qdb, *e, *ed, *WA, *WB, *tmp, *buf;
    orc_contour_counts(p->fA, p->ds, &NsA, &NsB);
    Ns = NsA + NsB;
    q = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    qd = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    qb = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    qdb = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    e = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    ed = (double *)malloc(sizeof(double) * (size_t)N * (Ns + 1));
    WA = (double *)malloc(sizeof(double) * (size_t)N);
    WB = (double *)malloc(sizeof(double) * (size_t)N);
    tmp = (double *)malloc(sizeof(double) * (size_t)N);
    buf = (double *)malloc(sizeof(double) * (size_t)N * 2);
    plan3_init(&p3, p->nx, p->ny, p->nz);
    for (r = 0; r < N; r++) {
        WA[r] = exp(-uA[r] * p->ds / 2);
        WB[r] = exp(-uB[r] * p->ds / 2);
    }
#define WOF(jj) ((jj) <= NsA ? WA : WB)
    for (r = 0; r < N; r++) { q[r] = 1.0; qd[(long)Ns * N + r] = 1.0; }
    for (j = 1; j <= Ns; j++)
        step_with_plan(p, &p3, kxyzdz, WOF(j), q + (long)(j - 1) * N, q + (long)j * N, buf, buf + N);
    for (j = Ns; j >= 1; j--)
        step_with_plan(p, &p3, kxyzdz, WOF(j), qd + (long)j * N, qd + (long)(j - 1) * N, buf, buf + N);
    Q = 0.0;
    for (r = 0; r < N; r++) Q += q[(long)Ns * N + r];
    Q /= N;
    for (r = 0; r < N; r++) {
        double a = 0.0, b = 0.0;
        int s;
        for (s = 0; s <= NsA; s++) a += ((s == 0 || s == NsA) ? 0.5 : 1.0) * q[(long)s * N + r] * qd[(long)s * N + r];
        for (s = NsA; s <= Ns; s++) b += ((s == NsA || s == Ns) ? 0.5 : 1.0) * q[(long)s * N + r] * qd[(long)s * N + r];
        phiA[r] = a * p->ds / Q;
        phiB[r] = b * p->ds / Q;
        lphi += lamA[r] * phiA[r] + lamB[r] * phiB[r];
    }
    L = lphi + mu * log(Q);
    dLdQ = -lphi / Q + mu / Q;
    /* explicit parts: dphi_b(r)/dq(s)(r) = (ds/Q) c_b(s) qd(s)(r) */
    for (j = 0; j <= Ns; j++) {
        const double cA = (j <= NsA) ? ((j == 0 || j == NsA) ? 0.5 : 1.0) : 0.0;
        const double cB = (j >= NsA) ? ((j == NsA || j == Ns) ? 0.5 : 1.0) : 0.0;
        for (r = 0; r < N; r++) {
            const double g = p->ds / Q * (cA * lamA[r] + cB * lamB[r]);
            e[(long)j * N + r] = g * qd[(long)j * N + r] + (j == Ns ? dLdQ / N : 0.0);
            ed[(long)j * N + r] = g * q[(long)j * N + r];
        }
    }
    memcpy(qb + (long)Ns * N, e + (long)Ns * N, sizeof(double) * (size_t)N);
    for (j = Ns; j >= 1; j--) {
        step_with_plan(p, &p3, kxyzdz, WOF(j), qb + (long)j * N, tmp, buf, buf + N);
        for (r = 0; r < N; r++) qb[(long)(j - 1) * N + r] = e[(long)(j - 1) * N + r] + tmp[r];
    }
    memcpy(qdb, ed, sizeof(double) * (size_t)N);
    for (j = 1; j <= Ns; j++) {
        step_with_plan(p, &p3, kxyzdz, WOF(j), qdb + (long)(j - 1) * N, tmp, buf, buf + N);
        for (r = 0; r < N; r++) qdb[(long)j * N + r] = ed[(long)j * N + r] + tmp[r];
    }
    for (r = 0; r < N; r++) { gA[r] = 0.0; gB[r] = 0.0; }
    for (j = 1; j <= Ns; j++) {
        double *g = (j <= NsA) ? gA : gB;
        const long o1 = (long)j * N, o0 = (long)(j - 1) * N;
        for (r = 0; r < N; r++)
            g[r] += -(p->ds / 2) * (qb[o1 + r] * q[o1 + r] + q[o0 + r] * (qb[o0 + r] - e[o0 + r]) +
                                    qdb[o0 + r] * qd[o0 + r] + qd[o1 + r] * (qdb[o1 + r] - ed[o1 + r]));
    }
#undef WOF
    *Qout = Q;
    plan3_free(&p3);
    free(q); free(qd); free(qb); free(qdb); free(e); free(ed);
    free(WA); free(WB); free(tmp); free(buf);
    return L;
}

/* ------------------------------------------------------------------------ */
/* Anderson mixing (addition; the reference has only simple mixing).          */
/* Residual per field, the same one simple mixing steps along:               */
/*   eta = (wA+wB-chiN)/2, psi = 1-phiA-phiB,                                */
/*   dA = chiN*phiB + eta - wA - wcmp*psi,  dB = chiN*phiA + eta - wB - wcmp*psi */
/* History k = 0..m' (0 newest), m' = min(updates done, m).                  */
/*   U_ij = <d0-di, d0-dj>,  V_i = <d0-di, d0>  (i,j = 1..m')                */
/*   solve U c = V (Gaussian elimination, partial pivoting, FP64)            */
/*   w+ = w0 + sum c_i (wi - w0) + mix*(d0 + sum c_i (di - d0))              */
/* <a,b> sums over both fields and all grid points.  m' = 0 is simple mixing */
/* with fraction `mix`.                                                      */
/* Optional relative ridge rho (0 by default): U + rho*tr(U)/m' * I, the    */
/* Tikhonov form the HIP solver applies (scft_params.anderson_reg).          */
/* ------------------------------------------------------------------------ */

struct orc_anderson {
    long N;
    int m;
    double mix, wcmp;
    double ridge;     /* relative ridge of the least squares (0: none) */
    int nhist;        /* number of stored entries */
    double *w;        /* [(m+1)][2N] ring, entry 0 = newest */
    double *d;
    int last_mp;      /* history entries used by the last update */
    double last_c[64];/* its least-squares coefficients c_1..c_m' */
};

orc_anderson *orc_anderson_create(long N, int m, double mix, double wcmp)
{
    orc_anderson *a;
    if (m < 0 || m > 64) return NULL; /* the least squares is solved in fixed 64 x 64 storage */
    a = (orc_anderson *)calloc(1, sizeof(orc_anderson));
    a->N = N;
    a->m = m;
    a->mix = mix;
    a->wcmp = wcmp;
    a->w = (double *)malloc(sizeof(double) * (size_t)(m + 1) * 2 * N);
    a->d = (double *)malloc(sizeof(double) * (size_t)(m + 1) * 2 * N);
    return a;
}

void orc_anderson_set_ridge(orc_anderson *a, double ridge)
{
    a->ridge = ridge;
}

int orc_anderson_coef(const orc_anderson *a, double *c)
{
    int i;
    for (i = 0; i < a->last_mp; i++) c[i] = a->last_c[i];
    return a->last_mp;
}

int orc_anderson_history(const orc_anderson *a, int k, double *w, double *d)
{
    const long N2 = 2 * a->N;
    if (k < 0 || k >= a->nhist) return -1;
    memcpy(w, a->w + (long)k * N2, sizeof(double) * (size_t)N2);
    memcpy(d, a->d + (long)k * N2, sizeof(double) * (size_t)N2);
    return 0;
}

void orc_anderson_destroy(orc_anderson *a)
{
    if (!a) return;
    free(a->w);
    free(a->d);
    free(a);
}

/* solve A x = b in place (n <= 64), Gaussian elimination with partial pivoting.
   Returns 0 on success, -1 if a pivot is exactly zero. */
static int small_solve(int n, double *A, double *b)
{
    int i, j, k;
    for (k = 0; k < n; k++) {
        int piv = k;
        double best = fabs(A[k * n + k]);
        for (i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); piv = i; }
        if (best == 0.0) return -1;
        if (piv != k) {
            for (j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[piv * n + j]; A[piv * n + j] = t; }
            { double t = b[k]; b[k] = b[piv]; b[piv] = t; }
        }
        for (i = k + 1; i < n; i++) {
            double f = A[i * n + k] / A[k * n + k];
            for (j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (k = n - 1; k >= 0; k--) {
        double s = b[k];
        for (j = k + 1; j < n; j++) s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
    return 0;
}

double orc_anderson_update(orc_anderson *a, const orc_params *p, double *wA, double *wB,
                           const double *phiA, const double *phiB)
{
    const long N = a->N, N2 = 2 * N;
    const double hAB = p->chiN;
    int k, i, j, mp;
    double num = 0.0, den = 0.0;
    long r;
    /* shift the ring: entry k -> k+1 (drop the oldest) */
    for (k = (a->nhist < a->m + 1 ? a->nhist : a->m); k >= 1; k--) {
        memcpy(a->w + (long)k * N2, a->w + (long)(k - 1) * N2, sizeof(double) * (size_t)N2);
        memcpy(a->d + (long)k * N2, a->d + (long)(k - 1) * N2, sizeof(double) * (size_t)N2);
    }
    if (a->nhist < a->m + 1) a->nhist++;
    for (r = 0; r < N; r++) {
        double eta = (wA[r] + wB[r] - hAB) / 2;
        double psum = 1.0 - phiA[r] - phiB[r];
        double dA = hAB * phiB[r] + eta - wA[r] - a->wcmp * psum;
        double dB = hAB * phiA[r] + eta - wB[r] - a->wcmp * psum;
        a->w[r] = wA[r];
        a->w[N + r] = wB[r];
        a->d[r] = dA;
        a->d[N + r] = dB;
        num += dA * dA + dB * dB;
        den += wA[r] * wA[r] + wB[r] * wB[r];
    }
    mp = a->nhist - 1;
    {
        double U[64 * 64], V[64], c[64];
        const double *d0 = a->d, *w0 = a->w;
        for (i = 1; i <= mp; i++) {
            const double *di = a->d + (long)i * N2;
            double v = 0.0;
            for (r = 0; r < N2; r++) v += (d0[r] - di[r]) * d0[r];
            V[i - 1] = v;
            for (j = i; j <= mp; j++) {
                const double *dj = a->d + (long)j * N2;
                double u = 0.0;
                for (r = 0; r < N2; r++) u += (d0[r] - di[r]) * (d0[r] - dj[r]);
                U[(i - 1) * mp + (j - 1)] = u;
                U[(j - 1) * mp + (i - 1)] = u;
            }
        }
        if (a->ridge > 0.0 && mp > 0) {
            double tr = 0.0, lam;
            for (i = 0; i < mp; i++) tr += U[i * mp + i];
            lam = a->ridge * tr / mp;
            for (i = 0; i < mp; i++) U[i * mp + i] += lam;
        }
        for (i = 0; i < mp; i++) c[i] = V[i];
        if (mp > 0 && small_solve(mp, U, c) != 0)
            for (i = 0; i < mp; i++) c[i] = 0.0;
        a->last_mp = mp;
        for (i = 0; i < mp; i++) a->last_c[i] = c[i];
        for (r = 0; r < N2; r++) {
            double wn = w0[r], dn = d0[r];
            for (i = 1; i <= mp; i++) {
                wn += c[i - 1] * (a->w[(long)i * N2 + r] - w0[r]);
                dn += c[i - 1] * (a->d[(long)i * N2 + r] - d0[r]);
            }
            if (r < N) wA[r] = wn + a->mix * dn;
            else wB[r - N] = wn + a->mix * dn;
        }
    }
    return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

double orc_iterate_anderson(orc_anderson *a, const orc_params *p, const double *kxyzdz,
                            double *wA, double *wB, double *phiA, double *phiB, double *fe)
{
    const long N = npts(p);
    double Q;
    orc_minus_average(N, wA);
    orc_minus_average(N, wB);
    Q = orc_conc(p, kxyzdz, wA, wB, phiA, phiB, NULL, NULL, NULL, NULL);
    orc_anderson_update(a, p, wA, wB, phiA, phiB);
    if (fe) orc_free_energy(p, N, wA, wB, phiA, phiB, Q, fe);
    return Q;
}
