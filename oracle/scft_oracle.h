/*
 * scft_oracle.h -- CPU FP64 restatement of the reference SCFT method.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference and the timed
 * CPU baseline ("port").  Nothing in the product (libscft_hip.so, the driver,
 * the Python host package) links, loads or calls it; only tests/, the smoke
 * check and bench.py's cpu_baseline leg do.
 *
 * It follows the reference line by line (file:line cited at each function in
 * scft_oracle.c): the truncated Pi (struct.h:11), the truncated contour counts
 * (init_cuda.cu:330-331), the k^2 tables (init_cuda.cu:408-428), the
 * operator-split step with its arithmetic order (cuda_aid.cu:227-241, 306-328,
 * 346-361), the four marches of getConc (cuda_scft.cu:61-138), the trapezoid
 * contour sum (cuda_aid.cu:61-118), Q (cuda_aid.cu:119-168), the simple mixing
 * update (cuda_aid.cu:382-416) and the free energy evaluated after the update
 * (cuda_scft.cu:182-220).  The FFT is this file's own mixed-radix Stockham
 * transform (the reference calls cuFFT, a closed library): forward uses
 * e^{-2 pi i jk/n}, inverse e^{+2 pi i jk/n}, both unnormalised, as cuFFT's
 * D2Z/Z2D are.
 *
 * Grid index: r = (k*Ny + j)*Nx + i (x fastest), as init_cuda.cu:425.
 * Propagator storage here: q[s][r] (one slice per contour point).
 */
#ifndef SCFT_ORACLE_H
#define SCFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    double lx, ly, lz;
    double chiN;   /* xN in ab.txt (hAB) */
    double fA;     /* fa in ab.txt */
    double ds;     /* ds0 */
    double pi;     /* value of pi used for wavenumbers: 3.1415926535 reproduces the reference */
    double wopt;   /* simple-mixing fraction (0.05 in the reference) */
    double wcmp;   /* incompressibility penalty (0.10 in the reference) */
} orc_params;

/* threads for the OpenMP loops (<=0: leave OpenMP's default) */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* ---- FFT (own implementation) ---- */
/* in-place complex 1-D FFT of length n on split arrays; sign -1 forward, +1 inverse; unnormalised */
int orc_fft1d(int n, double *re, double *im, int sign);
/* in-place complex 3-D FFT on [nz][ny][nx] split arrays */
int orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign);
/* brute-force O(n^2) DFT in long double, 1-D, out of place (for pinning the FFT) */
void orc_dft1d_brute(int n, const double *re, const double *im, double *ore, double *oim, int sign);

/* ---- constants (init_cuda.cu:306-331, 398-428) ---- */
void orc_contour_counts(double fA, double ds, int *NsA, int *NsB);
/* k^2 per axis, reference formula (length n each) */
void orc_ksq_axis(int n, double l, double pi, double *ksq);
/* kxyzdz[r] = exp(-ds*(kx^2+ky^2+kz^2)) over the full grid */
void orc_ktable(const orc_params *p, double *kxyzdz);

/* ---- propagator (cuda_scft.cu:435-563) ---- */
/* one contour step: qout = (IFFT[k * FFT[qin*wdz]] * wdz)/N   (cuda_aid.cu:237, 323-325, 358) */
void orc_step(const orc_params *p, const double *kxyzdz, const double *wdz,
              const double *qin, double *qout);
/* march: q is [(ns+1)][N]; sign +1 seeds slice 0 with qinit and marches up,
   sign -1 seeds slice ns and marches down (cuda_scft.cu:456-555) */
void orc_march(const orc_params *p, const double *kxyzdz, const double *w,
               const double *qinit, int ns, int sign, double *q);

/* mean subtraction, w -= sum(w)/N  (cuda_scft.cu:14-59, cuda_aid.cu:365-380) */
void orc_minus_average(long N, double *w);

/* getConc (cuda_scft.cu:61-138).  wA is mean-subtracted in place (P6).
   qa,qca: [(NsA+1)][N]; qb,qcb: [(NsB+1)][N] -- may be NULL (then allocated
   internally and freed).  Returns Q (= ql/N). */
double orc_conc(const orc_params *p, const double *kxyzdz, double *wA, const double *wB,
                double *phiA, double *phiB, double *qa, double *qca, double *qb, double *qcb);

/* simple mixing phi_w (cuda_aid.cu:382-416), in place on wA,wB */
void orc_phi_w(const orc_params *p, long N, double *wA, double *wB,
               const double *phiA, const double *phiB);

/* free-energy terms as Free() computes them (cuda_scft.cu:199-220):
   out[0]=F, out[1]=F_AB, out[2]=F_W, out[3]=F_S, out[4]=inCompMax */
void orc_free_energy(const orc_params *p, long N, const double *wA, const double *wB,
                     const double *phiA, const double *phiB, double Q, double *out);

/* One SCFT iteration exactly as the reference's loop body
   (cuda_scft.cu:167-182): mean-subtract wA, wB; getConc; phi_w.  Then the
   free energy terms (evaluated with the updated w, P4) into fe[5].  Returns Q. */
double orc_iterate(const orc_params *p, const double *kxyzdz, double *wA, double *wB,
                   double *phiA, double *phiB, double *fe);

/* Run `iters` iterations of orc_iterate; fe_hist (may be NULL) receives
   5 doubles per iteration. */
void orc_run(const orc_params *p, double *wA, double *wB, double *phiA, double *phiB,
             int iters, double *fe_hist);

/* ---- discrete adjoint of one chain evaluation (addition) ----
   For fields uA, uB used as given (no mean subtraction), forward
   phi = getConc(uA, uB) and L = sum_r(lamA phiA + lamB phiB) + mu ln Q;
   returns L and gA = dL/duA, gB = dL/duB (reverse mode through the marches,
   trapezoid sum and Q).  phiA/phiB/Q receive the forward values. */
double orc_conc_adjoint(const orc_params *p, const double *kxyzdz, const double *uA, const double *uB,
                        const double *lamA, const double *lamB, double mu, double *phiA, double *phiB,
                        double *Q, double *gA, double *gB);

/* ---- Anderson mixing (addition; no reference counterpart) ----
   History of m previous (w, residual) pairs; FP64 least squares. */
typedef struct orc_anderson orc_anderson;
orc_anderson *orc_anderson_create(long N, int m, double mix, double wcmp);
void orc_anderson_destroy(orc_anderson *a);
/* relative ridge rho of the least squares, U + rho*tr(U)/m' * I (0, the default: none) */
void orc_anderson_set_ridge(orc_anderson *a, double ridge);
/* coefficients c_1..c_m' of the last update into c; returns m' */
int orc_anderson_coef(const orc_anderson *a, double *c);
/* history entry k (0 = newest input of the last update): w and residual d, 2N each
   (A field, then B).  Returns -1 if there is no entry k. */
int orc_anderson_history(const orc_anderson *a, int k, double *w, double *d);
/* one Anderson update of (wA, wB) given the phi of the current w; returns the
   residual norm err = sqrt(sum(dA^2+dB^2)/sum(wA^2+wB^2)) */
double orc_anderson_update(orc_anderson *a, const orc_params *p, double *wA, double *wB,
                           const double *phiA, const double *phiB);
/* iteration with Anderson: mean-subtract, getConc, Anderson update, F */
double orc_iterate_anderson(orc_anderson *a, const orc_params *p, const double *kxyzdz,
                            double *wA, double *wB, double *phiA, double *phiB, double *fe);

#ifdef __cplusplus
}
#endif
#endif
