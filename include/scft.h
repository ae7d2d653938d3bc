/*
 * scft.h -- C ABI of the MI355X SCFT solver for AB diblock-copolymer melts.
 *
 * The library (libscft_hip.so) replaces the reference's CUDA solver layer of
 * ren5885a/SCFT_Ver1.01.  Every entry point returns an int status (SCFT_OK = 0)
 * and never exits the process; scft_last_error() gives the message.  There is
 * no global state: one context per host thread.  Host arrays are caller-owned,
 * FP64, laid out [instance][z][y][x] with x fastest (point index
 * r = (k*Ny + j)*Nx + i, as init_cuda.cu:425).  Device buffers are owned by
 * the context.
 *
 * Correspondence with the reference (file:line):
 *   scft_create / scft_destroy  <- init_cuda + initialize_cufft (init_cuda.h:11-13,
 *                                  init_cuda.cu:227-546) / finalize_cufft (init_cuda.cu:551-628)
 *   scft_set_fields             <- the phi_<n>.dat upload of initialize_cufft
 *                                  (init_cuda.cu:435-456, 508-513)
 *   scft_get_fields             <- host reads of the managed pha/phb/wa/wb arrays in Free
 *                                  (cuda_scft.cu:199-243)
 *   scft_propagate              <- sovDifFft (cuda_scft.h:1, cuda_scft.cu:435-563)
 *   scft_concentrations         <- getConc (cuda_scft.cu:61-138)
 *   scft_update_fields          <- phi_w launch in Free (cuda_scft.cu:176-187, cuda_aid.cu:382-416)
 *   scft_free_energy            <- the F block of Free (cuda_scft.cu:191-223)
 *   scft_iterate                <- the do-while loop of Free (cuda_scft.cu:167-253)
 *   scft_time_iterations        <- fft_test's cudaEvent timing (cuda_scft.cu:382-432)
 *   scft_contour_counts         <- NsA/dNsB of initialize_cufft (init_cuda.cu:330-331)
 */
#ifndef SCFT_H
#define SCFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SCFT_OK 0
#define SCFT_E_INVALID 1     /* bad argument or parameter */
#define SCFT_E_UNSUPPORTED 2 /* grid length with no kernel tiling */
#define SCFT_E_HIP 3         /* HIP runtime error (message names the call) */
#define SCFT_E_NOMEM 4       /* device allocation failed */
#define SCFT_E_NONFINITE 5   /* Q non-finite or <= 0, or non-finite input field */
#define SCFT_E_IO 6          /* file missing, short or malformed */
#define SCFT_E_STATE 7       /* call out of order (e.g. update before concentrations) */

/* propagator precision */
#define SCFT_FP64 0 /* everything FP64 (the reference's arithmetic) */
#define SCFT_FP32 1 /* propagator, FFT and stored q in FP32; Q, phi, F, fields FP64 */
#define SCFT_FP32X 2 /* stored q, W and every pass's loads and stores in FP32 (the HBM
                        bytes of SCFT_FP32; the W factors applied before a step stored as
                        W - 1, which FP32 keeps to about 1e-9), the transform arithmetic in
                        FP64 registers and LDS (rounded to FP32 once per store); Q, phi,
                        F, fields FP64 */

/* field update */
#define SCFT_MIX_SIMPLE 0   /* phi_w of the reference (cuda_aid.cu:382-416) */
#define SCFT_MIX_ANDERSON 1 /* Anderson mixing, FP64 least squares (addition) */

/* value of pi the reference uses for wavenumbers (struct.h:11) */
#define SCFT_REFERENCE_PI 3.1415926535

typedef struct scft_ctx scft_ctx;

typedef struct {
    int nx, ny, nz;       /* grid (Nx, Ny, Nz of ab.txt) */
    double lx, ly, lz;    /* box lengths in Rg (lx, ly, lz of ab.txt) */
    double chiN;          /* xN of ab.txt (hAB) */
    double fA;            /* fa of ab.txt */
    double ds;            /* ds0 of ab.txt */
    int batch;            /* independent instances sharing these parameters (batch of ab.txt) */
    int precision;        /* SCFT_FP64, SCFT_FP32 or SCFT_FP32X */
    int mixing;           /* SCFT_MIX_SIMPLE or SCFT_MIX_ANDERSON */
    double wopt;          /* simple-mixing fraction; reference 0.05 (cuda_aid.cu:395) */
    double wcmp;          /* incompressibility penalty; reference 0.10 (cuda_aid.cu:396) */
    int anderson_m;       /* Anderson history length (0 = simple mixing with anderson_mix) */
    double anderson_mix;  /* Anderson mixing fraction applied to the combined residual */
    double pi;            /* pi used for wavenumbers; 0 selects SCFT_REFERENCE_PI */
    int use_graph;        /* 1: iterations replay one captured HIP graph; 0: eager launches */
    int xy_plane;         /* -1 auto (scft_autotune decides), 0 separate x and y passes,
                             1 fused xy-plane kernel (one workgroup per plane), 2 split
                             xy-plane kernel (one workgroup per plane and real part), 3 the
                             split kernel with the three-level line tiling (n = 32, 64, 128),
                             4 and 5 forms 2 and 3 with each y line on consecutive lanes of
                             one wave; 1..5 need nx == ny and a plane that fits LDS */
    int tiling;           /* -1 default (one line per thread group until scft_autotune),
                             0 or 1 force one or two lines per thread group on every pass,
                             2 the three-level tiling (2x the threads per line; n = 32, 64,
                             128, 256) */
    int checkpoint;       /* 0: keep every q(s) slice; 1: sqrt(Ns) checkpointing -- keep
                             z(kK) for a spacing K near sqrt(Ns+1) (the one whose contour
                             re-marches the fewest steps, with no more slices than
                             K = ceil(sqrt(Ns+1)) would take); the contour is summed as the
                             march passes each pair's second slice, re-marching only the
                             first half's segments (about 1.4x the propagator steps,
                             O(sqrt Ns) slices of memory) */
    int fuse_steps;       /* 1: the packed march alternates its axis order and runs each
                             step boundary (the last axis of step t, the first of t+1) as
                             one kernel; 0: every step's passes separately */
    double anderson_reg;  /* relative ridge of the Anderson least squares: U + anderson_reg *
                             tr(U)/m * I (0, the default: plain least squares).  Bounds the
                             condition number, so that derivatives through many unrolled
                             Anderson iterations (scft_adjoint_loop) stay finite */
} scft_params;

/* Fill p with the reference's defaults (ab.txt sample sizes aside):
   FP64, simple mixing, wopt 0.05, wcmp 0.10, pi = SCFT_REFERENCE_PI, batch 1,
   Anderson m 10 / mix 0.1 (used only with SCFT_MIX_ANDERSON), HIP graph on,
   xy_plane auto, step-boundary fusion on. */
void scft_default_params(scft_params *p);

/* NsA and NsB exactly as the reference truncates them (init_cuda.cu:330-331). Host only. */
int scft_contour_counts(double fA, double ds, int *NsA, int *NsB);

/* Create a context on HIP device `device`: validates p, builds the wavenumber
   and twiddle tables, allocates every device buffer.  *out is NULL on error. */
int scft_create(const scft_params *p, int device, scft_ctx **out);
int scft_destroy(scft_ctx *ctx);
/* message of the last failing call on ctx (or of the last failed scft_create
   when ctx is NULL, per host thread) */
const char *scft_last_error(const scft_ctx *ctx);

/* wA, wB: host [batch][N].  Rejects non-finite values (SCFT_E_NONFINITE). */
int scft_set_fields(scft_ctx *ctx, const double *wA, const double *wB);
/* any pointer may be NULL; phi is the last computed concentration */
int scft_get_fields(scft_ctx *ctx, double *phiA, double *phiB, double *wA, double *wB);

/* One march of block `block` (0 = A with wA and NsA steps, 1 = B with wB and
   NsB steps) as sovDifFft: wdz = exp(-w*ds/2) from the current field (no mean
   subtraction); sign +1 seeds slice 0 with q_init and marches up, sign -1
   seeds slice ns and marches down.  q_init: host [batch][N], or NULL for 1.
   q_out: host [batch][ns+1][N]. */
int scft_propagate(scft_ctx *ctx, int block, int sign, const double *q_init, double *q_out);

/* getConc: subtract the mean of wA (as getConc does, cuda_scft.cu:71), march the
   chain, compute Q and phiA/phiB by the trapezoid rule.  Q: host [batch] or NULL. */
int scft_concentrations(scft_ctx *ctx, double *Q);

/* Field update from the last concentrations (simple mixing or Anderson per
   params), then the free-energy terms with the updated fields. */
int scft_update_fields(scft_ctx *ctx);

/* Free-energy terms of the last update, host [batch][5]:
   F, F_AB, F_W, F_S, inCompMax (cuda_scft.cu:199-221). */
int scft_free_energy(scft_ctx *ctx, double *out);

/* n_iter SCFT iterations, each: subtract the means of wA and wB, getConc, field
   update, free energy (cuda_scft.cu:167-223).  If tol > 0, stop after the first
   iteration whose inCompMax (all instances) is below tol; *done (may be NULL)
   receives the number of iterations run.  fe (may be NULL): host [batch][5] of
   the last iteration. */
int scft_iterate(scft_ctx *ctx, int n_iter, double tol, int *done, double *fe);

/* Time n_iter iterations with HIP events on the context's stream (no host
   synchronisation inside); *ms receives the elapsed milliseconds. */
int scft_time_iterations(scft_ctx *ctx, int n_iter, float *ms);

/* Kernel-level timing for the roofline: runs `reps` back-to-back launches of
   the propagator axis pass `axis` (0 x, 1 y, 2 z; 3, 4, 10, 12, 14 the
   xy-plane kernel forms 1, 2, 3, 4, 5; 5, 6, 7 the step-boundary x, y, z
   kernels; 8, 9, 11, 13, 15 the step-boundary forms of 3, 4, 10, 12, 14;
   16 the contour kernel over the current slice store, into scratch, which
   needs the full store: SCFT_E_INVALID in checkpoint mode)
   between two HIP events and returns the mean duration in ms.  A boundary
   kernel runs as in the march: in place on the work field, launch r with the
   arguments of contour step 1 + r % (Ns - 1), z(t) to slice t of the store.
   Changes the work field and those slices (the next evaluation recomputes
   both), nothing else. */
int scft_time_pass(scft_ctx *ctx, int axis, int reps, float *ms);

/* Discrete adjoint of one chain evaluation (no reference counterpart; the
   closest reference idea is the unused targeted update phi_w_constrained,
   cuda_aid.cu:418-452).  For fields uA, uB (host [batch][N], used as given:
   no mean subtraction) computes phi = getConc(u), Q, and
   L = sum_r (lamA phiA + lamB phiB) + mu ln Q per instance, and its gradient
   gA = dL/duA, gB = dL/duB by reverse mode through the march, the trapezoid
   sum and Q.  lamA/lamB may be NULL (zero).  Outputs may be NULL.  With
   lam = 0, mu = 1 the gradient is dlnQ/du = -phi/N. */
int scft_conc_adjoint(scft_ctx *ctx, const double *uA, const double *uB, const double *lamA,
                      const double *lamB, double mu, double *L, double *phiA, double *phiB, double *Q,
                      double *gA, double *gB);

/* Gradient of a structure-targeting loss through the unrolled SCFT loop (no
   reference counterpart).  External field h (host [batch][N]) acts on the A
   blocks: iteration k evaluates the chain at (P wA + h, P wB) (P removes the
   mean), and between evaluations updates w with the context's mixing (simple
   or Anderson, in FP64).  Starting from (w0A, w0B), n_iter chain
   evaluations are made; loss[b] = sum_r (phiA - target)^2 of the last one.
   grad_h (may be NULL: forward only) receives dloss/dh by the discrete
   adjoint of every step (mixing, least squares, chain, mean removal).
   phiA_last (may be NULL) receives the last phiA.  The trajectory and its
   adjoint stay in device memory: 2 * batch * N doubles per iteration for
   each of 3 arrays (5 with grad_h; in checkpoint mode with grad_h also the
   iteration's checkpoints and the march's opened states beside them, about
   2 sqrt(Ns) propagator slices), held by the
   context after the call and grown when a later call needs more; only the
   m x m least-squares solves run on the host. */
int scft_adjoint_loop(scft_ctx *ctx, const double *h, const double *w0A, const double *w0B, int n_iter,
                      const double *target, double *loss, double *grad_h, double *phiA_last);

/* Autotuner (no reference counterpart; the reference's launch shapes are fixed
   by factor_decompose_1024, init_cuda.cu:21-80): times every tiling variant of
   each axis pass (`reps` launches each) and the fused xy-plane kernel, keeps
   the fastest.  choice (may be NULL): int[4] = variant of x, y, z pass (0..2),
   plane form used (0 none, 1 xy-plane, 2 split xy-plane, 3 split three-level,
   4 and 5 the split forms with lane-local y lines);
   ms (may be NULL): float[4] = best x, y, z pass and chosen plane-kernel times.
   The plane form and the final axis variants are chosen by timing whole
   marches on stand-in factors, which overwrites the W factors and the slice
   store; fields, phi, Q and F are kept, and the next concentrations or
   iteration recomputes the rest. */
int scft_autotune(scft_ctx *ctx, int reps, int *choice, float *ms);

/* Apply a tiling choice of the form scft_autotune returns (int[4]: variant of
   the x, y, z pass, plane form 0..5) without timing anything, e.g. one
   cached from an earlier scft_autotune on the same grid.  SCFT_E_UNSUPPORTED if
   a variant or the plane kernel does not fit this grid.  No reference
   counterpart (its launch shapes are fixed, init_cuda.cu:21-80). */
int scft_set_tiling(scft_ctx *ctx, const int *choice);

/* Switch an FP32-storage context between SCFT_FP32 and SCFT_FP32X (no
   reference counterpart).  Both forms keep the same FP32 slices and FP64
   fields, phi, Q, F and Anderson history (each evaluation forms the W factors
   in the current form's storage), so a solve can run its bulk
   iterations in SCFT_FP32 and finish in SCFT_FP32X, whose FP64 transform
   arithmetic removes the FP32 form's rounding floor.  The tiling carries over
   where the new form supports it (else the default); scft_autotune or
   scft_set_tiling may follow.  SCFT_E_STATE for an FP64 context. */
int scft_set_precision(scft_ctx *ctx, int precision);

/* ---- heterogeneous parameter sweep (no reference counterpart: the
   reference's batch shares chiN, f, box and grid, struct.h:42-68) ----
   Independent instances with their own cubic grid n^3, box L^3, chiN and fA
   (so their own NsA, NsB); everything else (ds, precision, mixing, pi) comes
   from `base`.  Instances are grouped by grid size, one SCFT iteration of the
   whole sweep is one captured HIP graph, and scft_sweep_autotune tunes each
   group's passes.  Instance indices follow the input order. */
typedef struct scft_sweep scft_sweep;
typedef struct {
    int n;       /* grid n x n x n */
    double L;    /* box L x L x L (Rg) */
    double chiN;
    double fA;
} scft_instance;

int scft_sweep_create(const scft_params *base, const scft_instance *inst, int count, int device,
                      scft_sweep **out);
int scft_sweep_destroy(scft_sweep *sw);
const char *scft_sweep_last_error(const scft_sweep *sw);
/* fields of instance i (host, n^3 each); restarts the Anderson history of its group */
int scft_sweep_set_fields(scft_sweep *sw, int i, const double *wA, const double *wB);
int scft_sweep_get_fields(scft_sweep *sw, int i, double *phiA, double *phiB, double *wA, double *wB);
/* n_iter iterations of every instance; fe (may be NULL): [count][5] as scft_iterate */
int scft_sweep_iterate(scft_sweep *sw, int n_iter, double *fe);
int scft_sweep_time_iterations(scft_sweep *sw, int n_iter, float *ms);
int scft_sweep_autotune(scft_sweep *sw, int reps);
/* N, NsA, NsB of instance i */
int scft_sweep_size(const scft_sweep *sw, int i, long *N, int *NsA, int *NsB);

/* scft_propagate without allocation.  No counterpart in the reference: for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  One march of block
   `block` as scft_propagate does it (same p, block, sign and the same
   kernels), on caller-owned DEVICE arrays: w [batch][N] (the block's field,
   used as given), q_init [batch][N] or NULL for 1, q_out [batch][ns+1][N]
   (ns = NsA or NsB of p), all FP64.  The W factors, the march slices, the
   work field and the k/twiddle tables live in one static device array of the
   library (about 21 MB on every GPU it runs on, held until the process ends);
   batch * N may be at most 262144 (64^3) points, larger sizes return
   SCFT_E_INVALID.  Calls on the same GPU share that array and must not
   overlap (one host thread at a time per GPU); the call returns when q_out
   is written.  `stream` is a hipStream_t (NULL: the null stream).  Errors:
   scft_last_error(NULL). */
int scft_propagate_noalloc(const scft_params *p, int device, int block, int sign, const double *w,
                           const double *q_init, double *q_out, void *stream);

/* Attainable HBM bandwidth of `device` (no reference counterpart; the
   measured peak beside the nominal one in a roofline): `reps` copies of
   `bytes` bytes (a multiple of 16) by a copy kernel with 16-byte nontemporal
   loads and stores, between HIP events, best of three timings; *gbps
   receives (read + write bytes) / s / 1e9.  Allocates and frees two buffers
   of `bytes`. */
int scft_copy_bandwidth(int device, size_t bytes, int reps, double *gbps);

/* Problem sizes: N points, NsA, NsB, device bytes allocated. */
int scft_sizes(const scft_ctx *ctx, long *N, int *NsA, int *NsB, long *device_bytes);

#ifdef __cplusplus
}
#endif
#endif /* SCFT_H */
