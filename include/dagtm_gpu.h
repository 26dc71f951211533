/*
 * dagtm_gpu.h -- C ABI of the MI355X DA-GTM engine (libdagtm_gpu.so).
 *
 * The reference program (jychoi-hpc/da-gtm) has no plugin or FFI layer: its
 * hot path is a set of internal C functions of libdagtmcomm called from
 * dagtm/dagtm.cxx with GSL matrices and MPI communicators.  Each entry point
 * below replaces one of those call sites on a 1x1 process grid (one GPU holds
 * the whole K x N problem, so the communicators disappear).  The replaced
 * reference interface is cited per function.
 *
 * Conventions (mirroring the reference's, comm.h:122-131):
 *   - every function returns 0 (DG_OK) or a positive DG_E* code and never
 *     aborts; dg_last_error() gives the message of the last failure;
 *   - matrices are row-major doubles with an explicit row stride (ld*), so a
 *     padded gsl_matrix (tda, comm.c:1866-1883) can be passed directly;
 *   - host buffers are borrowed only for the duration of a call; device
 *     buffers belong to the context; calls are synchronous at the boundary;
 *   - one context per host thread; a context owns one HIP stream.
 */
#ifndef DAGTM_GPU_H
#define DAGTM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define DG_OK       0
#define DG_EINVAL   1   /* bad argument or unsupported shape */
#define DG_ENOMEM   2   /* device allocation failed */
#define DG_EHIP     3   /* HIP runtime error (no device, launch failure) */
#define DG_ESTATE   4   /* call out of order (e.g. mstep before estep) */
#define DG_ESOLVE   5   /* M-step factorisation failed even with jitter */

typedef struct dg_ctx dg_ctx;

enum dg_precision {
    DG_FP64     = 0,  /* every step in FP64 (the FP64 GPU oracle) */
    DG_MIXED    = 1,  /* 16-bit split MFMA distances, FP32 epilogue, FP64 M-step */
    DG_MIXED_R1 = 2   /* DG_MIXED with R carried as one FP16 term in the R T
                         product (not a hi/lo pair): a quarter fewer pass-2
                         MFMAs; G and R T to ~2e-5 relative instead of ~1e-8 */
};

typedef struct {
    int    device;      /* HIP device ordinal */
    int    precision;   /* enum dg_precision */
    double lambda;      /* M-step ridge, A + (lambda/beta) I; 0 = reference */
} dg_config;

/* Replaces the set-up of dagtm.cxx:893-1068 (MPI_Init, grid, allocation). */
int  dg_create(const dg_config *cfg, dg_ctx **out);
int  dg_destroy(dg_ctx *ctx);
const char *dg_last_error(const dg_ctx *ctx);
/* Version / build string of the library. */
const char *dg_version(void);

/* Copies the problem to the device.  Replaces h5read_byrow of T
 * (dagtm.cxx:1076-1102) and the initial dagtm_dist (dagtm.cxx:1232):
 *   T  N x D   (ldT >= D)      data
 *   FI K x M1  (ldFI >= M1)    basis matrix, M1 = M + 1
 *   Y  K x D   (ldY >= D)      centres (FI W0, or Y from a checkpoint)
 *   beta                       inverse noise variance */
int dg_upload(dg_ctx *ctx, const double *T, long N, int D, long ldT,
              const double *FI, int K, int M1, long ldFI,
              const double *Y, long ldY, double beta);

/* Replaces the restart read of Y and beta (dagtm.cxx:1161-1213). */
int dg_set_state(dg_ctx *ctx, const double *Y, long ldY, double beta);

/* New latent basis with the same M + 1, as in the HDA expansion: replaces
 * the re-set-up of dagtm.cxx:2112-2160 (new FI, re-allocation of every
 * K-sized buffer, Y = FI W, dist).  FI is K x M1 (ldFI >= M1, M1 must equal
 * the uploaded M + 1); W (M1 x D, ldW >= D) may be NULL to keep the W of the
 * last dg_mstep.  T stays resident; beta is unchanged; the G of the last
 * M-step (used by dg_next_ctemp) restarts at zero; a dg_estep is needed
 * before the next dg_mstep. */
int dg_set_basis(dg_ctx *ctx, const double *FI, int K, int M1, long ldFI,
                 const double *W, long ldW);

typedef struct {
    double llh;        /* normalised log-likelihood (mpicomm.c:463-500, from g) */
    double nfree;      /* normalised -F (from h; equals -llh in EM mode)        */
    long   nzero;      /* columns that took the zero-column fall-back (Q2)      */
} dg_estep_out;

/* One E-step at the context's beta: distances, responsibilities and both
 * qualities.  Replaces dagtm_dist_ws (comm.c:560-629) for the current Y,
 * dagtm_mpi_resp (mpicomm.c:316-414) and the two dagtm_mpi_qual calls
 * (mpicomm.c:463-500) of dagtm.cxx:1537-1603.  R is never materialised:
 * the call leaves G = rowsum(R) and RT = R T on the device for dg_mstep. */
int dg_estep(dg_ctx *ctx, int isDA, double temp, dg_estep_out *out);

/* M-step and beta update: G, A = FI^T diag(G) FI (+ lambda/beta I),
 * B = FI^T (R T), W by Cholesky, Y = FI W and beta = N D / sum R o dist.
 * Replaces dagtm.cxx:1715-1785 (rowsum, dgemm, dagtm_solve2, dgemm) and
 * dagtm_mpi_beta (mpicomm.c:276-302).  *beta_out may be NULL. */
int dg_mstep(dg_ctx *ctx, double *beta_out);

enum { DG_SCHED_EXP = 0, DG_SCHED_LINEAR = 1, DG_SCHED_AUTO = 2 };

typedef struct {
    int    isDA;        /* 1 = deterministic annealing, 0 = plain EM (-m) */
    int    schedule;    /* DG_SCHED_EXP (-c exp), DG_SCHED_LINEAR (-c linear) or
                           DG_SCHED_AUTO (-c auto) */
    int    maxloop;     /* -j, 0 = unlimited */
    double alpha;       /* -a, EXP cooling factor */
    double delta;       /* LINEAR step (dagtm.cxx:2169-2171) */
    double eps;         /* -e, convergence threshold on |diff| */
    double start_temp;  /* -S or the first critical temperature / alpha^10 */
    int    nloop0;      /* iterations done before this run: 0, or the iter of the
                           checkpoint a restart read (dagtm.cxx:1197-1212); nloop
                           counts on from it and maxloop is absolute */
} dg_anneal_params;

typedef struct {
    int    nloop;        /* total E+M iterations */
    int    ntemps;       /* temperatures visited */
    int    exit_maxloop; /* 1 when the run ended on maxloop */
    double temp, llh, nfree, qual, beta;
} dg_anneal_result;

/* Progress row of dagtm.cxx:1570-1571, called after every E-step. */
typedef void (*dg_iter_cb)(void *user, int nloop, int lloop, double temp,
                           double nfree, double llh, double diff,
                           double beta);

/* The annealing loop dagtm.cxx:1508-2206 for the EXP, LINEAR and AUTO
 * schedules and EM mode, with the reference's convergence tests (|diff| <
 * eps, the 50-iteration moving average, qual_old = 0 at start).  AUTO moves
 * to dg_next_ctemp's temperature, or multiplies by alpha when that is not
 * below the current one (dagtm.cxx:1919-1935). */
int dg_anneal(dg_ctx *ctx, const dg_anneal_params *p, dg_anneal_result *res,
              dg_iter_cb cb, void *user);

/* ---- the annealing control over caller-supplied steps --------------------
 * dg_anneal and the dagtm driver both run the reference's temperature loop
 * (dagtm.cxx:1508-2206) through dg_anneal_ctl, which calls back for every
 * step, so a driver with its own per-iteration work (progress rows,
 * checkpoints, the forced checkpoint of dagtm.cxx:1609-1687) or with HDA
 * grid doubling (dagtm.cxx:1938-2164) uses the same control.  Per
 * temperature: clamp temp < 1 to 1 and mark the last pass (1510-1514),
 * refill the 50-entry quality queue with DBL_MAX (1521), remember
 * qual_before_temp = llh (1523); per iteration: E-step, qual_old = qual (0
 * before the first), nloop++ and lloop++, qual = nfree (DA) or llh (EM),
 * diff = qual - qual_old, on_iter, then break when |diff| < eps or, once the
 * global nloop exceeds 50, when the queue mean moved by less than eps after
 * pushing qual (1689-1709; both before the M-step); M-step; stop after it
 * when nloop reaches maxloop (1790-1795).  After a temperature (DA, not
 * the last): LINEAR temp -= delta; AUTO temp = next_ctemp's value, or temp *
 * alpha when that is not below temp (1919-1935); EXP runs the HDA rule
 * below, then temp *= alpha. */
#define DG_ESTOP    6   /* a hook stopped dg_anneal_ctl (not a failure of the hook) */

typedef struct {
    /* required: one E-step at (isDA, temp) giving llh and nfree */
    int (*estep)(void *user, int isDA, double temp, double *llh, double *nfree);
    /* required: the M-step and beta update */
    int (*mstep)(void *user);
    /* AUTO schedule and HDA arming: the largest critical temperature below
     * temp of the current state (dagtm_mpi_next_ctemp); may be NULL otherwise */
    int (*next_ctemp)(void *user, double temp, double eps, double *ctemp_max);
    /* optional, after each E-step's qualities and before the convergence
     * tests; a nonzero return ends the run with that status (DG_ESTOP for a
     * requested stop such as the forced checkpoint) */
    int (*on_iter)(void *user, int nloop, int lloop, double temp, double nfree,
                   double llh, double diff);
    /* HDA: double the grid side, K <- (2 round(K^(1/L)))^L, and move the
     * model to the new basis (dagtm.cxx:2101-2160); *K is in/out.  nloop is
     * the iteration count at the expansion. */
    int (*expand)(void *user, int nloop, double temp, int *K);
} dg_anneal_ops;

/* Hierarchical DA (-H, -k, -E; EXP schedule only, as in the reference).
 * After a temperature that is not the last, while K < Kmax: an armed
 * expansion fires when the llh gained over that temperature (llh -
 * qual_before_temp, dagtm.cxx:1912) is at least eps2 and is cancelled
 * otherwise; when not armed (or just cancelled), next_ctemp is called and
 * the expansion is armed when its value is >= temp * alpha (1938-1977).  A
 * firing calls expand. */
typedef struct {
    int    enable;      /* -H */
    int    K;           /* in: current K; out: K after the expansions */
    int    Kmax;        /* -k */
    double eps2;        /* -E */
    int    nexpanded;   /* out: expansions done */
} dg_hda_state;

/* Runs the loop with ops (hda may be NULL: no HDA).  Returns DG_OK, DG_EINVAL
 * (missing hook, unknown schedule), or the first nonzero status a hook
 * returned; res is filled in every case except DG_EINVAL, with res->beta = 0
 * (beta belongs to the steps; dg_anneal fills in the context's). */
int dg_anneal_ctl(const dg_anneal_params *p, dg_hda_state *hda, const dg_anneal_ops *ops,
                  void *user, dg_anneal_result *res);

/* Critical temperatures of the current state, for AUTO cooling.  For each
 * centre k, with S1 = sum_n R_kn x x^T and S2 = sum_n R_kn^2 x x^T over
 * x = t_n - y_k, the eigenvalues of H = [[2S1 - S2, -S2], [-S2, 2S1 - S2]]
 * at or above limit = (temp - eps) g_k 2 / beta count as 0, and
 * ctemp_k = beta / (2 g_k) * max of them (HUGE_VAL when g_k = 0);
 * *ctemp_max = max_k ctemp_k.  Replaces dagtm_mpi_next_ctemp and its helper
 * (mpicomm.c:561-686) with dagtm_vector_max (comm.c:2149-2161).
 *   gg        K weights g_k, or NULL for the G of the last dg_mstep (zero
 *             before the first), which is what the reference passes
 *             (vggsub, dagtm.cxx:1715 and 1922)
 *   ctemp_k   K outputs, may be NULL
 * R is the responsibility matrix of the last dg_estep (which must be at the
 * current state), computed in FP64 in both precisions.  Cost O(K N D^2) on
 * the device plus 2K symmetric D x D eigenvalue problems on the host. */
int dg_next_ctemp(dg_ctx *ctx, double temp, double eps, const double *gg,
                  double *ctemp_max, double *ctemp_k);

/* ---- many independent instances ------------------------------------------
 * One annealing run per item (any N, D, K, M per item), on one GPU.  Every
 * item gets its own context with its data resident on the device and its own
 * annealing control.  In the mixed precisions with workers = 0 every step of
 * all the instances runs as one sequence of batched launches: each E-pass of
 * a template class (D range, data terms, R terms) is one persistent launch
 * over a device work queue holding the work items of every instance, and
 * the other kernels take the instance from a grid dimension; the sequence of
 * each phase (E-step, M-step) is captured once as a hipGraph and replayed
 * every step, with per-instance parameters and on/off flags in a device
 * table.  workers = -1 runs the same launches without graph capture.  With
 * workers >= 1 (and in FP64) a pool of host threads takes the items from a
 * shared queue in order of decreasing K*N*D and runs their annealing loops on
 * their own streams.  Results do not depend on the mode, the worker count or
 * the order: each item's result equals its own dg_anneal run bit for bit.
 * The reference runs one instance per process; this has no counterpart
 * there. */
typedef struct {
    const double *T;  long N; int D; long ldT;     /* data, N x D */
    const double *FI; int K; int M1; long ldFI;    /* basis, K x M1 */
    const double *Y0; long ldY0;                   /* initial centres, K x D */
    double beta0;
    dg_anneal_params params;
    double *W; long ldW;                           /* out: M1 x D, may be NULL */
    double *Y; long ldY;                           /* out: K x D, may be NULL */
} dg_batch_item;

/* cfg: device, precision and lambda shared by all items; workers: 0 or -1
 * (above; in FP64 they mean 4), else clamped to 1..4.  res[i] and status[i] (DG_OK or the item's error code; a failing item
 * does not stop the others) receive each item's outcome; phase_ms (may be
 * NULL) the wall time of the upload phase and of the annealing phase.
 * Returns DG_OK, or DG_EINVAL / DG_EHIP when the batch could not start. */
int dg_batch_anneal(const dg_config *cfg, int workers, const dg_batch_item *items, int n,
                    dg_anneal_result *res, int *status, double *phase_ms);

/* MAP projection M = R^T X (N x L, ldM >= L) with the responsibilities at
 * the current model state and the last E-step's temperature (an E-step is
 * run first when an M-step has moved the model since).  Computed in FP64
 * in both precisions.  Replaces dagtm.cxx:2347-2355; unlike the reference
 * after a maxloop exit (which reuses R o dist, Q4), R here is always the
 * responsibility matrix. */
int dg_map(dg_ctx *ctx, const double *X, int L, long ldX, double *M,
           long ldM);

/* dg_map without device allocation, for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, result and
 * errors as dg_map; the latent coordinates (K x L) and the projection (N x L)
 * are staged through a fixed device scratch array of 2^21 doubles (16 MiB,
 * held by every GPU the library runs on until the process ends) instead of
 * per-call buffers, so (K + N) * L must not exceed 2^21 (DG_EINVAL
 * otherwise).  Every context on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (from different host threads) race. */
int dg_map_noalloc(dg_ctx *ctx, const double *X, int L, long ldX, double *M,
                   long ldM);

/* Materialises R (K x N, ldR >= N) of the last E-step, in FP64 in both
 * precisions.  Debug / -R output (dagtm.cxx:2400-2414); intended for small
 * problems (K x N doubles are allocated on the device for the call). */
int dg_resp_dense(dg_ctx *ctx, double *R, long ldR);

/* Pass-2 chunk count tuning (no counterpart in the reference).  Times the
 * mixed E-step of the uploaded problem for a few chunk counts within the
 * accuracy limit (chunks of at most 8192 points), keeps the fastest for this
 * context, stores it in *nc_out (0 in DG_FP64, which has no tuned choice)
 * and, when table_path is not NULL, appends "<precision> <D> <K> <nbits(N)>
 * <count>" to that file.  dg_upload and dg_set_basis read the file named by
 * the DG_TUNE_FILE environment variable and use its entry for their shape.
 * Leaves no E-step behind (a dg_estep is needed before dg_mstep). */
int dg_tune(dg_ctx *ctx, const char *table_path, int *nc_out);

/* Copies the model back.  Replaces the h5save_byrow of W, Y and beta
 * (dagtm.cxx:2420-2480).  Any pointer may be NULL. */
int dg_download(dg_ctx *ctx, double *W, long ldW, double *Y, long ldY,
                double *beta);

/* Copies G (K) and RT (K x D, ldRT >= D) of the last E-step (testing). */
int dg_get_stats(dg_ctx *ctx, double *G, double *RT, long ldRT);

/* Measured peaks of the context's device (no counterpart in the reference;
 * the denominators of roofline fractions): FP16 MFMA TFLOP/s
 * (v_mfma_f32_32x32x16_f16 chains), transcendental exponentials per second
 * (v_exp_f32) and vector FP64 FMA TFLOP/s, each from one timed launch of a
 * dependency-chain kernel on every CU.  Any output pointer may be NULL. */
int dg_device_peaks(dg_ctx *ctx, double *mfma_f16_tflops, double *exp_per_s,
                    double *fp64_tflops);

/* Blocks of the last mixed-precision E-pass 2 (no counterpart in the
 * reference; a measurement): *visited (centre group, 32-point block) pairs
 * that it computed, of *total.  Pass 2 leaves out a pair only when every
 * responsibility there rounds to zero in its arithmetic, so results do not
 * depend on the count.  DG_ESTATE for an FP64 context or before an E-step. */
int dg_pass2_blocks(dg_ctx *ctx, long *visited, long *total);

/* Device time of the last dg_estep / dg_mstep, per kernel phase, in ms:
 * t[0] E-pass 1, t[1] E-pass 2, t[2] reductions, t[3] M-step.  Timing is
 * recorded only after dg_set_timing(ctx, 1). */
int dg_set_timing(dg_ctx *ctx, int on);
int dg_get_timing(dg_ctx *ctx, double t[4]);

#ifdef __cplusplus
}
#endif
#endif
