/*
 * dagtm_oracle.h -- CPU FP64 restatement of the reference DA-GTM algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path
 * (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).  It is
 * never linked into, loaded by, or called from the product library
 * (da-gtm_amd/), which fails loudly instead of falling back to the CPU.
 *
 * Every function restates one reference routine in the reference's own
 * step order (file:line into jychoi-hpc/da-gtm given per function).  Matrices
 * are dense row-major doubles with an explicit leading dimension where the
 * reference works on padded gsl_matrix rows (comm.c:1866-1883).
 *
 * Pinning (what checks each function; the tests hold them):
 *   orc_rctg / orc_stp / orc_gbf      data/Oil.out.h5 X.dat, FI.dat (s=2)
 *   orc_dist + orc_resp (tau=1)       Oil.out.h5 R.dat from Y.dat, T.dat, beta
 *   orc_qual                          Oil state llh -0.1692069355836593
 *   orc_resp (tau!=1, zero column)    brute force on tiny inputs + closed forms
 *   orc_mstep pieces / orc_beta       brute force, closed-form beta identity,
 *                                     Oil state N*D/sum(R o dist) = 32.3357...
 *   orc_solve2 / orc_chol_solve       numpy.linalg.solve on well-conditioned A
 *   orc_mt_*                          std::mt19937 known answer (10000th
 *                                     output of seed 5489 = 4123659995)
 *   orc_yrnd2 / orc_bi / orc_ctemp    closed forms on tiny inputs; the GSL
 *                                     seed mapping itself is parity unpinned
 *                                     (GSL is absent from this image)
 */
#ifndef DAGTM_ORACLE_H
#define DAGTM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model set-up (comm.c:992-1185, 1401-1536) ----------------------- */
int orc_rctg(int xdim, int ydim, int zdim, int L, double *X);
int orc_gbf(const double *MU, int M, const double *X, int K, int L,
            double sigma, double *FI);
int orc_stp(int L, int K, int M, double s, double *X, double *FI);

typedef struct { uint32_t mt[624]; int mti; } orc_mt;
void     orc_mt_seed(orc_mt *r, uint32_t s);
uint32_t orc_mt_next(orc_mt *r);
double   orc_mt_uniform(orc_mt *r);

int    orc_yrnd2(const double *mu, const double *sd, int K, int D,
                 unsigned long seed, double *Yrnd);
double orc_bi(const double *Y, int K, int D);
void   orc_mean_sd(const double *T, long N, int D, double *mu, double *sd);
double orc_ctemp(const double *T, long N, int D);
int    orc_stp3_rnd(const double *mu, const double *sd, int D, int L, int K,
                    int M, double s, unsigned long seed, double *X, double *FI,
                    double *W, double *beta);

/* ---- linear algebra ------------------------------------------------- */
int orc_eig_sym(int n, double *A, double *evals, double *V);
int orc_solve2(int n, const double *A, int k, const double *B, double *X);
int orc_solve(int m, int n, const double *A, int k, const double *B,
              double *X);
int orc_chol_solve(int n, const double *A, double ridge, int k,
                   const double *B, double *X, double *jitter_used);
void orc_gemm(int tA, int tB, long m, long n, long k, double alpha,
              const double *A, long lda, const double *B, long ldb,
              double beta, double *C, long ldc);

/* ---- hot path (mpicomm.c:276-500, comm.c:399-456, 560-629) ---------- */
void   orc_dist(const double *Y, long K, const double *T, long N, int D,
                double *Dm);
long   orc_resp(int K, long N, double beta, const double *Dm, int isDA,
                double temp, double *R, double *g, double *h);
double orc_qual(long N, int K, int D, double beta, double *v, int isDA,
                double temp);
double orc_beta(const double *Y, const double *T, double *R, long N, int K,
                int D, double *Dm);

/* ---- one materialised E+M iteration and the annealing loop ----------- */
enum { ORC_SOLVE_PINV = 0, ORC_SOLVE_CHOL = 1 };
enum { ORC_SCHED_EXP = 0, ORC_SCHED_LINEAR = 1, ORC_SCHED_AUTO = 2 };

typedef struct {
    long N; int D, K, M1;
    const double *T;     /* N x D   */
    const double *FI;    /* K x M1  */
    double *Y;           /* K x D   (in/out) */
    double *W;           /* M1 x D  (out)    */
    double beta;         /* in/out  */
    double lambda;       /* ridge: A + (lambda/beta) I ; 0 = reference */
    int solver;          /* ORC_SOLVE_PINV (reference) or ORC_SOLVE_CHOL */
    /* workspace, allocated by orc_model_alloc */
    double *Dm, *R;      /* K x N */
    double *g, *h;       /* N     */
    double *G;           /* K     */
    double *A, *B, *FIR; /* M1 x M1, M1 x D, M1 x N */
    double jitter;       /* last Cholesky jitter (CHOL only) */
    long   nzero;        /* zero columns in the last E-step */
} orc_model;

int  orc_model_alloc(orc_model *m);
void orc_model_free(orc_model *m);

/* E-step at the current distances: resp + both qualities.  Returns
 * the number of zero columns (Q2 fall-back), fills llh and nfree. */
long orc_estep(orc_model *m, int isDA, double temp, double *llh,
               double *nfree);
/* M-step and beta update (dagtm.cxx:1715-1785). */
int  orc_mstep(orc_model *m);

typedef struct {
    int isDA, schedule, maxloop;
    double alpha, delta, eps, start_temp;
    int nloop0;         /* restart: the checkpoint's iter (dagtm.cxx:1197-1212) */
} orc_anneal_params;

typedef struct {
    int nloop;          /* total E+M iterations (dagtm.cxx:1553) */
    int ntemps;         /* temperatures visited */
    int exit_maxloop;   /* 1 when the run ended on maxloop (Q4) */
    double temp, llh, nfree, qual, beta;
} orc_anneal_result;

/* per-iteration record, emitted after the E-step (the reference's progress
 * row, dagtm.cxx:1570-1571): nloop, lloop, temp, nfree, llh, diff and the
 * beta that E-step used */
typedef void (*orc_iter_cb)(void *user, int nloop, int lloop, double temp,
                            double nfree, double llh, double diff,
                            double beta);

int orc_anneal(orc_model *m, const orc_anneal_params *p,
               orc_anneal_result *res, orc_iter_cb cb, void *user);

/* The loop control of dagtm.cxx:1508-2206 alone, over step callbacks, so
 * that it can be driven by scripted qualities.  The steps are called in the
 * reference's order; ev (may be NULL) receives every event. */
enum { ORC_EV_ESTEP = 1, ORC_EV_MSTEP = 2, ORC_EV_CTEMP = 3, ORC_EV_EXPAND = 4 };
typedef struct {
    int  (*estep)(void *user, int isDA, double temp, double *llh, double *nfree);
    int  (*mstep)(void *user);
    int  (*next_ctemp)(void *user, double temp, double eps, double *ctemp_max);
    int  (*on_iter)(void *user, int nloop, int lloop, double temp, double nfree,
                    double llh, double diff);
    int  (*expand)(void *user, int nloop, double temp, int *K);
} orc_steps;

/* HDA (dagtm.cxx:1938-2164, EXP schedule): enable, K (in/out), Kmax, eps2,
 * nexpanded (out) */
typedef struct { int enable, K, Kmax; double eps2; int nexpanded; } orc_hda;

int orc_anneal_ctl(const orc_anneal_params *p, orc_hda *hda, const orc_steps *st,
                   void *user, orc_anneal_result *res);

/* Next critical temperature below temp (dagtm_mpi_next_ctemp and its helper,
 * mpicomm.c:561-675): per centre k, x_n = t_n - y_k, S1 = sum_n R_kn x x^T,
 * S2 = sum_n R_kn^2 x x^T, H = [[2 S1 - S2, -S2], [-S2, 2 S1 - S2]] (2D x 2D);
 * eigenvalues of H at or above limit = (temp - eps) gg_k 2 / beta count as 0
 * (dagtm_vector_max, comm.c:2149-2161) and ctemp_k = beta / (2 gg_k) times the
 * largest; HUGE_VAL when gg_k = 0.  Returns max_k ctemp_k; ctemp_k may be NULL. */
double orc_next_ctemp(const double *T, long N, int D, const double *Y, int K,
                      const double *R, const double *gg, double temp, double beta,
                      double eps, double *ctemp_k);

/* MAP projection M = R^T X (dagtm.cxx:2347-2355): N x L */
void orc_map(const orc_model *m, const double *X, int L, double *Mout);

void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
