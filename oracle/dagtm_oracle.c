/*
 * dagtm_oracle.c -- CPU FP64 restatement of the reference DA-GTM algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the CPU baseline that bench.py
 * times beside the GPU).  Nothing in da-gtm_amd/ links or loads this file.
 *
 * Each function follows one reference routine in its step order; the
 * citations are file:line into jychoi-hpc/da-gtm.  Sums run sequentially in
 * index order like the reference's vec.c loops, so results do not depend on
 * the OpenMP thread count.
 */
#include "dagtm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, ld) ((size_t)(i) * (size_t)(ld) + (size_t)(j))

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

/* ===================================================================== */
/* model set-up                                                          */
/* ===================================================================== */

/* comm.c:1453-1504 (dagtm_rctg): x fastest, then y, then z; coordinates
 * scaled by 2/m and shifted by -1, m = max(dims)-1. */
int orc_rctg(int xdim, int ydim, int zdim, int L, double *X)
{
    int m = xdim;
    if (ydim > m) m = ydim;
    if (zdim > m) m = zdim;
    m -= 1;
    size_t idx = 0;
    if (zdim == 0) {
        if (ydim == 0) {
            for (int x = 0; x < xdim; x++) { X[IDX(idx, 0, L)] = x; idx++; }
        } else {
            for (int y = 0; y < ydim; y++)
                for (int x = 0; x < xdim; x++) {
                    X[IDX(idx, 0, L)] = x; X[IDX(idx, 1, L)] = y; idx++;
                }
        }
    } else {
        for (int z = 0; z < zdim; z++)
            for (int y = 0; y < ydim; y++)
                for (int x = 0; x < xdim; x++) {
                    X[IDX(idx, 0, L)] = x; X[IDX(idx, 1, L)] = y;
                    X[IDX(idx, 2, L)] = z; idx++;
                }
    }
    /* gsl_matrix_scale then gsl_matrix_add_constant */
    double sc = 2 / (double)m;
    for (size_t i = 0; i < idx * (size_t)L; i++) {
        X[i] = X[i] * sc;
        X[i] = X[i] + (-1.0);
    }
    return 0;
}

/* comm.c:1506-1536 (dagtm_gbf): FI = [exp(-dist(X,MU)/(2 sigma^2)), 1] */
int orc_gbf(const double *MU, int M, const double *X, int K, int L,
            double sigma, double *FI)
{
    double *Dm = (double *)malloc(sizeof(double) * (size_t)K * M);
    if (!Dm) return -1;
    orc_dist(X, K, MU, M, L, Dm);
    double sc = -0.5 / pow(sigma, 2);
    for (int i = 0; i < K; i++) {
        for (int j = 0; j < M; j++)
            FI[IDX(i, j, M + 1)] = exp(Dm[IDX(i, j, M)] * sc);
        FI[IDX(i, M, M + 1)] = 1.0;
    }
    free(Dm);
    return 0;
}

/* comm.c:992-1103 (dagtm_stp, GRID_SWEEP branch): latent grid X (K x L),
 * RBF centres MU = rctg(g_M) * g_M/(g_M-1), sigma = s * (MU[1,0]-MU[0,0]). */
int orc_stp(int L, int K, int M, double s, double *X, double *FI)
{
    if (s <= 0 || L < 1 || L > 3) return -1;
    int gx = (int)rint(pow(K, 1 / (double)L));
    int gm = (int)rint(pow(M, 1 / (double)L));
    if (pow(gx, L) != K) return -2;   /* "K is not acceptable" */
    if (pow(gm, L) != M) return -3;   /* "M is not acceptable" */
    double *MU = (double *)malloc(sizeof(double) * (size_t)M * L);
    if (!MU) return -1;
    orc_rctg(gx, L >= 2 ? gx : 0, L >= 3 ? gx : 0, L, X);
    orc_rctg(gm, L >= 2 ? gm : 0, L >= 3 ? gm : 0, L, MU);
    double scale = (double)gm / (gm - 1);
    for (int i = 0; i < M * L; i++) MU[i] *= scale;
    double sigma = s * (MU[IDX(1, 0, L)] - MU[IDX(0, 0, L)]);
    orc_gbf(MU, M, X, K, L, sigma, FI);
    free(MU);
    return 0;
}

/* MT19937 (Matsumoto & Nishimura, init_genrand of 2002).  GSL's
 * gsl_rng_mt19937 maps seed 0 to 4357 and returns u/2^32 from
 * gsl_rng_uniform; the reference uses it through gsl_rng_default
 * (comm.c:1417-1424). */
void orc_mt_seed(orc_mt *r, uint32_t s)
{
    if (s == 0) s = 4357;
    r->mt[0] = s;
    for (int i = 1; i < 624; i++)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->mti = 624;
}

uint32_t orc_mt_next(orc_mt *r)
{
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    const uint32_t UPPER = 0x80000000u, LOWER = 0x7fffffffu;
    if (r->mti >= 624) {
        int kk;
        for (kk = 0; kk < 624 - 397; kk++) {
            uint32_t y = (r->mt[kk] & UPPER) | (r->mt[kk + 1] & LOWER);
            r->mt[kk] = r->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < 623; kk++) {
            uint32_t y = (r->mt[kk] & UPPER) | (r->mt[kk + 1] & LOWER);
            r->mt[kk] = r->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        uint32_t y = (r->mt[623] & UPPER) | (r->mt[0] & LOWER);
        r->mt[623] = r->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
        r->mti = 0;
    }
    uint32_t y = r->mt[r->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

double orc_mt_uniform(orc_mt *r) { return orc_mt_next(r) / 4294967296.0; }

/* comm.c:1401-1448 (dagtm_yrnd2): uniform on mu +- 6 sd, with the
 * u' = ceil(u/1e-15)*1e-15 transform.  Seed 0 is replaced by time(NULL)
 * (comm.c:1419-1423). */
int orc_yrnd2(const double *mu, const double *sd, int K, int D,
              unsigned long seed, double *Yrnd)
{
    orc_mt r;
    if (seed == 0) seed = (unsigned long)time(NULL);
    orc_mt_seed(&r, (uint32_t)seed);
    const double roundoff = 1E-15;
    for (int i = 0; i < K; i++)
        for (int j = 0; j < D; j++) {
            double dev = sd[j] * 6;
            double rnd = ceil(orc_mt_uniform(&r) / roundoff) * roundoff;
            double m = mu[j];
            Yrnd[IDX(i, j, D)] = 2 * dev * rnd - dev + m;
        }
    return 0;
}

/* gsl_stats_mean's published running-mean recurrence (long double). */
static double running_mean(const double *v, size_t n, size_t stride)
{
    long double mean = 0;
    for (size_t i = 0; i < n; i++)
        mean += (v[i * stride] - mean) / (long double)(i + 1);
    return (double)mean;
}

/* comm.c:1161-1185 (dagtm_bi): 2 / mean nearest-neighbour sq. distance. */
double orc_bi(const double *Y, int K, int D)
{
    double *Dm = (double *)malloc(sizeof(double) * (size_t)K * K);
    double *v = (double *)malloc(sizeof(double) * (size_t)K);
    orc_dist(Y, K, Y, K, D, Dm);
    for (int i = 0; i < K; i++) Dm[IDX(i, i, K)] = DBL_MAX;
    for (int j = 0; j < K; j++) {          /* gsl_vector_min per column */
        double mn = Dm[IDX(0, j, K)];
        for (int i = 1; i < K; i++) if (Dm[IDX(i, j, K)] < mn) mn = Dm[IDX(i, j, K)];
        v[j] = mn;
    }
    double m = running_mean(v, (size_t)K, 1);
    if (m == 0) m = 1;
    free(Dm); free(v);
    return 2 / m;
}

/* mpicomm.c:686-730 (dagtm_mpi_mean / dagtm_mpi_sd) on a 1x1 grid. */
void orc_mean_sd(const double *T, long N, int D, double *mu, double *sd)
{
    for (int j = 0; j < D; j++) mu[j] = 0.0;
    for (long i = 0; i < N; i++)
        for (int j = 0; j < D; j++) mu[j] += T[IDX(i, j, D)];
    for (int j = 0; j < D; j++) mu[j] *= 1.0 / N;
    for (int j = 0; j < D; j++) sd[j] = 0.0;
    for (long i = 0; i < N; i++)
        for (int j = 0; j < D; j++) {
            double c = T[IDX(i, j, D)] - mu[j];
            sd[j] += pow(c, 2.0);
        }
    for (int j = 0; j < D; j++) sd[j] = sqrt(sd[j] * (1.0 / (N - 1)));
}

/* mpicomm.c:511-559 (dagtm_mpi_ctemp): tau_c1 = lambda_max(Cov) * beta_T,
 * beta_T = N*D / sum(centred^2).  gsl_eigen_symm is replaced by the Jacobi
 * solver below (only the largest eigenvalue is used). */
double orc_ctemp(const double *T, long N, int D)
{
    double *c = (double *)calloc((size_t)D, sizeof(double));
    double *S = (double *)calloc((size_t)D * D, sizeof(double));
    double *ev = (double *)malloc(sizeof(double) * (size_t)D);
    double *V = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *C = (double *)malloc(sizeof(double) * (size_t)N * D);
    for (long i = 0; i < N; i++)
        for (int j = 0; j < D; j++) c[j] += T[IDX(i, j, D)];
    for (int j = 0; j < D; j++) c[j] *= 1.0 / N;
    for (long i = 0; i < N; i++)       /* dger(-1, ones, c) then add T */
        for (int j = 0; j < D; j++) C[IDX(i, j, D)] = -1.0 * c[j] + T[IDX(i, j, D)];
    orc_gemm(1, 0, D, D, N, 1.0, C, D, C, D, 0.0, S, D);
    orc_eig_sym(D, S, ev, V);
    double lmax = ev[0];
    for (int j = 1; j < D; j++) if (ev[j] > lmax) lmax = ev[j];
    lmax /= N;
    double s = 0;
    for (long i = 0; i < N; i++)
        for (int j = 0; j < D; j++) s += pow(C[IDX(i, j, D)], 2.0);
    double beta = N * D / s;
    free(c); free(S); free(ev); free(V); free(C);
    return lmax * beta;
}

/* comm.c:1118-1156 (dagtm_stp3_rnd): X, FI, random Y, W0 by least squares,
 * beta0 = 0.1 when K > 1000 else dagtm_bi(Yrnd) (Q10: from Yrnd). */
int orc_stp3_rnd(const double *mu, const double *sd, int D, int L, int K,
                 int M, double s, unsigned long seed, double *X, double *FI,
                 double *W, double *beta)
{
    int rc = orc_stp(L, K, M, s, X, FI);
    if (rc) return rc;
    double *Yrnd = (double *)malloc(sizeof(double) * (size_t)K * D);
    if (!Yrnd) return -1;
    orc_yrnd2(mu, sd, K, D, seed, Yrnd);
    rc = orc_solve(K, M + 1, FI, D, Yrnd, W);
    *beta = (K > 1000) ? 0.1 : orc_bi(Yrnd, K, D);
    free(Yrnd);
    return rc;
}

/* ===================================================================== */
/* linear algebra                                                        */
/* ===================================================================== */

/* C = alpha * op(A) op(B) + beta * C, summation over k in index order. */
void orc_gemm(int tA, int tB, long m, long n, long k, double alpha,
              const double *A, long lda, const double *B, long ldb,
              double beta, double *C, long ldc)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < m; i++) {
        double *c = C + (size_t)i * ldc;
        if (beta == 0.0) for (long j = 0; j < n; j++) c[j] = 0.0;
        else if (beta != 1.0) for (long j = 0; j < n; j++) c[j] *= beta;
        if (!tB) {
            for (long l = 0; l < k; l++) {
                double a = alpha * (tA ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)]);
                const double *b = B + (size_t)l * ldb;
                for (long j = 0; j < n; j++) c[j] += a * b[j];
            }
        } else {
            for (long j = 0; j < n; j++) {
                double s = 0;
                const double *b = B + (size_t)j * ldb;
                for (long l = 0; l < k; l++)
                    s += (alpha * (tA ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)])) * b[l];
                c[j] += s;
            }
        }
    }
}

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (A is
 * destroyed).  V holds eigenvectors in columns.  Stands in for GSL's
 * gsl_eigen_symm / SVD of a symmetric matrix. */
int orc_eig_sym(int n, double *A, double *evals, double *V)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[IDX(i, j, n)] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; i++) {
            diag += A[IDX(i, i, n)] * A[IDX(i, i, n)];
            for (int j = i + 1; j < n; j++) off += A[IDX(i, j, n)] * A[IDX(i, j, n)];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[IDX(p, q, n)];
                if (apq == 0.0) continue;
                double app = A[IDX(p, p, n)], aqq = A[IDX(q, q, n)];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) /
                           (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < n; r++) {       /* A <- A J */
                    double arp = A[IDX(r, p, n)], arq = A[IDX(r, q, n)];
                    A[IDX(r, p, n)] = cs * arp - sn * arq;
                    A[IDX(r, q, n)] = sn * arp + cs * arq;
                }
                for (int r = 0; r < n; r++) {       /* A <- J^T A */
                    double apr = A[IDX(p, r, n)], aqr = A[IDX(q, r, n)];
                    A[IDX(p, r, n)] = cs * apr - sn * aqr;
                    A[IDX(q, r, n)] = sn * apr + cs * aqr;
                }
                A[IDX(p, q, n)] = 0.0; A[IDX(q, p, n)] = 0.0;
                for (int r = 0; r < n; r++) {       /* V <- V J */
                    double vrp = V[IDX(r, p, n)], vrq = V[IDX(r, q, n)];
                    V[IDX(r, p, n)] = cs * vrp - sn * vrq;
                    V[IDX(r, q, n)] = sn * vrp + cs * vrq;
                }
            }
    }
    for (int i = 0; i < n; i++) evals[i] = A[IDX(i, i, n)];
    return 0;
}

/* comm.c:1299-1348 (dagtm_solve2): A X = B by SVD, dropping exactly-zero
 * singular values (gsl_linalg_SV_solve).  For symmetric A the SVD is the
 * eigen-decomposition with |lambda|, so X = V diag(1/lambda) V^T B with
 * lambda == 0 dropped. */
int orc_solve2(int n, const double *A, int k, const double *B, double *X)
{
    double *U = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *V = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *ev = (double *)malloc(sizeof(double) * (size_t)n);
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    if (!U || !V || !ev || !w) return -1;
    memcpy(U, A, sizeof(double) * (size_t)n * n);
    orc_eig_sym(n, U, ev, V);
    for (int c = 0; c < k; c++) {
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int r = 0; r < n; r++) s += V[IDX(r, i, n)] * B[IDX(r, c, k)];
            w[i] = (ev[i] != 0.0) ? s / ev[i] : 0.0;
        }
        for (int r = 0; r < n; r++) {
            double s = 0;
            for (int i = 0; i < n; i++) s += V[IDX(r, i, n)] * w[i];
            X[IDX(r, c, k)] = s;
        }
    }
    free(U); free(V); free(ev); free(w);
    return 0;
}

/* comm.c:1247-1294 (dagtm_solve): least squares A X = B through the normal
 * equations (A^T A) X = A^T B and the SVD solve above.  A is m x n. */
int orc_solve(int m, int n, const double *A, int k, const double *B,
              double *X)
{
    double *AtA = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *AtB = (double *)malloc(sizeof(double) * (size_t)n * k);
    orc_gemm(1, 0, n, n, m, 1.0, A, n, A, n, 0.0, AtA, n);
    orc_gemm(1, 0, n, k, m, 1.0, A, n, B, k, 0.0, AtB, k);
    int rc = orc_solve2(n, AtA, k, AtB, X);
    free(AtA); free(AtB);
    return rc;
}

/* Cholesky solve of (A + ridge I) X = B -- the method the HIP M-step uses.
 * On a non-positive pivot the factorisation restarts with an added jitter
 * 1e-13 * trace/n, growing x10 per retry (at most 8 retries). */
int orc_chol_solve(int n, const double *A, double ridge, int k,
                   const double *B, double *X, double *jitter_used)
{
    double *Lm = (double *)malloc(sizeof(double) * (size_t)n * n);
    if (!Lm) return -1;
    double tr = 0;
    for (int i = 0; i < n; i++) tr += A[IDX(i, i, n)];
    double jit = 0.0;
    int ok = 0;
    for (int attempt = 0; attempt < 9 && !ok; attempt++) {
        if (attempt == 1) jit = 1e-13 * tr / n;
        else if (attempt > 1) jit *= 10.0;
        ok = 1;
        for (int j = 0; j < n; j++) {
            for (int i = j; i < n; i++) {
                double s = A[IDX(i, j, n)] + ((i == j) ? ridge + jit : 0.0);
                for (int p = 0; p < j; p++) s -= Lm[IDX(i, p, n)] * Lm[IDX(j, p, n)];
                if (i == j) {
                    if (!(s > 0.0)) { ok = 0; break; }
                    Lm[IDX(j, j, n)] = sqrt(s);
                } else {
                    Lm[IDX(i, j, n)] = s / Lm[IDX(j, j, n)];
                }
            }
            if (!ok) break;
        }
    }
    if (jitter_used) *jitter_used = jit;
    if (!ok) { free(Lm); return 1; }
    for (int c = 0; c < k; c++) {
        for (int i = 0; i < n; i++) {       /* L y = b */
            double s = B[IDX(i, c, k)];
            for (int p = 0; p < i; p++) s -= Lm[IDX(i, p, n)] * X[IDX(p, c, k)];
            X[IDX(i, c, k)] = s / Lm[IDX(i, i, n)];
        }
        for (int i = n - 1; i >= 0; i--) {  /* L^T x = y */
            double s = X[IDX(i, c, k)];
            for (int p = i + 1; p < n; p++) s -= Lm[IDX(p, i, n)] * X[IDX(p, c, k)];
            X[IDX(i, c, k)] = s / Lm[IDX(i, i, n)];
        }
    }
    free(Lm);
    return 0;
}

/* ===================================================================== */
/* hot path                                                              */
/* ===================================================================== */

/* comm.c:399-456 / 560-629 (dagtm_dist, dagtm_dist_ws):
 * Dm = -2 Y T^T (dgemm), then += |y_k|^2 (dger), then += |t_n|^2 (dger).
 * Each dot product runs over d in index order. */
void orc_dist(const double *Y, long K, const double *T, long N, int D,
              double *Dm)
{
    double *yy = (double *)malloc(sizeof(double) * (size_t)K);
    double *tt = (double *)malloc(sizeof(double) * (size_t)N);
    for (long k = 0; k < K; k++) {
        double s = 0;
        for (int d = 0; d < D; d++) s += Y[IDX(k, d, D)] * Y[IDX(k, d, D)];
        yy[k] = s;
    }
    for (long n = 0; n < N; n++) {
        double s = 0;
        for (int d = 0; d < D; d++) s += T[IDX(n, d, D)] * T[IDX(n, d, D)];
        tt[n] = s;
    }
    /* d-major copy of T so the n loop is unit stride */
    const long NB = 256;
    #pragma omp parallel
    {
        double *tb = (double *)malloc(sizeof(double) * (size_t)D * NB);
        #pragma omp for schedule(dynamic, 1)
        for (long n0 = 0; n0 < N; n0 += NB) {
            long nb = (N - n0 < NB) ? N - n0 : NB;
            for (long j = 0; j < nb; j++)
                for (int d = 0; d < D; d++) tb[IDX(d, j, NB)] = T[IDX(n0 + j, d, D)];
            for (long k = 0; k < K; k++) {
                double *c = Dm + IDX(k, n0, N);
                for (long j = 0; j < nb; j++) c[j] = 0.0;
                for (int d = 0; d < D; d++) {
                    double a = -2.0 * Y[IDX(k, d, D)];
                    const double *t = tb + (size_t)d * NB;
                    for (long j = 0; j < nb; j++) c[j] += a * t[j];
                }
                for (long j = 0; j < nb; j++) c[j] += yy[k];
                for (long j = 0; j < nb; j++) c[j] += tt[n0 + j];
            }
        }
        free(tb);
    }
    free(yy); free(tt);
}

/* mpicomm.c:316-414 (dagtm_mpi_resp) on a 1x1 grid, in its pass order:
 * R = D; R *= -beta/2; R = exp(R); g = colsum(R); [DA: R = pow(R, 1/temp);
 * h = colsum(R)]; zero-column fall-back (R column = 1, v = K); R /= v.
 * In EM mode v aliases g, so the fall-back patches g (Q2).
 * Returns the number of zero columns. */
long orc_resp(int K, long N, double beta, const double *Dm, int isDA,
              double temp, double *R, double *g, double *h)
{
    const size_t KN = (size_t)K * (size_t)N;
    const double sc = -beta / 2.0;
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < K; k++)
        for (long n = 0; n < N; n++) {
            size_t i = IDX(k, n, N);
            R[i] = exp(Dm[i] * sc);
        }
    (void)KN;
    /* colsum in row order: V[j] += M[i][j] for i ascending (vec.c:316) */
    #pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) {
        double s = 0.0;
        for (long k = 0; k < K; k++) s += R[IDX(k, n, N)];
        g[n] = s;
    }
    double *v = g;
    if (isDA) {
        const double e = 1.0 / temp;
        #pragma omp parallel for schedule(static)
        for (long k = 0; k < K; k++)
            for (long n = 0; n < N; n++) {
                size_t i = IDX(k, n, N);
                R[i] = pow(R[i], e);
            }
        #pragma omp parallel for schedule(static)
        for (long n = 0; n < N; n++) {
            double s = 0.0;
            for (long k = 0; k < K; k++) s += R[IDX(k, n, N)];
            h[n] = s;
        }
        v = h;
    }
    long nzero = 0;
    for (long n = 0; n < N; n++)
        if (v[n] == 0.0) {
            v[n] = (double)K;
            for (long k = 0; k < K; k++) R[IDX(k, n, N)] = 1.0;
            nzero++;
        }
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < K; k++)
        for (long n = 0; n < N; n++) R[IDX(k, n, N)] = R[IDX(k, n, N)] / v[n];
    return nzero;
}

/* mpicomm.c:463-500 (dagtm_mpi_qual): v <- log v in place (Q3), then
 * qual = [N((D/2) log(beta/2pi) - log K) + (DA ? temp : 1) * sum] / N. */
double orc_qual(long N, int K, int D, double beta, double *v, int isDA,
                double temp)
{
    double sum = 0;
    for (long n = 0; n < N; n++) v[n] = log(v[n]);
    for (long n = 0; n < N; n++) sum += v[n];
    double q = N * ((D / 2.0) * log(beta / (2.0 * M_PI)) - log((double)K));
    if (isDA) q += temp * sum;
    else q += sum;
    return q / (double)N;
}

/* mpicomm.c:276-302 (dagtm_mpi_beta): Dm = dist(Y,T); R o= Dm in place
 * (Q4); beta = N*D / sum(R). */
double orc_beta(const double *Y, const double *T, double *R, long N, int K,
                int D, double *Dm)
{
    orc_dist(Y, K, T, N, D, Dm);
    const size_t KN = (size_t)K * (size_t)N;
    for (size_t i = 0; i < KN; i++) R[i] = R[i] * Dm[i];
    double s = 0;
    for (size_t i = 0; i < KN; i++) s += R[i];
    return ((double)N * D) / s;
}

/* ===================================================================== */
/* iteration + annealing loop                                            */
/* ===================================================================== */

int orc_model_alloc(orc_model *m)
{
    size_t KN = (size_t)m->K * (size_t)m->N;
    m->Dm = (double *)malloc(sizeof(double) * KN);
    m->R = (double *)malloc(sizeof(double) * KN);
    m->g = (double *)malloc(sizeof(double) * (size_t)m->N);
    m->h = (double *)malloc(sizeof(double) * (size_t)m->N);
    m->G = (double *)calloc((size_t)m->K, sizeof(double));   /* zero before the first M-step */
    m->A = (double *)malloc(sizeof(double) * (size_t)m->M1 * m->M1);
    m->B = (double *)malloc(sizeof(double) * (size_t)m->M1 * m->D);
    m->FIR = (double *)malloc(sizeof(double) * (size_t)m->M1 * (size_t)m->N);
    m->jitter = 0; m->nzero = 0;
    if (!m->Dm || !m->R || !m->g || !m->h || !m->G || !m->A || !m->B || !m->FIR)
        return -1;
    /* initial distances, dagtm.cxx:1232 */
    orc_dist(m->Y, m->K, m->T, m->N, m->D, m->Dm);
    return 0;
}

void orc_model_free(orc_model *m)
{
    free(m->Dm); free(m->R); free(m->g); free(m->h); free(m->G);
    free(m->A); free(m->B); free(m->FIR);
    m->Dm = m->R = m->g = m->h = m->G = m->A = m->B = m->FIR = NULL;
}

/* dagtm.cxx:1537-1603: resp then qual (llh from g, -F from h). */
long orc_estep(orc_model *m, int isDA, double temp, double *llh,
               double *nfree)
{
    long nz = orc_resp(m->K, m->N, m->beta, m->Dm, isDA, temp, m->R, m->g, m->h);
    m->nzero = nz;
    *llh = orc_qual(m->N, m->K, m->D, m->beta, m->g, 0, temp);
    if (isDA) *nfree = orc_qual(m->N, m->K, m->D, m->beta, m->h, 1, temp);
    else *nfree = -*llh;
    return nz;
}

/* dagtm.cxx:1715-1785: G = rowsum(R); A = FI^T diag(G) FI; B = (FI^T R) T;
 * W = solve(A, B); Y = FI W; then dagtm_mpi_beta. */
int orc_mstep(orc_model *m)
{
    const long N = m->N; const int K = m->K, M1 = m->M1, D = m->D;
    for (int k = 0; k < K; k++) {                      /* vec.c:298 */
        double s = 0.0;
        const double *r = m->R + IDX(k, 0, N);
        for (long n = 0; n < N; n++) s += r[n];
        m->G[k] = s;
    }
    double *GFI = (double *)malloc(sizeof(double) * (size_t)K * M1);
    for (int k = 0; k < K; k++)                        /* scale_by_row */
        for (int j = 0; j < M1; j++) GFI[IDX(k, j, M1)] = m->FI[IDX(k, j, M1)] * m->G[k];
    orc_gemm(1, 0, M1, M1, K, 1.0, m->FI, M1, GFI, M1, 0.0, m->A, M1);
    free(GFI);
    orc_gemm(1, 0, M1, N, K, 1.0, m->FI, M1, m->R, N, 0.0, m->FIR, N);
    orc_gemm(0, 0, M1, D, N, 1.0, m->FIR, N, m->T, D, 0.0, m->B, D);
    int rc;
    double ridge = (m->lambda > 0) ? m->lambda / m->beta : 0.0;
    if (m->solver == ORC_SOLVE_CHOL) {
        rc = orc_chol_solve(M1, m->A, ridge, D, m->B, m->W, &m->jitter);
    } else {
        if (ridge > 0) {
            double *Ar = (double *)malloc(sizeof(double) * (size_t)M1 * M1);
            memcpy(Ar, m->A, sizeof(double) * (size_t)M1 * M1);
            for (int i = 0; i < M1; i++) Ar[IDX(i, i, M1)] += ridge;
            rc = orc_solve2(M1, Ar, D, m->B, m->W);
            free(Ar);
        } else {
            rc = orc_solve2(M1, m->A, D, m->B, m->W);
        }
    }
    orc_gemm(0, 0, K, D, M1, 1.0, m->FI, M1, m->W, D, 0.0, m->Y, D);
    m->beta = orc_beta(m->Y, m->T, m->R, N, K, D, m->Dm);
    return rc;
}

#define QUEUE_SIZE 50    /* INIT_QUEUE_SIZE, dagtm.cxx:45 */

/* comm.c:1730-1746 (dagtm_queue_*): ring of 50 doubles, mean = seq. sum/50 */
typedef struct { double v[QUEUE_SIZE]; int next; } qqueue;
static void q_reset(qqueue *q, double init) { for (int i = 0; i < QUEUE_SIZE; i++) q->v[i] = init; q->next = 0; }
static void q_push(qqueue *q, double x) { q->v[q->next] = x; q->next = (q->next + 1) % QUEUE_SIZE; }
static double q_mean(const qqueue *q) { double s = 0; for (int i = 0; i < QUEUE_SIZE; i++) s += q->v[i]; return s / QUEUE_SIZE; }
#define ABSV(a) (((a) > (0)) ? (a) : (-(a)))

/* dagtm.cxx:1508-1795 + 1909-2206, control only: the temperature loop
 * (do { ... } while (!isLast)) with
 *   - clamp temp < 1 to 1 and mark the last pass (1510-1514);
 *   - the quality queue refilled with DBL_MAX per temperature (1521) and
 *     qual_before_temp = llh (1523);
 *   - per iteration (1531-1795): resp + qual (the E-step), qual_old = qual,
 *     lloop++, nloop++, qual = nfree (DA) / llh (EM), diff = qual -
 *     qual_old (qual starts at 0: Q7), the progress row and checkpoint
 *     (on_iter), |diff| < eps -> break (1689-1693), nloop > 50 -> push qual
 *     and break when |mean after - mean before| < eps (1696-1709), the
 *     M-step, maxloop -> last pass (1790-1795);
 *   - after the temperature, if DA and not the last: AUTO (1918-1934), EXP
 *     with the HDA arm/cancel/fire rule (1936-2166), LINEAR (2168-2170);
 *     else leave the loop (2176-2179). */
int orc_anneal_ctl(const orc_anneal_params *p, orc_hda *hda, const orc_steps *st,
                   void *user, orc_anneal_result *res)
{
    double temp = p->isDA ? p->start_temp : 1.0;
    int isLast = 0, nloop = p->nloop0, ntemps = 0, exit_maxloop = 0, rc = 0;
    int bReadyToExpand = 0, bDoExpand = 0;
    double qual = 0, qual_old = 0, llh = 0, nfree = 0, diff;
    const int hda_on = hda && hda->enable;
    qqueue Q;
    if (hda) hda->nexpanded = 0;
    do {
        if (temp < 1.0) { temp = 1.0; isLast = 1; }
        q_reset(&Q, DBL_MAX);
        ntemps++;
        double qual_before_temp = llh;
        int lloop = 0;
        while (1) {
            if ((rc = st->estep(user, p->isDA, temp, &llh, &nfree))) goto out;
            qual_old = qual;
            lloop++; nloop++;
            qual = p->isDA ? nfree : llh;
            diff = qual - qual_old;
            if (st->on_iter && (rc = st->on_iter(user, nloop, lloop, temp, nfree, llh, diff)))
                goto out;
            if (ABSV(diff) < p->eps) break;
            if (nloop > QUEUE_SIZE) {
                double old = q_mean(&Q);
                q_push(&Q, qual);
                double avg = q_mean(&Q);
                double da = avg - old;
                if (ABSV(da) < p->eps) break;
            }
            if ((rc = st->mstep(user))) goto out;
            if (p->maxloop && nloop >= p->maxloop) { isLast = 1; exit_maxloop = 1; break; }
        }
        double qual_diff_at_temp = llh - qual_before_temp;
        if (p->isDA && !isLast) {
            if (p->schedule == ORC_SCHED_AUTO) {
                double ct;
                if ((rc = st->next_ctemp(user, temp, p->eps, &ct))) goto out;
                if (ct >= temp) ct = temp * p->alpha;
                temp = ct;
            } else if (p->schedule == ORC_SCHED_EXP) {
                if (hda_on && hda->K < hda->Kmax) {
                    if (bReadyToExpand) {
                        if (qual_diff_at_temp < hda->eps2) bReadyToExpand = 0;
                        else bDoExpand = 1;
                    }
                    if (!bReadyToExpand) {
                        double ct;
                        if ((rc = st->next_ctemp(user, temp, p->eps, &ct))) goto out;
                        if (ct >= temp * p->alpha) bReadyToExpand = 1;
                    }
                    if (bDoExpand) {
                        bDoExpand = 0;
                        bReadyToExpand = 0;
                        hda->nexpanded++;
                        if ((rc = st->expand(user, nloop, temp, &hda->K))) goto out;
                    }
                }
                temp *= p->alpha;
            } else if (p->schedule == ORC_SCHED_LINEAR) {
                temp -= p->delta;
            }
        } else {
            break;
        }
    } while (!isLast);
out:
    res->nloop = nloop; res->ntemps = ntemps; res->exit_maxloop = exit_maxloop;
    res->temp = temp; res->llh = llh; res->nfree = nfree; res->qual = qual;
    res->beta = 0;
    return rc;
}

/* orc_anneal's steps: the materialised model */
typedef struct { orc_model *m; orc_iter_cb cb; void *user; } orc_model_steps;
static int ms_estep(void *u, int isDA, double temp, double *llh, double *nfree)
{
    orc_estep(((orc_model_steps *)u)->m, isDA, temp, llh, nfree);
    return 0;
}
static int ms_mstep(void *u) { orc_mstep(((orc_model_steps *)u)->m); return 0; }
static int ms_ctemp(void *u, double temp, double eps, double *ct)
{
    /* R is the last E-step's, G the last M-step's (the reference's vggsub is
     * refreshed only after the convergence test, dagtm.cxx:1715) */
    orc_model *m = ((orc_model_steps *)u)->m;
    *ct = orc_next_ctemp(m->T, m->N, m->D, m->Y, m->K, m->R, m->G, temp, m->beta, eps, NULL);
    return 0;
}
static int ms_iter(void *u, int nloop, int lloop, double temp, double nfree, double llh,
                   double diff)
{
    orc_model_steps *s = (orc_model_steps *)u;
    if (s->cb) s->cb(s->user, nloop, lloop, temp, nfree, llh, diff, s->m->beta);
    return 0;
}

/* The whole loop on the materialised model (no HDA). */
int orc_anneal(orc_model *m, const orc_anneal_params *p,
               orc_anneal_result *res, orc_iter_cb cb, void *user)
{
    orc_model_steps s = {m, cb, user};
    orc_steps st = {ms_estep, ms_mstep, ms_ctemp, ms_iter, NULL};
    int rc = orc_anneal_ctl(p, NULL, &st, &s, res);
    res->beta = m->beta;
    return rc;
}

double orc_next_ctemp(const double *T, long N, int D, const double *Y, int K,
                      const double *R, const double *gg, double temp, double beta,
                      double eps, double *ctemp_k)
{
    const double GG_MIN = 1e-250;                 /* mpicomm.c:617-628 */
    const int n2 = 2 * D;
    double *H = (double *)malloc(sizeof(double) * (size_t)n2 * n2);
    double *lam = (double *)malloc(sizeof(double) * (size_t)n2);
    double *S1 = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *S2 = (double *)malloc(sizeof(double) * (size_t)D * D);
    double *xs = (double *)malloc(sizeof(double) * (size_t)D);
    double *x2 = (double *)malloc(sizeof(double) * (size_t)D);
    double *V = (double *)malloc(sizeof(double) * (size_t)n2 * n2);
    double cmax = -HUGE_VAL;
    for (int k = 0; k < K; k++) {
        double ck;
        const double g = gg[k];
        if (g == 0.0) {
            ck = HUGE_VAL;
        } else {
            memset(S1, 0, sizeof(double) * (size_t)D * D);
            memset(S2, 0, sizeof(double) * (size_t)D * D);
            for (long n = 0; n < N; n++) {
                const double r = R[(size_t)k * N + n];
                const double sr = sqrt(r);
                for (int a = 0; a < D; a++) {
                    xs[a] = (T[IDX(n, a, D)] - Y[IDX(k, a, D)]) * sr;   /* sqrt(r) x */
                    x2[a] = xs[a] * sr;                                    /* r x */
                }
                for (int a = 0; a < D; a++)
                    for (int b = 0; b < D; b++) {
                        S1[IDX(a, b, D)] += 2.0 * xs[a] * xs[b];
                        S2[IDX(a, b, D)] += x2[a] * x2[b];
                    }
            }
            for (int a = 0; a < D; a++)
                for (int b = 0; b < D; b++) {
                    const double s11 = S1[IDX(a, b, D)] - S2[IDX(a, b, D)];
                    const double s12 = -S2[IDX(a, b, D)];
                    H[IDX(a, b, n2)] = s11;
                    H[IDX(a + D, b + D, n2)] = s11;
                    H[IDX(a, b + D, n2)] = s12;
                    H[IDX(a + D, b, n2)] = s12;
                }
            if (g < GG_MIN)
                for (long e = 0; e < (long)n2 * n2; e++) H[e] /= GG_MIN;
            orc_eig_sym(n2, H, lam, V);
            if (g < GG_MIN)
                for (int i = 0; i < n2; i++) lam[i] *= GG_MIN;
            const double limit = (temp - eps) * g * 2.0 / beta;
            double lmax = -HUGE_VAL;
            for (int i = 0; i < n2; i++) {
                const double l = (lam[i] >= limit) ? 0.0 : lam[i];
                if (l > lmax) lmax = l;
            }
            ck = beta / 2.0 / g * lmax;
        }
        if (ctemp_k) ctemp_k[k] = ck;
        if (ck > cmax) cmax = ck;
    }
    free(H); free(lam); free(S1); free(S2); free(xs); free(x2); free(V);
    return cmax;
}

/* dagtm.cxx:2347-2355: M = R^T X (N x L) with the R left by the last
 * iteration (R o Dm after a maxloop exit, Q4). */
void orc_map(const orc_model *m, const double *X, int L, double *Mout)
{
    orc_gemm(1, 0, m->N, L, m->K, 1.0, m->R, m->N, X, L, 0.0, Mout, L);
}
