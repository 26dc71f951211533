/*
 * A C caller of libdagtm_gpu that passes every matrix the way the reference's
 * gsl_matrix does: (m->data, m->size1, m->size2, m->tda), rows padded to 64
 * bytes (gsl_matrix_alloc, comm.c:1866-1883), as the swap shown in
 * INTEGRATION.md does.  It runs one problem twice, once with dense rows and
 * once with padded rows whose padding holds NaN, through dg_upload, dg_estep,
 * dg_get_stats, dg_mstep, dg_download, dg_set_state, dg_map and dg_set_basis,
 * and checks that every output agrees bit for bit and that no padding
 * element of an output was written.
 *
 *   tda_caller <precision 0|1|2> <N> <D> <K> <M1> <binary 0|1> <seed>
 *
 * Prints one line; exit status 0 when everything agrees.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dagtm_gpu.h"

typedef struct { size_t size1, size2, tda; double *data; } mat;   /* gsl_matrix's fields */

static mat mat_alloc(size_t r, size_t c, int padded)
{
    mat m = {r, c, padded ? (c * 8 + 63) / 64 * 8 : c, NULL};
    if (padded && m.tda == c) m.tda = c + 8;          /* always exercise a stride > cols */
    m.data = (double *)malloc(sizeof(double) * r * m.tda);
    for (size_t i = 0; i < r * m.tda; i++) m.data[i] = NAN;
    return m;
}

static void mat_copy(mat *dst, const mat *src)
{
    for (size_t i = 0; i < src->size1; i++)
        memcpy(dst->data + i * dst->tda, src->data + i * src->tda, src->size2 * sizeof(double));
}

/* bitwise equality of the data parts; padding of b must still be NaN */
static int mat_same(const char *what, const mat *a, const mat *b)
{
    for (size_t i = 0; i < a->size1; i++) {
        if (memcmp(a->data + i * a->tda, b->data + i * b->tda, a->size2 * sizeof(double))) {
            printf("MISMATCH %s row %zu\n", what, i);
            return 0;
        }
        for (size_t j = b->size2; j < b->tda; j++)
            if (!isnan(b->data[i * b->tda + j])) {
                printf("PADDING WRITTEN %s row %zu col %zu\n", what, i, j);
                return 0;
            }
    }
    return 1;
}

static uint64_t rng_state;
static double urand(void)                              /* splitmix64 -> [0, 1) */
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
    mat T, FI, Y0, X, FI2;                              /* inputs */
    mat G, RT, W, Y, M, Y2, W2;                         /* outputs */
    double llh[3], nfree[3], beta[2];
} run_t;

#define CK(call) do { int rc_ = (call); if (rc_) { \
    printf("FAIL %s: %d %s\n", #call, rc_, dg_last_error(ctx)); exit(1); } } while (0)

static void run(int precision, run_t *r, int L)
{
    dg_ctx *ctx = NULL;
    dg_config cfg = {0, precision, 1e-3};
    if (dg_create(&cfg, &ctx)) { printf("FAIL dg_create\n"); exit(1); }
    const int N = (int)r->T.size1, D = (int)r->T.size2, K = (int)r->FI.size1;
    const int M1 = (int)r->FI.size2;
    dg_estep_out eo;
    CK(dg_upload(ctx, r->T.data, N, D, (long)r->T.tda, r->FI.data, K, M1, (long)r->FI.tda,
                 r->Y0.data, (long)r->Y0.tda, 8.0));
    CK(dg_estep(ctx, 1, 2.0, &eo));
    r->llh[0] = eo.llh; r->nfree[0] = eo.nfree;
    double *G = r->G.data;
    CK(dg_get_stats(ctx, G, r->RT.data, (long)r->RT.tda));
    CK(dg_mstep(ctx, &r->beta[0]));
    CK(dg_download(ctx, r->W.data, (long)r->W.tda, r->Y.data, (long)r->Y.tda, NULL));
    /* restart-style state from the downloaded (padded) Y */
    CK(dg_set_state(ctx, r->Y.data, (long)r->Y.tda, r->beta[0]));
    CK(dg_estep(ctx, 1, 1.5, &eo));
    r->llh[1] = eo.llh; r->nfree[1] = eo.nfree;
    CK(dg_map(ctx, r->X.data, L, (long)r->X.tda, r->M.data, (long)r->M.tda));
    CK(dg_mstep(ctx, &r->beta[1]));
    /* an HDA-style new basis with the W just computed */
    CK(dg_download(ctx, r->W2.data, (long)r->W2.tda, NULL, 0, NULL));
    CK(dg_set_basis(ctx, r->FI2.data, (int)r->FI2.size1, M1, (long)r->FI2.tda, r->W2.data,
                    (long)r->W2.tda));
    CK(dg_estep(ctx, 1, 1.2, &eo));
    r->llh[2] = eo.llh; r->nfree[2] = eo.nfree;
    CK(dg_download(ctx, NULL, 0, r->Y2.data, (long)r->Y2.tda, NULL));
    dg_destroy(ctx);
}

static void alloc_run(run_t *r, int N, int D, int K, int M1, int K2, int L, int padded)
{
    r->T = mat_alloc(N, D, padded); r->FI = mat_alloc(K, M1, padded);
    r->Y0 = mat_alloc(K, D, padded); r->X = mat_alloc(K, L, padded);
    r->FI2 = mat_alloc(K2, M1, padded);
    r->G = mat_alloc(K, 1, 0); r->RT = mat_alloc(K, D, padded); r->W = mat_alloc(M1, D, padded);
    r->Y = mat_alloc(K, D, padded); r->M = mat_alloc(N, L, padded);
    r->Y2 = mat_alloc(K2, D, padded); r->W2 = mat_alloc(M1, D, padded);
}

int main(int argc, char **argv)
{
    if (argc != 8) {
        fprintf(stderr, "usage: %s precision N D K M1 binary seed\n", argv[0]);
        return 2;
    }
    const int precision = atoi(argv[1]), N = atoi(argv[2]), D = atoi(argv[3]);
    const int K = atoi(argv[4]), M1 = atoi(argv[5]), binary = atoi(argv[6]);
    rng_state = (uint64_t)atoll(argv[7]);
    const int L = 3, K2 = K + K / 2;
    run_t a, b;
    alloc_run(&a, N, D, K, M1, K2, L, 0);
    alloc_run(&b, N, D, K, M1, K2, L, 1);
    /* inputs: data (binary or float), a positive basis with a bias column,
     * centres inside the data's range, latent coordinates */
    for (int n = 0; n < N; n++)
        for (int d = 0; d < D; d++)
            a.T.data[n * D + d] = binary ? (double)(urand() < 0.3) : 4.0 * urand() - 2.0;
    for (int k = 0; k < K2; k++)
        for (int m = 0; m < M1; m++) {
            const double v = (m == M1 - 1) ? 1.0 : urand();
            a.FI2.data[k * M1 + m] = v;
            if (k < K) a.FI.data[k * M1 + m] = v;
        }
    for (int k = 0; k < K; k++) {
        const int n = (int)(urand() * N);
        for (int d = 0; d < D; d++)
            a.Y0.data[k * D + d] = 0.5 * a.T.data[n * D + d] + 0.5 * (binary ? 0.3 : 0.0);
        for (int l = 0; l < L; l++) a.X.data[k * L + l] = 2.0 * urand() - 1.0;
    }
    mat_copy(&b.T, &a.T); mat_copy(&b.FI, &a.FI); mat_copy(&b.Y0, &a.Y0);
    mat_copy(&b.X, &a.X); mat_copy(&b.FI2, &a.FI2);
    run(precision, &a, L);
    run(precision, &b, L);
    int ok = mat_same("G", &a.G, &b.G) && mat_same("RT", &a.RT, &b.RT) &&
             mat_same("W", &a.W, &b.W) && mat_same("Y", &a.Y, &b.Y) &&
             mat_same("M", &a.M, &b.M) && mat_same("W2", &a.W2, &b.W2) &&
             mat_same("Y2", &a.Y2, &b.Y2);
    for (int i = 0; i < 3 && ok; i++)
        if (memcmp(&a.llh[i], &b.llh[i], sizeof(double)) ||
            memcmp(&a.nfree[i], &b.nfree[i], sizeof(double))) {
            printf("MISMATCH llh/nfree %d\n", i);
            ok = 0;
        }
    for (int i = 0; i < 2 && ok; i++)
        if (memcmp(&a.beta[i], &b.beta[i], sizeof(double))) { printf("MISMATCH beta %d\n", i); ok = 0; }
    if (ok && !(isfinite(a.llh[2]) && a.beta[1] > 0)) { printf("NOT FINITE\n"); ok = 0; }
    if (ok)
        printf("tda OK precision %d N %d D %d (tda %zu) K %d M1 %d (tda %zu) llh %.15g\n", precision,
               N, D, b.T.tda, K, M1, b.FI.tda, a.llh[2]);
    return ok ? 0 : 1;
}
