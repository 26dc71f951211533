// Internal declarations shared by the HIP translation units of libdagtm_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dagtm_gpu.h"

// exp(x) of an FP64 x below this value rounds to +0 (2^-1075 is the midpoint
// under the smallest subnormal).  The reference applies exp before pow
// (mpicomm.c:334, 350), so an element whose exponent -beta*d/2 falls below
// it contributes nothing at any temperature; both precisions reproduce that.
constexpr double kExpUnderflowX = -745.1332191019412;

struct dg_ctx {
    int device = 0;
    dg_config cfg{};
    hipStream_t stream = nullptr;
    std::string err;

    // problem shape
    long N = 0;
    int D = 0, K = 0, M1 = 0;
    bool loaded = false;
    bool have_estep = false;       // G/RT/pm/pl valid for the current Y
    bool have_w = false;           // W holds the last M-step's (or a caller's) weights
    bool last_isDA = true;
    double last_temp = 1.0;
    double beta = 0.0;
    double sum_tt = 0.0;           // sum_n |t_n|^2 (FP64)

    // ---- FP64 state (always present) ----
    int ldt = 0;                   // padded row stride of T64 / Y64
    double *T64 = nullptr;         // N x ldt
    double *tt = nullptr;          // N      |t_n|^2
    double *FI = nullptr;          // K x M1
    double *Y64 = nullptr;         // K x ldt
    double *yy = nullptr;          // K      |y_k|^2
    double *W = nullptr;           // M1 x D
    double *A = nullptr;           // M1 x M1 (lower triangle holds the factor after a solve)
    double *A0 = nullptr;          // M1 x M1 unfactored copy
    double *Bm = nullptr;          // M1 x D
    double *G = nullptr;           // K
    double *Gm = nullptr;          // K   G used by the last M-step (zero before one)
    double *RT = nullptr;          // K x D
    double *pm = nullptr;          // N   per-point max exponent x_max (pass 1)
    double *pl = nullptr;          // N   per-point log S2 (pass 1)
    double *logg = nullptr;        // N
    double *logh = nullptr;        // N
    double *partRT = nullptr;      // NC x K x D
    double *partG = nullptr;       // NC x K
    int NC = 1;                    // N chunks of E-pass 2
    double *partAB = nullptr;      // nchAB x (M1*M1 + M1*D) split-K partials of A, B
    int nchAB = 1;                 // centre chunks of the A/B reduction
    double *scal = nullptr;        // device scalars (see kScal*)
    double *h_scal = nullptr;      // pinned host mirror

    // ---- mixed-precision state ----
    int mixed_ready = 0;
    int ldh = 0;                   // padded row stride of the 16-bit operands
    int t_exact = 0;               // T exactly representable in the 16-bit type
    float t_scale = 1.f;           // power-of-two scale applied to T and Y
    void *mix = nullptr;           // padded shapes and split operands (dg_mixed.hip)
    void *T16 = nullptr;           // TT x Np x ldh   FP16 terms of T
    float *tt32 = nullptr;         // Np   |t'|^2
    float *yy32 = nullptr;         // Kp   |y'|^2
    float *pP32 = nullptr;         // 2 Np: per 32-point block, 32 pass-2 exponent
                                   // offsets then 32 underflow thresholds
    float *partRT32 = nullptr;     // NC x D x K
    float *partG32 = nullptr;      // NC x K
    int *perm = nullptr;           // Np: input position of the mixed path's point n
    int *p1start = nullptr;        // Np/128: the tile each pass-1 workgroup starts from
                                   // (points ordered by their best centre tile), or null

    int nc_override = 0;           // pass-2 chunk count chosen by the tuner (0: built-in)
    double *map_scratch = nullptr; // this device's dg_map_noalloc scratch (static device array)

    // either mixed-precision E-step
    bool mixed() const { return cfg.precision != DG_FP64; }

    // timing
    int timing = 0;
    hipEvent_t ev[6] = {};
    double t_ms[4] = {0, 0, 0, 0};
};

// scalar slots in ctx->scal
enum {
    kScalSumLogG = 0,   // sum_n log g_n
    kScalSumLogH = 1,   // sum_n log h_n
    kScalNZero = 2,     // zero columns (as double)
    kScalBetaSum = 3,   // sum R o dist(Y_new, T)
    kScalJitter = 4,    // jitter used by the last factorisation
    kScalSolveOK = 5,   // 1 if the factorisation succeeded
    kScalCount = 8
};

// ---- batched instances ----------------------------------------------------
// Batched light kernels run one flat grid over the instances of their list,
// each instance with exactly its own workgroup count (bk_count): a device map
// per launch gives workgroup g its (table index, workgroup within instance).
enum BKind {
    BK_MAKE_AK, BK_REDUCE32, BK_ZERO_FIX, BK_COPY_G, BK_PART_AB, BK_SUM_AB, BK_TRSM_LDS,
    BK_PHIW, BK_ROWNORM, BK_BETA_PART, BK_ABSMAX_PART, BK_MAKE_Y16, BK_BLOCK_LISTS, BK_COUNT
};
// tile sizes the flat grids are counted in (checked against the kernels)
constexpr int kBRK = 64, kBRD = 16;     // k_reduce32: centres x dims per workgroup
constexpr int kBGT = 64;                 // M-step products: output tile
constexpr int kBBetaWG = 64;             // k_beta_part: centres per workgroup
constexpr int kBAmaxWG = 256;            // k_absmax_part: most workgroups
// One instance of a batched step, as the batched kernels read it from a
// device table: the shape, the device buffers of its context and the
// parameters of the current step.  Built on the host from each context.
struct BInst {
    // shape
    long N, Np, CL;
    int D, K, Kp, M1, ldt, ldh, NC, nchAB, KS, TT;
    // FP64 state (dg_ctx)
    const double *T64, *tt, *FI;
    double *Y64, *yy, *W, *A, *A0, *Bm, *G, *Gm, *RT, *pm, *pl, *logg, *logh, *partAB, *scal;
    // mixed state
    const _Float16 *T16, *T16t;
    _Float16 *Y16;
    float *tt32, *yy32, *aK32, *pP32, *blkThr, *partRT32, *partG32, *gmx, *bnd;
    const int *perm;
    const int *p1start;            // pass-1 start tile per point tile (null: tile 0)
    int screen;                    // pass 1 screens tiles
    int2 *brange;                  // pass-2 block runs per (chunk, group)
    float skipBelow;
    double *mscal, *partLog, *amaxPart;
    // step parameters
    float beta_f, it;
    int isDA, two, do_e, do_m;
    double logK, ridge;
    // first work item of this instance in its pass-1 / pass-2 launch
    int item_e1, item_e2;
};

__host__ __device__ inline int bk_cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// workgroups of instance I in the flat grid of a kind
__host__ __device__ inline int bk_count(const BInst &I, int kind) {
    const auto cdiv = bk_cdiv;
    switch (kind) {
        case BK_MAKE_AK: return cdiv(I.Kp, 256);
        case BK_REDUCE32: return cdiv(I.K, kBRK) * cdiv(I.D, kBRD);
        case BK_ZERO_FIX: return I.D;
        case BK_COPY_G: return cdiv(I.K, 256);
        case BK_PART_AB: return cdiv(I.M1, kBGT) * cdiv(I.M1 + I.D, kBGT) * I.nchAB;
        case BK_SUM_AB: return cdiv((long)I.M1 * I.M1 + (long)I.M1 * I.D, 256);
        case BK_TRSM_LDS: return cdiv(I.D, 32);
        case BK_PHIW: return cdiv(I.K, kBGT) * cdiv(I.D, kBGT);
        case BK_ROWNORM: return cdiv(I.K, 256);
        case BK_BETA_PART: return cdiv(I.K, kBBetaWG);
        case BK_ABSMAX_PART: {
            const long w = ((long)I.K * I.D + 4095) / 4096;
            return (int)(w < kBAmaxWG ? w : kBAmaxWG);
        }
        case BK_MAKE_Y16: return cdiv(I.Kp, 4);
        case BK_BLOCK_LISTS: return (I.Kp / 128) * I.NC;
        default: return 0;
    }
}

// instances of one E-pass template class in a batched launch: their table
// indices are list[first .. first + n), with items_e1 / items_e2 work items
// in all; index numbers the class's two work-queue counters
struct BClass {
    int KS, TT, D, two, rlo;
    int first, n, items_e1, items_e2, index;
    // flat grids of the chunk reduction and the zero-column fix over the
    // class's instances: sizes and device maps
    int wg_reduce32, wg_zero_fix, wg_lists;
    const int2 *map_reduce32, *map_zero_fix, *map_lists;
};

// flat grids of the batched launches over their whole lists: sizes and
// device maps (entry g = {table index, workgroup within that instance})
struct BGrids {
    int wg[BK_COUNT] = {};
    const int2 *map[BK_COUNT] = {};
};

// batched E-step pieces (dg_mixed.hip)
int dg_batch_fill_mixed(const dg_ctx *c, BInst *I);
int dg_batch_items(const BInst &I, int *e1, int *e2);
// make_ak of the n listed instances on st; then per class, on side stream
// (class index) % nside: pass 1, the log sums, pass 2, the chunk reductions
// and the zero-column fix; st waits for every side stream at the end.
// counters holds 4 ncls ints of device memory, ev 1 + nside events.
int dg_batch_estep(hipStream_t st, const BInst *tab, const int *list, int n, const BClass *cls,
                   int ncls, int *counters, const BGrids &gr, int cus, hipStream_t *side,
                   int nside, hipEvent_t *ev);
// the FP16 split of the new centres (dg_mixed_prepare_y) of the listed instances
int dg_batch_prepare_y(hipStream_t st, const BInst *tab, const int *list, int n,
                       const BGrids &gr);
// out[i][kScalCount] = the scalars of table entry i, for i < n
int dg_batch_gather_scal(hipStream_t st, const BInst *tab, int n, double *out);

// M-step of the listed instances (dg_mstep.hip): list holds n table indices,
// the nlds instances with M + 1 <= 126 (LDS Cholesky) first
int dg_batch_mstep(hipStream_t st, const BInst *tab, const int *list, int n, int nlds,
                   const BGrids &gr);

// annealing of n uploaded mixed-precision contexts as batched launches per
// step (dg_batch.cpp); graphs: capture each phase once as a hipGraph
int dg_batch_run_mixed(std::vector<dg_ctx *> &ctx, const dg_anneal_params *params,
                       dg_anneal_result *res, int *status, bool graphs);

int dg_fail(dg_ctx *ctx, int code, const std::string &msg);
int dg_hip_check(dg_ctx *ctx, hipError_t e, const char *what);

#define DG_HIP(ctx, call)                                             \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return dg_hip_check((ctx), e_, #call);  \
    } while (0)

// ---- FP64 kernels (dg_fp64.hip) ----
int dg_estep_fp64(dg_ctx *ctx, int isDA, double temp);
int dg_resp_dense_fp64(dg_ctx *ctx, double *dR, long ldR);
int dg_map_fp64(dg_ctx *ctx, const double *dX, int L, double *dM);
// S1, S2 of centres k0 .. k0+KB-1 into dS (KB x 2 x D x D, lower triangle
// valid) from nc N-chunk partials in dPart (dS == dPart when nc == 1).
int dg_ctemp_moments_fp64(dg_ctx *ctx, int k0, int KB, int nc, double *dPart, double *dS);

// ---- host eigenvalues (dg_eig.cpp) ----
extern "C" int dg_sym_eigvals(int n, const double *A, long lda, double *w);

// ---- mixed kernels (dg_mixed.hip) ----
int dg_mixed_prepare_t(dg_ctx *ctx);
// Built-in pass-2 chunk count for the context's shape; *lo / *hi (may be
// NULL) receive the smallest count that keeps chunks <= 8192 points and the
// largest useful one (both multiples of 8).
long dg_mixed_nc_limits(const dg_ctx *ctx, long *lo, long *hi);
int dg_mixed_nc(const dg_ctx *ctx);   // chunk count in use
int dg_mixed_prepare_y(dg_ctx *ctx);
// orders the mixed path's points by the centre tile holding their largest
// responsibility under the current centres (one pass-1 sweep), then rebuilds
// the FP16 operands in that order; DG_MIXED_ORDER=0 keeps the input order
int dg_mixed_order(dg_ctx *ctx);
// (centre group, block) pairs the last pass 2 visited, and all of them
int dg_mixed_pass2_blocks(dg_ctx *ctx, long *visited, long *total);
int dg_estep_mixed(dg_ctx *ctx, int isDA, double temp);
void dg_mixed_free(dg_ctx *ctx);

// ---- M-step (dg_mstep.hip) ----
// The M-step kernels; for M + 1 > 126 the factorisation uses A0 + (ridge +
// jitter) I (M + 1 <= 126 retries with a jitter inside its one kernel).
int dg_mstep_run(dg_ctx *ctx, double jitter);
// tr(A0) into scal[kScalJitter] (for the retries of M + 1 > 126)
int dg_mstep_trace(dg_ctx *ctx);
// Y = FI W and |y_k|^2 from the context's FI and W.
int dg_centres_from_w(dg_ctx *ctx);
int dg_rownorms(dg_ctx *ctx, const double *X, long rows, int ld, int D, double *out);
// timed dependency-chain kernels (dg_peaks.hip): FP16 MFMA, v_exp_f32, FP64 FMA
int dg_peaks_run(dg_ctx *ctx, int wgs, double *mfma_tflops, double *exp_per_s,
                 double *f64_tflops);
// dst[r][d] = src[r lds + d] for d < D, 0 for D <= d < ld (device arrays)
int dg_pad_rows(dg_ctx *ctx, const double *src, long rows, int D, long lds, int ld, double *dst);
int dg_reduce_logs(dg_ctx *ctx);
int dg_sum_into(dg_ctx *ctx, const double *v, long n, int slot);
