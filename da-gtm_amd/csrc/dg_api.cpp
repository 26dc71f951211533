// C ABI of libdagtm_gpu: context, data movement, and the host-side control
// of one DA-GTM iteration and of the annealing loop.  Kernels live in
// dg_fp64.hip, dg_mixed.hip and dg_mstep.hip.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "dg_internal.h"

#ifndef DG_VERSION_STRING
#define DG_VERSION_STRING "dagtm_gpu gfx950"
#endif

int dg_fail(dg_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int dg_hip_check(dg_ctx *ctx, hipError_t e, const char *what) {
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return DG_EHIP;
}

namespace {

template <typename T>
int dalloc(dg_ctx *c, T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return dg_fail(c, DG_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return DG_OK;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Runs fn(t) for t = 0 .. nth - 1 on up to nth threads; a thread that cannot
// be started runs its share on the calling thread.  fn must not throw.
template <class F>
void run_threads(unsigned nth, F &&fn) {
    std::vector<std::thread> pool;
    pool.reserve(nth);
    unsigned started = 1;
    try {
        for (; started < nth; ++started) pool.emplace_back(fn, started);
    } catch (...) {
    }
    fn(0u);
    for (unsigned t = started; t < nth; ++t) fn(t);
    for (auto &th : pool) th.join();
}

// A device buffer freed when it goes out of scope (error and exception paths).
struct DevBuf {
    double *p = nullptr;
    ~DevBuf() { dfree(p); }
};

// Device state whose size depends on K: basis, centres, per-centre
// statistics and the partial sums of their reductions.
void free_kstate(dg_ctx *c) {
    dfree(c->FI); dfree(c->Y64); dfree(c->yy); dfree(c->G); dfree(c->Gm); dfree(c->RT);
    dfree(c->partRT); dfree(c->partG); dfree(c->partAB);
}

void free_problem(dg_ctx *c) {
    free_kstate(c);
    dfree(c->T64); dfree(c->tt); dfree(c->W); dfree(c->A); dfree(c->A0); dfree(c->Bm);
    dfree(c->pm); dfree(c->pl); dfree(c->logg); dfree(c->logh);
    dg_mixed_free(c);
    dfree(c->perm);
    dfree(c->p1start);
    c->loaded = false;
    c->have_estep = false;
    c->have_w = false;
}

// Allocates the K-dependent state for c->K (c->N, c->D, c->M1 set); Y64 and
// Gm start at zero.
int alloc_kstate(dg_ctx *c) {
    const long N = c->N, K = c->K;
    const int D = c->D, M1 = c->M1;
    const size_t ld = (size_t)c->ldt;
    // N chunks of E-pass 2: enough (centre tile x chunk) workgroups to fill
    // 256 CUs several times, each chunk at least one 64-point tile.
    const long tiles = (N + 63) / 64;
    const long ktiles = (K + 63) / 64;
    long nc = (1024 + ktiles - 1) / ktiles;
    if (nc > tiles) nc = tiles;
    if (nc < 1) nc = 1;
    c->NC = (int)nc;
    // A/B reduction: at most 64 chunks of centres, at least 64 centres each
    long nch = (K + 63) / 64;
    if (nch > 64) nch = 64;
    c->nchAB = (int)(nch < 1 ? 1 : nch);
    int rc;
    if ((rc = dalloc(c, &c->FI, (size_t)K * M1)) || (rc = dalloc(c, &c->Y64, (size_t)K * ld)) ||
        (rc = dalloc(c, &c->yy, K)) || (rc = dalloc(c, &c->G, K)) || (rc = dalloc(c, &c->Gm, K)) ||
        (rc = dalloc(c, &c->RT, (size_t)K * D)) ||
        (rc = dalloc(c, &c->partRT, (size_t)c->NC * K * D)) ||
        (rc = dalloc(c, &c->partG, (size_t)c->NC * K)) ||
        (rc = dalloc(c, &c->partAB, std::max((size_t)c->nchAB * ((size_t)M1 * M1 + (size_t)M1 * D),
                                             (size_t)(K + 63) / 64))))
        return rc;
    DG_HIP(c, hipMemsetAsync(c->Y64, 0, (size_t)K * ld * sizeof(double), c->stream));
    DG_HIP(c, hipMemsetAsync(c->Gm, 0, (size_t)K * sizeof(double), c->stream));
    return DG_OK;
}

// Host rows (stride lds) -> device rows (stride ldd >= cols, zero padding).
// Pageable host memory moves far faster as one block than as many short
// rows (10^6 rows of 166 doubles: 9.4 s strided, 0.2 s as one block), so the
// host block -- (rows - 1) lds + cols doubles, nothing past the last row's
// data -- goes over in one copy and is repacked on the device.
int upload_rows(dg_ctx *c, double *dst, int ldd, const double *src, long lds, long rows,
                int cols) {
    const size_t n = (size_t)(rows - 1) * lds + cols;
    if (lds == cols && ldd == cols) {
        DG_HIP(c, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        return DG_OK;
    }
    DevBuf tmp;
    int rc;
    if ((rc = dalloc(c, &tmp.p, n))) return rc;
    DG_HIP(c, hipMemcpyAsync(tmp.p, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = dg_pad_rows(c, tmp.p, rows, cols, lds, ldd, dst))) return rc;
    // the block is freed on return: wait for the repack that reads it
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int sync_scalars(dg_ctx *c) {
    DG_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, kScalCount * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int select_device(dg_ctx *c) {
    DG_HIP(c, hipSetDevice(c->device));
    return DG_OK;
}

// Quality of mpicomm.c:463-500: [N((D/2) log(beta/2pi) - log K) + w * sum] / N
double quality(long N, int K, int D, double beta, double sum, double w) {
    double q = N * ((D / 2.0) * log(beta / (2.0 * M_PI)) - log((double)K));
    q += w * sum;
    return q / (double)N;
}

// ---- tuned pass-2 chunk counts ------------------------------------------
// A text table, one line per shape: "<precision> <D> <K> <nbits(N)> <NC>"
// (the last line for a shape wins), named by the DG_TUNE_FILE environment
// variable and written by dg_tune.  Shared by every context of the process.
std::mutex g_tune_mu;
std::string g_tune_path;
std::map<std::array<long, 4>, int> g_tune;

std::array<long, 4> tune_key(const dg_ctx *c) {
    long nb = 0;
    while (nb < 62 && (1L << nb) < c->N) ++nb;
    return {(long)c->cfg.precision, (long)c->D, (long)c->K, nb};
}

// caller holds g_tune_mu
void tune_load(const std::string &path) {
    if (path == g_tune_path) return;
    g_tune.clear();
    g_tune_path = path;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return;
    long p, d, k, nb, nc;
    while (std::fscanf(f, "%ld %ld %ld %ld %ld", &p, &d, &k, &nb, &nc) == 5)
        g_tune[{p, d, k, nb}] = (int)nc;
    std::fclose(f);
}

int tune_find(const dg_ctx *c) {
    const char *path = std::getenv("DG_TUNE_FILE");
    if (!path || !*path) return 0;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load(path);
    auto it = g_tune.find(tune_key(c));
    return it == g_tune.end() ? 0 : it->second;
}

// ---- dg_map_noalloc's scratch ---------------------------------------------
// A file-scope device array: every GPU the library runs on holds its own copy
// (16 MiB) until the process ends.  Its device address is looked up once per
// device (hipGetSymbolAddress with that device current) and kept; host code
// uses only that address, never the array's name.
constexpr size_t kMapScratchDoubles = (size_t)1 << 21;
__device__ __align__(256) double scratch[kMapScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// The calling thread has `device` current.  Returns this device's scratch
// address, or nullptr when the lookup failed (reported by dg_map_noalloc).
double *scratch_for_device(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) {
            g_scratch_ptr[device] = static_cast<double *>(ptr);
        } else {
            (void)hipGetLastError();
            g_scratch_ptr[device] = nullptr;
        }
        g_scratch_done[device] = true;
    }
    return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

const char *dg_version(void) { return DG_VERSION_STRING; }

const char *dg_last_error(const dg_ctx *ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int dg_create(const dg_config *cfg, dg_ctx **out) {
    if (!out) return DG_EINVAL;
    *out = nullptr;
    dg_config def{0, DG_FP64, 0.0};
    const dg_config &cf = cfg ? *cfg : def;
    if (cf.precision != DG_FP64 && cf.precision != DG_MIXED && cf.precision != DG_MIXED_R1)
        return DG_EINVAL;
    if (cf.lambda < 0.0) return DG_EINVAL;
    dg_ctx *c = new (std::nothrow) dg_ctx();
    if (!c) return DG_ENOMEM;
    c->cfg = cf;
    c->device = cf.device;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0 || cf.device < 0 || cf.device >= ndev) {
        delete c;
        (void)hipGetLastError();
        return DG_EHIP;
    }
    int rc = DG_OK;
    if ((rc = select_device(c)) != DG_OK) { delete c; return rc; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DG_EHIP;
    }
    for (auto &ev : c->ev)
        if (hipEventCreate(&ev) != hipSuccess) { dg_destroy(c); return DG_EHIP; }
    if (dalloc(c, &c->scal, kScalCount) != DG_OK ||
        hipHostMalloc((void **)&c->h_scal, kScalCount * sizeof(double), hipHostMallocDefault) !=
            hipSuccess) {
        dg_destroy(c);
        return DG_ENOMEM;
    }
    std::memset(c->h_scal, 0, kScalCount * sizeof(double));
    c->map_scratch = scratch_for_device(c->device);
    *out = c;
    return DG_OK;
}

int dg_destroy(dg_ctx *c) {
    if (!c) return DG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_problem(c);
    dfree(c->scal);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DG_OK;
}

int dg_upload(dg_ctx *c, const double *T, long N, int D, long ldT, const double *FI, int K,
              int M1, long ldFI, const double *Y, long ldY, double beta) {
    if (!c) return DG_EINVAL;
    if (!T || !FI || !Y || N < 1 || D < 1 || K < 1 || M1 < 1 || ldT < D || ldFI < M1 || ldY < D ||
        !(beta > 0.0))
        return dg_fail(c, DG_EINVAL, "dg_upload: bad shape, stride or beta");
    if (D > 192) return dg_fail(c, DG_EINVAL, "dg_upload: D > 192 is not supported");
    if (M1 > 2048) return dg_fail(c, DG_EINVAL, "dg_upload: M+1 > 2048 is not supported");
    int rc;
    if ((rc = select_device(c)) != DG_OK) return rc;
    free_problem(c);
    c->N = N; c->D = D; c->K = K; c->M1 = M1; c->beta = beta;
    c->ldt = ((D + 15) / 16) * 16;
    const size_t ld = (size_t)c->ldt;
    if ((rc = dalloc(c, &c->T64, (size_t)N * ld)) || (rc = dalloc(c, &c->tt, N)) ||
        (rc = dalloc(c, &c->W, (size_t)M1 * D)) || (rc = dalloc(c, &c->A, (size_t)M1 * M1)) ||
        (rc = dalloc(c, &c->A0, (size_t)M1 * M1)) || (rc = dalloc(c, &c->Bm, (size_t)M1 * D)) ||
        (rc = dalloc(c, &c->pm, N)) || (rc = dalloc(c, &c->pl, N)) ||
        (rc = dalloc(c, &c->logg, N)) || (rc = dalloc(c, &c->logh, N)) ||
        (rc = alloc_kstate(c))) {
        free_problem(c);
        return rc;
    }
    if ((rc = upload_rows(c, c->T64, (int)ld, T, ldT, N, D)) ||
        (rc = upload_rows(c, c->FI, M1, FI, ldFI, K, M1)) ||
        (rc = upload_rows(c, c->Y64, (int)ld, Y, ldY, K, D))) {
        free_problem(c);
        return rc;
    }
    if ((rc = dg_rownorms(c, c->T64, N, c->ldt, D, c->tt))) return rc;
    if ((rc = dg_rownorms(c, c->Y64, K, c->ldt, D, c->yy))) return rc;
    // sum_n |t_n|^2 for the closed-form beta update
    if ((rc = dg_sum_into(c, c->tt, N, kScalBetaSum)) || (rc = sync_scalars(c))) return rc;
    c->sum_tt = c->h_scal[kScalBetaSum];
    c->loaded = true;
    if (c->mixed()) {
        c->nc_override = tune_find(c);
        if ((rc = dg_mixed_prepare_t(c)) || (rc = dg_mixed_prepare_y(c)) ||
            (rc = dg_mixed_order(c)))
            return rc;
    }
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int dg_set_state(dg_ctx *c, const double *Y, long ldY, double beta) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_set_state before dg_upload");
    if (!Y || ldY < c->D || !(beta > 0.0)) return dg_fail(c, DG_EINVAL, "dg_set_state: bad args");
    int rc;
    if ((rc = select_device(c))) return rc;
    if ((rc = upload_rows(c, c->Y64, c->ldt, Y, ldY, c->K, c->D))) return rc;
    if ((rc = dg_rownorms(c, c->Y64, c->K, c->ldt, c->D, c->yy))) return rc;
    DG_HIP(c, hipMemsetAsync(c->Gm, 0, (size_t)c->K * sizeof(double), c->stream));
    c->beta = beta;
    c->have_w = false;
    c->have_estep = false;
    if (c->mixed() && (rc = dg_mixed_prepare_y(c))) return rc;
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int dg_estep(dg_ctx *c, int isDA, double temp, dg_estep_out *out) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_estep before dg_upload");
    if (isDA && !(temp > 0.0)) return dg_fail(c, DG_EINVAL, "dg_estep: temp must be > 0");
    int rc;
    if ((rc = select_device(c))) return rc;
    if (c->mixed()) rc = dg_estep_mixed(c, isDA, temp);
    else rc = dg_estep_fp64(c, isDA, temp);
    if (rc) return rc;
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[3], c->stream));
    if ((rc = sync_scalars(c))) return rc;
    if (c->timing) {
        float a = 0, b = 0, d = 0;
        (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&d, c->ev[2], c->ev[3]);
        c->t_ms[0] = a; c->t_ms[1] = b; c->t_ms[2] = d;
    }
    c->have_estep = true;
    c->last_isDA = isDA != 0;
    c->last_temp = isDA ? temp : 1.0;
    const double sg = c->h_scal[kScalSumLogG], sh = c->h_scal[kScalSumLogH];
    if (out) {
        out->llh = quality(c->N, c->K, c->D, c->beta, sg, 1.0);
        out->nfree = isDA ? quality(c->N, c->K, c->D, c->beta, sh, temp) : -out->llh;
        out->nzero = (long)c->h_scal[kScalNZero];
    }
    return DG_OK;
}

int dg_mstep(dg_ctx *c, double *beta_out) {
    if (!c) return DG_EINVAL;
    if (!c->loaded || !c->have_estep) return dg_fail(c, DG_ESTATE, "dg_mstep needs a dg_estep first");
    int rc;
    if ((rc = select_device(c))) return rc;
    // the G of this M-step is the one the critical-temperature search reads
    // (vggsub of dagtm.cxx:1715, still set when the next E-step converges)
    DG_HIP(c, hipMemcpyAsync(c->Gm, c->G, (size_t)c->K * sizeof(double),
                             hipMemcpyDeviceToDevice, c->stream));
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[4], c->stream));
    if ((rc = dg_mstep_run(c, 0.0))) return rc;
    if (c->timing) DG_HIP(c, hipEventRecord(c->ev[5], c->stream));
    if ((rc = sync_scalars(c))) return rc;
    if (c->h_scal[kScalSolveOK] != 1.0 && c->M1 > 126) {
        // a non-positive pivot: retry with a jitter of 1e-13 tr(A) / n, growing
        // x10 per retry, at most 8 retries (the policy of the one-kernel
        // factorisation that M + 1 <= 126 uses)
        if ((rc = dg_mstep_trace(c)) || (rc = sync_scalars(c))) return rc;
        double jit = 1e-13 * c->h_scal[kScalJitter] / c->M1;
        for (int attempt = 1; attempt <= 8; ++attempt, jit *= 10.0) {
            if ((rc = dg_mstep_run(c, jit)) || (rc = sync_scalars(c))) return rc;
            if (c->h_scal[kScalSolveOK] == 1.0) break;
        }
    }
    if (c->timing) {
        float a = 0;
        (void)hipEventElapsedTime(&a, c->ev[4], c->ev[5]);
        c->t_ms[3] = a;
    }
    c->have_estep = false;
    if (c->h_scal[kScalSolveOK] != 1.0)
        return dg_fail(c, DG_ESOLVE, "M-step Cholesky failed after jitter retries");
    c->have_w = true;
    const double s = c->sum_tt + c->h_scal[kScalBetaSum];
    c->beta = ((double)c->N * c->D) / s;
    if (c->mixed() && (rc = dg_mixed_prepare_y(c))) return rc;
    if (beta_out) *beta_out = c->beta;
    return DG_OK;
}

namespace {
// dg_anneal's steps for dg_anneal_ctl: the context's own E-step, M-step and
// critical temperatures, and the progress callback with the current beta.
struct AnnealSteps {
    dg_ctx *c;
    dg_iter_cb cb;
    void *user;
};
int st_estep(void *u, int isDA, double temp, double *llh, double *nfree) {
    dg_estep_out eo;
    const int rc = dg_estep(static_cast<AnnealSteps *>(u)->c, isDA, temp, &eo);
    *llh = eo.llh; *nfree = eo.nfree;
    return rc;
}
int st_mstep(void *u) { return dg_mstep(static_cast<AnnealSteps *>(u)->c, nullptr); }
int st_ctemp(void *u, double temp, double eps, double *ct) {
    return dg_next_ctemp(static_cast<AnnealSteps *>(u)->c, temp, eps, nullptr, ct, nullptr);
}
int st_iter(void *u, int nloop, int lloop, double temp, double nfree, double llh, double diff) {
    const AnnealSteps *a = static_cast<AnnealSteps *>(u);
    if (a->cb) a->cb(a->user, nloop, lloop, temp, nfree, llh, diff, a->c->beta);
    return DG_OK;
}
}  // namespace

int dg_anneal(dg_ctx *c, const dg_anneal_params *p, dg_anneal_result *res, dg_iter_cb cb,
              void *user) {
    if (!c || !p || !res) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_anneal before dg_upload");
    if (p->schedule != DG_SCHED_EXP && p->schedule != DG_SCHED_LINEAR &&
        p->schedule != DG_SCHED_AUTO)
        return dg_fail(c, DG_EINVAL, "dg_anneal: unknown schedule");
    AnnealSteps steps{c, cb, user};
    const dg_anneal_ops ops{st_estep, st_mstep, st_ctemp, st_iter, nullptr};
    const int rc = dg_anneal_ctl(p, nullptr, &ops, &steps, res);
    res->beta = c->beta;
    return rc;
}

int dg_tune(dg_ctx *c, const char *table_path, int *nc_out) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_tune before dg_upload");
    if (nc_out) *nc_out = 0;
    if (!c->mixed()) return DG_OK;             // the FP64 path has no tuned choice
    int rc;
    if ((rc = select_device(c))) return rc;
    long lo = 0, hi = 0;
    const long heur = dg_mixed_nc_limits(c, &lo, &hi);
    std::vector<long> cand;
    for (long v : {heur, heur / 2, heur / 4, heur / 8, heur / 16, lo, 2 * heur}) {
        v = std::max(lo, std::min(hi, (v + 7) / 8 * 8));
        if (std::find(cand.begin(), cand.end(), v) == cand.end()) cand.push_back(v);
    }
    // time three DA E-steps at temperature 2 (three exponentials per
    // element) per candidate, after one untimed step
    using clk = std::chrono::steady_clock;
    long best = heur;
    double best_ms = HUGE_VAL;
    for (long nc : cand) {
        c->nc_override = (int)nc;
        if ((rc = dg_mixed_prepare_t(c)) || (rc = dg_mixed_prepare_y(c))) return rc;
        if ((rc = dg_estep(c, 1, 2.0, nullptr))) return rc;
        double ms = HUGE_VAL;
        for (int r = 0; r < 3; ++r) {
            const auto t0 = clk::now();
            if ((rc = dg_estep(c, 1, 2.0, nullptr))) return rc;
            ms = std::min(ms, std::chrono::duration<double, std::milli>(clk::now() - t0).count());
        }
        if (ms < best_ms) { best_ms = ms; best = nc; }
    }
    c->nc_override = (int)best;
    if ((rc = dg_mixed_prepare_t(c)) || (rc = dg_mixed_prepare_y(c))) return rc;
    c->have_estep = false;
    if (nc_out) *nc_out = (int)best;
    if (table_path && *table_path) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        FILE *f = std::fopen(table_path, "a");
        if (!f) return dg_fail(c, DG_EINVAL, std::string("dg_tune: cannot write ") + table_path);
        const auto k = tune_key(c);
        std::fprintf(f, "%ld %ld %ld %ld %ld\n", k[0], k[1], k[2], k[3], best);
        std::fclose(f);
        if (g_tune_path == table_path) g_tune[k] = (int)best;
    }
    return DG_OK;
}

int dg_set_basis(dg_ctx *c, const double *FI, int K, int M1, long ldFI, const double *W, long ldW) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_set_basis before dg_upload");
    if (!FI || K < 1 || M1 != c->M1 || ldFI < M1 || (W && ldW < c->D))
        return dg_fail(c, DG_EINVAL, "dg_set_basis: bad shape or stride (M+1 must not change)");
    if (!W && !c->have_w)
        return dg_fail(c, DG_ESTATE, "dg_set_basis: no M-step since the last upload; pass W");
    int rc;
    if ((rc = select_device(c))) return rc;
    DG_HIP(c, hipStreamSynchronize(c->stream));
    free_kstate(c);
    c->K = K;
    c->have_estep = false;
    if ((rc = alloc_kstate(c))) { free_problem(c); return rc; }
    if ((rc = upload_rows(c, c->FI, M1, FI, ldFI, K, M1))) return rc;
    if (W && (rc = upload_rows(c, c->W, c->D, W, ldW, M1, c->D))) return rc;
    c->have_w = true;
    if ((rc = dg_centres_from_w(c))) return rc;
    // the mixed path's operand layout and chunking depend on K: rebuild it
    // from the resident FP64 data
    if (c->mixed()) c->nc_override = tune_find(c);
    if (c->mixed() &&
        ((rc = dg_mixed_prepare_t(c)) || (rc = dg_mixed_prepare_y(c))))
        return rc;
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

static int next_ctemp_impl(dg_ctx *c, double temp, double eps, const double *gg,
                           double *ctemp_max, double *ctemp_k);

int dg_next_ctemp(dg_ctx *c, double temp, double eps, const double *gg, double *ctemp_max,
                  double *ctemp_k) {
    if (!c) return DG_EINVAL;
    try {
        return next_ctemp_impl(c, temp, eps, gg, ctemp_max, ctemp_k);
    } catch (...) {
        return dg_fail(c, DG_ENOMEM, "dg_next_ctemp: host allocation failed");
    }
}

static int next_ctemp_impl(dg_ctx *c, double temp, double eps, const double *gg,
                           double *ctemp_max, double *ctemp_k) {
    if (!c->loaded || !c->have_estep)
        return dg_fail(c, DG_ESTATE, "dg_next_ctemp needs a dg_estep at the current state");
    if (!ctemp_max || !std::isfinite(temp) || !std::isfinite(eps))
        return dg_fail(c, DG_EINVAL, "dg_next_ctemp: bad args");
    int rc;
    if ((rc = select_device(c))) return rc;
    // R in FP64: the mixed path refreshes the FP64 per-point normalisers
    if (c->mixed() && (rc = dg_estep_fp64(c, c->last_isDA, c->last_temp)))
        return rc;
    const int K = c->K, D = c->D;
    const long DD = (long)D * D;
    std::vector<double> g(K);
    if (gg) {
        std::copy(gg, gg + K, g.begin());
    } else {
        DG_HIP(c, hipMemcpyAsync(g.data(), c->Gm, (size_t)K * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        DG_HIP(c, hipStreamSynchronize(c->stream));
    }
    // centres per batch: S1 and S2 of a batch take at most 256 MB
    const int KB = (int)std::max(1L, std::min((long)K, (1L << 25) / (2 * DD)));
    const int nt = (D + 63) / 64, pairs = nt * (nt + 1) / 2;
    const long tiles = (c->N + 63) / 64;
    long nc = (2048 + (long)pairs * KB - 1) / ((long)pairs * KB);
    nc = std::max(1L, std::min({nc, 64L, tiles}));
    const size_t slen = (size_t)KB * 2 * DD;
    DevBuf bPart, bS;
    if ((rc = dalloc(c, &bPart.p, slen * nc))) return rc;
    if (nc > 1 && (rc = dalloc(c, &bS.p, slen))) return rc;
    double *dPart = bPart.p, *dS = bS.p;
    std::vector<double> hS(slen), ck(K);
    const double GG_MIN = 1e-250;   // mpicomm.c:617-628
    int bad = 0;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int k0 = 0; k0 < K && !rc; k0 += KB) {
        const int kb = std::min(KB, K - k0);
        rc = dg_ctemp_moments_fp64(c, k0, kb, (int)nc, dPart, nc > 1 ? dS : dPart);
        hipError_t e = hipSuccess;
        if (!rc) e = hipMemcpyAsync(hS.data(), nc > 1 ? dS : dPart, (size_t)kb * 2 * DD * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream);
        if (!rc && e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = dg_hip_check(c, e, "dg_next_ctemp");
        if (rc) break;
        // per centre: eig(H) of mpicomm.c:599-626 with H = [[2S1 - S2, -S2], [-S2, 2S1 - S2]]
        // is eig(2 S1) together with eig(2 S1 - 2 S2) (H is block-symmetric)
        std::atomic<int> fail{0};
        auto work = [&](unsigned t) noexcept {
            std::vector<double> A, w;
            try {
                A.resize((size_t)DD);
                w.resize(D);
            } catch (...) {
                fail.fetch_or(2);
                return;
            }
            for (int i = (int)t; i < kb; i += (int)nth) {
                const int k = k0 + i;
                if (g[k] == 0.0) { ck[k] = HUGE_VAL; continue; }
                const double *S1 = hS.data() + (size_t)i * 2 * DD, *S2 = S1 + DD;
                const double sc = g[k] < GG_MIN ? 1.0 / GG_MIN : 1.0;
                const double limit = (temp - eps) * g[k] * 2.0 / c->beta;
                double lmax = -HUGE_VAL;
                for (int part = 0; part < 2; ++part) {
                    for (int a = 0; a < D; ++a)
                        for (int b = 0; b <= a; ++b) {
                            const double s1 = 2.0 * S1[(size_t)a * D + b];
                            const double v = part ? s1 - 2.0 * S2[(size_t)a * D + b] : s1;
                            A[(size_t)a * D + b] = v * sc;
                        }
                    const int er = dg_sym_eigvals(D, A.data(), D, w.data());
                    if (er) { fail.fetch_or(er == 2 ? 2 : 1); return; }
                    for (int j = 0; j < D; ++j) {
                        const double l = (w[j] / sc >= limit) ? 0.0 : w[j] / sc;   // comm.c:2149-2161
                        if (l > lmax) lmax = l;
                    }
                }
                ck[k] = c->beta / 2.0 / g[k] * lmax;
            }
        };
        run_threads(nth, work);
        bad |= fail.load();
    }
    if (rc) return rc;
    // bit 1: an eigenvalue iteration did not converge; bit 2: host allocation failed
    if (bad & 2) return dg_fail(c, DG_ENOMEM, "dg_next_ctemp: host allocation failed");
    if (bad & 1) return dg_fail(c, DG_ESOLVE, "dg_next_ctemp: eigenvalue iteration did not converge");
    double cmax = -HUGE_VAL;
    for (int k = 0; k < K; ++k) cmax = std::max(cmax, ck[k]);
    *ctemp_max = cmax;
    if (ctemp_k) std::copy(ck.begin(), ck.end(), ctemp_k);
    return DG_OK;
}

constexpr int kMaxBatchWorkers = 4;

static int batch_anneal_impl(const dg_config *cfg, int workers, const dg_batch_item *items,
                             int n, dg_anneal_result *res, int *status, double *phase_ms);

int dg_batch_anneal(const dg_config *cfg, int workers, const dg_batch_item *items, int n,
                    dg_anneal_result *res, int *status, double *phase_ms) {
    try {
        return batch_anneal_impl(cfg, workers, items, n, res, status, phase_ms);
    } catch (...) {
        return DG_ENOMEM;
    }
}

static int batch_anneal_impl(const dg_config *cfg, int workers, const dg_batch_item *items,
                             int n, dg_anneal_result *res, int *status, double *phase_ms) {
    if (!cfg || n < 0 || (n > 0 && (!items || !res || !status))) return DG_EINVAL;
    if (phase_ms) phase_ms[0] = phase_ms[1] = 0.0;
    if (n == 0) return DG_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        (void)hipGetLastError();
        return DG_EHIP;
    }
    // workers 0 / -1: the batched launch sequence (with / without hipGraph
    // capture) on this thread, for the mixed precisions; otherwise host
    // workers, at most kMaxBatchWorkers: one per hardware queue of the
    // process (HIP's default of four)
    const bool batched = workers <= 0 && cfg->precision != DG_FP64;
    const bool graphs = workers == 0;
    if (batched || workers <= 0 || workers > kMaxBatchWorkers) workers = kMaxBatchWorkers;
    workers = batched ? 1 : std::min(workers, n);
    // largest work first, so the tail of the queue holds the short items
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double)items[a].K * items[a].N * items[a].D >
               (double)items[b].K * items[b].N * items[b].D;
    });
    for (int i = 0; i < n; ++i) { status[i] = DG_EHIP; std::memset(&res[i], 0, sizeof(res[i])); }
    std::vector<dg_ctx *> ctx(n, nullptr);
    auto run_pool = [&](const std::function<void(int)> &job) {
        std::atomic<int> next{0};
        run_threads((unsigned)workers, [&](unsigned) noexcept {
            for (int j; (j = next.fetch_add(1)) < n;) {
                try {
                    job(order[j]);
                } catch (...) {
                    status[order[j]] = DG_ENOMEM;
                }
            }
        });
    };
    using clk = std::chrono::steady_clock;
    // phase 1: one context per instance, data resident on the device, on the
    // calling thread: every device allocation and free of the batch happens
    // here or after phase 2, never while workers launch
    const auto t0 = clk::now();
    for (int j = 0; j < n; ++j) {
        const int i = order[j];
        const dg_batch_item &it = items[i];
        try {
            int rc = dg_create(cfg, &ctx[i]);
            if (!rc) rc = dg_upload(ctx[i], it.T, it.N, it.D, it.ldT, it.FI, it.K, it.M1, it.ldFI,
                                    it.Y0, it.ldY0, it.beta0);
            status[i] = rc;
        } catch (...) {
            status[i] = DG_ENOMEM;
        }
    }
    // phase 2: the annealing runs, concurrently on the workers' host threads,
    // or all instances' steps as batched launches
    const auto t1 = clk::now();
    int brc = DG_OK;
    if (batched) {
        std::vector<int> ok;
        for (int i = 0; i < n; ++i)
            if (status[i] == DG_OK) ok.push_back(i);
        std::vector<dg_ctx *> cs;
        std::vector<dg_anneal_params> ps;
        std::vector<dg_anneal_result> rs(ok.size());
        std::vector<int> st(ok.size(), DG_OK);
        for (int i : ok) { cs.push_back(ctx[i]); ps.push_back(items[i].params); }
        if (!ok.empty() && (brc = select_device(ctx[ok[0]])) == DG_OK)
            brc = dg_batch_run_mixed(cs, ps.data(), rs.data(), st.data(), graphs);
        for (size_t j = 0; j < ok.size(); ++j) {
            res[ok[j]] = rs[j];
            status[ok[j]] = brc ? brc : st[j];
        }
    } else {
        run_pool([&](int i) {
            if (status[i] != DG_OK) return;
            status[i] = dg_anneal(ctx[i], &items[i].params, &res[i], nullptr, nullptr);
        });
    }
    const auto t2 = clk::now();
    for (int i = 0; i < n; ++i) {
        const dg_batch_item &it = items[i];
        if (status[i] == DG_OK && (it.W || it.Y))
            status[i] = dg_download(ctx[i], it.W, it.ldW, it.Y, it.ldY, nullptr);
        if (ctx[i]) dg_destroy(ctx[i]);
    }
    if (phase_ms) {
        phase_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        phase_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return brc == DG_EHIP ? DG_EHIP : DG_OK;
}

int dg_map(dg_ctx *c, const double *X, int L, long ldX, double *M, long ldM) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_map before dg_upload");
    if (!X || !M || L < 1 || L > 3 || ldX < L || ldM < L) return dg_fail(c, DG_EINVAL, "dg_map: bad args");
    int rc;
    if ((rc = select_device(c))) return rc;
    if (!c->have_estep) {
        dg_estep_out eo;
        if ((rc = dg_estep(c, c->last_isDA, c->last_temp, &eo))) return rc;
    }
    // output projections run in FP64; the mixed path refreshes the FP64
    // per-point normalisers from the same model state first
    if (c->mixed() && (rc = dg_estep_fp64(c, c->last_isDA, c->last_temp)))
        return rc;
    double *dX = nullptr, *dM = nullptr;
    if ((rc = dalloc(c, &dX, (size_t)c->K * L))) return rc;
    if ((rc = dalloc(c, &dM, (size_t)c->N * L))) { dfree(dX); return rc; }
    hipError_t e = hipMemcpy2DAsync(dX, (size_t)L * sizeof(double), X, (size_t)ldX * sizeof(double),
                                    (size_t)L * sizeof(double), (size_t)c->K,
                                    hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        // the mixed path keeps its FP64 state current, so the FP64 kernel serves both
        rc = dg_map_fp64(c, dX, L, dM);
        if (!rc)
            e = hipMemcpy2DAsync(M, (size_t)ldM * sizeof(double), dM, (size_t)L * sizeof(double),
                                 (size_t)L * sizeof(double), (size_t)c->N, hipMemcpyDeviceToHost,
                                 c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    dfree(dX); dfree(dM);
    if (e != hipSuccess) return dg_hip_check(c, e, "dg_map");
    return rc;
}

int dg_map_noalloc(dg_ctx *c, const double *X, int L, long ldX, double *M, long ldM) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_map_noalloc before dg_upload");
    if (!X || !M || L < 1 || L > 3 || ldX < L || ldM < L)
        return dg_fail(c, DG_EINVAL, "dg_map_noalloc: bad args");
    if (((size_t)c->K + (size_t)c->N) * (size_t)L > kMapScratchDoubles)
        return dg_fail(c, DG_EINVAL,
                       "dg_map_noalloc: (K + N) * L exceeds the scratch maximum of 2^21 "
                       "doubles (use dg_map)");
    if (!c->map_scratch)
        return dg_fail(c, DG_EHIP, "dg_map_noalloc: no scratch address for this device");
    int rc;
    if ((rc = select_device(c))) return rc;
    // the steps below launch kernels on buffers the context already owns
    if (!c->have_estep) {
        dg_estep_out eo;
        if ((rc = dg_estep(c, c->last_isDA, c->last_temp, &eo))) return rc;
    }
    if (c->mixed() && (rc = dg_estep_fp64(c, c->last_isDA, c->last_temp))) return rc;
    double *dX = c->map_scratch;                       // K x L
    double *dM = c->map_scratch + (size_t)c->K * L;    // N x L
    const size_t row = (size_t)L * sizeof(double);
    if (ldX == L)
        DG_HIP(c, hipMemcpyAsync(dX, X, (size_t)c->K * row, hipMemcpyHostToDevice, c->stream));
    else
        DG_HIP(c, hipMemcpy2DAsync(dX, row, X, (size_t)ldX * sizeof(double), row, (size_t)c->K,
                                   hipMemcpyHostToDevice, c->stream));
    if ((rc = dg_map_fp64(c, dX, L, dM))) return rc;
    if (ldM == L)
        DG_HIP(c, hipMemcpyAsync(M, dM, (size_t)c->N * row, hipMemcpyDeviceToHost, c->stream));
    else
        DG_HIP(c, hipMemcpy2DAsync(M, (size_t)ldM * sizeof(double), dM, row, row, (size_t)c->N,
                                   hipMemcpyDeviceToHost, c->stream));
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int dg_resp_dense(dg_ctx *c, double *R, long ldR) {
    if (!c) return DG_EINVAL;
    if (!c->loaded || !c->have_estep) return dg_fail(c, DG_ESTATE, "dg_resp_dense needs a dg_estep");
    if (!R || ldR < c->N) return dg_fail(c, DG_EINVAL, "dg_resp_dense: bad args");
    int rc;
    if ((rc = select_device(c))) return rc;
    if (c->mixed() && (rc = dg_estep_fp64(c, c->last_isDA, c->last_temp)))
        return rc;
    double *dR = nullptr;
    if ((rc = dalloc(c, &dR, (size_t)c->K * c->N))) return rc;
    rc = dg_resp_dense_fp64(c, dR, c->N);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpy2DAsync(R, (size_t)ldR * sizeof(double), dR, (size_t)c->N * sizeof(double),
                             (size_t)c->N * sizeof(double), (size_t)c->K, hipMemcpyDeviceToHost,
                             c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    dfree(dR);
    if (e != hipSuccess) return dg_hip_check(c, e, "dg_resp_dense");
    return rc;
}

int dg_download(dg_ctx *c, double *W, long ldW, double *Y, long ldY, double *beta) {
    if (!c) return DG_EINVAL;
    if (!c->loaded) return dg_fail(c, DG_ESTATE, "dg_download before dg_upload");
    int rc;
    if ((rc = select_device(c))) return rc;
    if (W) {
        if (ldW < c->D) return dg_fail(c, DG_EINVAL, "dg_download: ldW < D");
        DG_HIP(c, hipMemcpy2DAsync(W, (size_t)ldW * sizeof(double), c->W,
                                   (size_t)c->D * sizeof(double), (size_t)c->D * sizeof(double),
                                   (size_t)c->M1, hipMemcpyDeviceToHost, c->stream));
    }
    if (Y) {
        if (ldY < c->D) return dg_fail(c, DG_EINVAL, "dg_download: ldY < D");
        DG_HIP(c, hipMemcpy2DAsync(Y, (size_t)ldY * sizeof(double), c->Y64,
                                   (size_t)c->ldt * sizeof(double), (size_t)c->D * sizeof(double),
                                   (size_t)c->K, hipMemcpyDeviceToHost, c->stream));
    }
    DG_HIP(c, hipStreamSynchronize(c->stream));
    if (beta) *beta = c->beta;
    return DG_OK;
}

int dg_get_stats(dg_ctx *c, double *G, double *RT, long ldRT) {
    if (!c) return DG_EINVAL;
    if (!c->loaded || !c->have_estep) return dg_fail(c, DG_ESTATE, "dg_get_stats needs a dg_estep");
    int rc;
    if ((rc = select_device(c))) return rc;
    if (G)
        DG_HIP(c, hipMemcpyAsync(G, c->G, (size_t)c->K * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    if (RT) {
        if (ldRT < c->D) return dg_fail(c, DG_EINVAL, "dg_get_stats: ldRT < D");
        DG_HIP(c, hipMemcpy2DAsync(RT, (size_t)ldRT * sizeof(double), c->RT,
                                   (size_t)c->D * sizeof(double), (size_t)c->D * sizeof(double),
                                   (size_t)c->K, hipMemcpyDeviceToHost, c->stream));
    }
    DG_HIP(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

int dg_device_peaks(dg_ctx *c, double *mfma_f16_tflops, double *exp_per_s, double *fp64_tflops) {
    if (!c) return DG_EINVAL;
    int rc;
    if ((rc = select_device(c))) return rc;
    int cus = 0;
    DG_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    double m = 0, e = 0, f = 0;
    // eight 4-wave workgroups per CU: eight waves per SIMD
    if ((rc = dg_peaks_run(c, 8 * std::max(cus, 1), &m, &e, &f))) return rc;
    if (mfma_f16_tflops) *mfma_f16_tflops = m;
    if (exp_per_s) *exp_per_s = e;
    if (fp64_tflops) *fp64_tflops = f;
    return DG_OK;
}

int dg_pass2_blocks(dg_ctx *c, long *visited, long *total) {
    if (!c || !visited || !total) return DG_EINVAL;
    if (!c->mixed() || !c->have_estep)
        return dg_fail(c, DG_ESTATE, "dg_pass2_blocks: no mixed-precision E-step yet");
    int rc;
    if ((rc = select_device(c))) return rc;
    return dg_mixed_pass2_blocks(c, visited, total);
}

int dg_set_timing(dg_ctx *c, int on) {
    if (!c) return DG_EINVAL;
    c->timing = on ? 1 : 0;
    return DG_OK;
}

int dg_get_timing(dg_ctx *c, double t[4]) {
    if (!c || !t) return DG_EINVAL;
    for (int i = 0; i < 4; ++i) t[i] = c->t_ms[i];
    return DG_OK;
}

}  // extern "C"
