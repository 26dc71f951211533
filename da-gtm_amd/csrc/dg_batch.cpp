// Batched independent instances in mixed precision (dg_batch_anneal with
// workers = 0): every instance keeps its own context (data resident on the
// device) and its own annealing control, and each step of all instances runs
// as one sequence of batched launches on one stream:
//
//   E-phase: the instance table (parameters of this step, which instances
//            take an E-step) goes to the device, then make_ak, and per
//            template class -- the classes concurrently on side streams --
//            pass 1 as one persistent launch over a work queue, the log sums,
//            pass 2 likewise, the chunk reductions and the zero-column fix;
//            then one copy of every instance's scalars to the host;
//   M-phase: the table again (which instances take an M-step, their ridge),
//            the M-step kernels of all instances per launch, the FP16 split of
//            the new centres, and the scalars.
//
// Each phase is captured once as a hipGraph and replayed every step: the
// instances a step leaves out are switched off by flags in the table, so the
// graphs never change.  Between the phases the host runs each instance's
// control (dgc::AnnealMachine) on its scalars.  Every batched kernel computes
// for an instance exactly what its single-instance launch computes, so each
// instance's results equal its own dg_anneal run bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "dg_control.h"
#include "dg_internal.h"

namespace {

// quality of mpicomm.c:463-500 (as dg_estep computes it)
double batch_quality(long N, int K, int D, double beta, double sum, double w) {
    double q = N * ((D / 2.0) * log(beta / (2.0 * M_PI)) - log((double)K));
    q += w * sum;
    return q / (double)N;
}

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~Graph() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

struct Batch {
    int n = 0;
    hipStream_t st = nullptr;
    int cus = 1;
    std::vector<dg_ctx *> ctx;
    // host (pinned) and device copies of the instance table and the lists
    BInst *h_tab = nullptr, *d_tab = nullptr;
    int *h_list = nullptr, *d_list = nullptr;    // [0, n): E order (by class); [n, 2n): M order
    int *d_counters = nullptr;
    double *h_scal = nullptr, *d_scal = nullptr;
    std::vector<BClass> cls;
    int nlds = 0;
    BGrids gr;
    int2 *d_map = nullptr;     // the flat-grid maps (BGrids, BClass)
    Graph ge, gm;
    // the E-phase runs the template classes concurrently on side streams
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    hipEvent_t ev[1 + kSide] = {};

    ~Batch() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t t : side)
            if (t) (void)hipStreamDestroy(t);
        if (h_tab) (void)hipHostFree(h_tab);
        if (h_list) (void)hipHostFree(h_list);
        if (h_scal) (void)hipHostFree(h_scal);
        if (d_tab) (void)hipFree(d_tab);
        if (d_list) (void)hipFree(d_list);
        if (d_counters) (void)hipFree(d_counters);
        if (d_map) (void)hipFree(d_map);
        if (d_scal) (void)hipFree(d_scal);
        if (st) (void)hipStreamDestroy(st);
    }
};

int fill_instance(const dg_ctx *c, BInst *I) {
    std::memset(I, 0, sizeof(*I));
    I->N = c->N; I->D = c->D; I->K = c->K; I->M1 = c->M1; I->ldt = c->ldt; I->nchAB = c->nchAB;
    I->T64 = c->T64; I->tt = c->tt; I->FI = c->FI;
    I->Y64 = c->Y64; I->yy = c->yy; I->W = c->W; I->A = c->A; I->A0 = c->A0; I->Bm = c->Bm;
    I->G = c->G; I->Gm = c->Gm; I->RT = c->RT; I->pm = c->pm; I->pl = c->pl;
    I->logg = c->logg; I->logh = c->logh; I->partAB = c->partAB; I->scal = c->scal;
    I->logK = log((double)c->K);
    return dg_batch_fill_mixed(c, I);
}

// E-phase and M-phase launch sequences (captured into the graphs)
int e_phase(Batch &b) {
    DG_HIP(nullptr, hipMemcpyAsync(b.d_tab, b.h_tab, sizeof(BInst) * b.n, hipMemcpyHostToDevice,
                                   b.st));
    int rc = dg_batch_estep(b.st, b.d_tab, b.d_list, b.n, b.cls.data(), (int)b.cls.size(),
                            b.d_counters, b.gr, b.cus, b.side, Batch::kSide, b.ev);
    if (rc) return rc;
    if ((rc = dg_batch_gather_scal(b.st, b.d_tab, b.n, b.d_scal))) return rc;
    DG_HIP(nullptr, hipMemcpyAsync(b.h_scal, b.d_scal, sizeof(double) * kScalCount * b.n,
                                   hipMemcpyDeviceToHost, b.st));
    return DG_OK;
}

int m_phase(Batch &b) {
    DG_HIP(nullptr, hipMemcpyAsync(b.d_tab, b.h_tab, sizeof(BInst) * b.n, hipMemcpyHostToDevice,
                                   b.st));
    int rc = dg_batch_mstep(b.st, b.d_tab, b.d_list + b.n, b.n, b.nlds, b.gr);
    if (rc) return rc;
    if ((rc = dg_batch_prepare_y(b.st, b.d_tab, b.d_list, b.n, b.gr))) return rc;
    if ((rc = dg_batch_gather_scal(b.st, b.d_tab, b.n, b.d_scal))) return rc;
    DG_HIP(nullptr, hipMemcpyAsync(b.h_scal, b.d_scal, sizeof(double) * kScalCount * b.n,
                                   hipMemcpyDeviceToHost, b.st));
    return DG_OK;
}

int capture(Batch &b, int (*phase)(Batch &), Graph &g) {
    DG_HIP(nullptr, hipStreamBeginCapture(b.st, hipStreamCaptureModeThreadLocal));
    const int rc = phase(b);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b.st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return DG_EHIP;
    g.g = graph;
    DG_HIP(nullptr, hipGraphInstantiate(&g.x, graph, nullptr, nullptr, 0));
    return DG_OK;
}

int run_phase(Batch &b, bool graphs, int (*phase)(Batch &), Graph &g) {
    if (graphs) {
        DG_HIP(nullptr, hipGraphLaunch(g.x, b.st));
    } else {
        const int rc = phase(b);
        if (rc) return rc;
    }
    DG_HIP(nullptr, hipStreamSynchronize(b.st));
    return DG_OK;
}

}  // namespace

// Mixed-precision batch of n already uploaded contexts (each of precision
// DG_MIXED or DG_MIXED_R1), annealed with params[i].  status[i] receives each
// instance's outcome; res[i] its result.  graphs: capture the phases.
int dg_batch_run_mixed(std::vector<dg_ctx *> &ctx, const dg_anneal_params *params,
                       dg_anneal_result *res, int *status, bool graphs) {
    const int n = (int)ctx.size();
    if (n == 0) return DG_OK;
    Batch b;
    b.n = n;
    b.ctx = ctx;
    int dev = 0;
    DG_HIP(nullptr, hipGetDevice(&dev));
    DG_HIP(nullptr, hipDeviceGetAttribute(&b.cus, hipDeviceAttributeMultiprocessorCount, dev));
    DG_HIP(nullptr, hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    for (hipStream_t &t : b.side) DG_HIP(nullptr, hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    for (hipEvent_t &e : b.ev) DG_HIP(nullptr, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    DG_HIP(nullptr, hipHostMalloc((void **)&b.h_tab, sizeof(BInst) * n, hipHostMallocDefault));
    DG_HIP(nullptr, hipHostMalloc((void **)&b.h_list, sizeof(int) * 2 * n, hipHostMallocDefault));
    DG_HIP(nullptr, hipHostMalloc((void **)&b.h_scal, sizeof(double) * kScalCount * n,
                                  hipHostMallocDefault));
    DG_HIP(nullptr, hipMalloc((void **)&b.d_tab, sizeof(BInst) * n));
    DG_HIP(nullptr, hipMalloc((void **)&b.d_list, sizeof(int) * 2 * n));
    DG_HIP(nullptr, hipMalloc((void **)&b.d_scal, sizeof(double) * kScalCount * n));

    // the table, the template classes of the E-passes and the work items
    std::map<std::tuple<int, int, int, int>, std::vector<int>> byclass;   // KS, TT, GM, RLO
    for (int i = 0; i < n; ++i) {
        BInst &I = b.h_tab[i];
        if (fill_instance(ctx[i], &I)) return DG_ESTATE;
        const int DR = 32 * ((16 * I.KS + 31) / 32);
        const int rlo = ctx[i]->cfg.precision == DG_MIXED ? 1 : 0;
        byclass[std::make_tuple(I.KS, I.TT, DR > I.D ? 1 : 0, rlo)].push_back(i);
    }
    int pos = 0;
    for (auto &kv : byclass) {
        BClass k{};
        k.KS = std::get<0>(kv.first);
        k.TT = std::get<1>(kv.first);
        k.rlo = std::get<3>(kv.first);
        k.two = 1;     // at temperature 1 the second sum equals the first, bit for bit
        k.first = pos;
        k.n = (int)kv.second.size();
        k.index = (int)b.cls.size();
        k.D = b.h_tab[kv.second[0]].D;
        for (int i : kv.second) {
            BInst &I = b.h_tab[i];
            int e1 = 0, e2 = 0;
            dg_batch_items(I, &e1, &e2);
            I.item_e1 = k.items_e1;
            I.item_e2 = k.items_e2;
            k.items_e1 += e1;
            k.items_e2 += e2;
            b.h_list[pos++] = i;
        }
        b.cls.push_back(k);
    }
    // M order: the LDS Cholesky instances (M + 1 <= 126) first
    int q = n;
    for (int i = 0; i < n; ++i)
        if (b.h_tab[i].M1 <= 126) b.h_list[q++] = i;
    b.nlds = q - n;
    for (int i = 0; i < n; ++i)
        if (b.h_tab[i].M1 > 126) b.h_list[q++] = i;
    // flat-grid maps: E-list kinds, per-class kinds, M-list kinds
    std::vector<int2> hmap;
    auto add_map = [&](const int *l, int cnt, int kind) {
        const int at = (int)hmap.size();
        for (int j = 0; j < cnt; ++j) {
            const int w = bk_count(b.h_tab[l[j]], kind);
            for (int g = 0; g < w; ++g) hmap.push_back(make_int2(l[j], g));
        }
        return std::make_pair(at, (int)hmap.size() - at);
    };
    int at[BK_COUNT] = {};
    std::vector<std::pair<int, int>> at_cls;
    auto whole = [&](const int *l, int cnt, int kind) {
        const auto r = add_map(l, cnt, kind);
        at[kind] = r.first;
        b.gr.wg[kind] = r.second;
    };
    whole(b.h_list, n, BK_MAKE_AK);
    whole(b.h_list, n, BK_ABSMAX_PART);
    whole(b.h_list, n, BK_MAKE_Y16);
    for (BClass &k : b.cls) {
        const auto r1 = add_map(b.h_list + k.first, k.n, BK_REDUCE32);
        const auto r2 = add_map(b.h_list + k.first, k.n, BK_ZERO_FIX);
        const auto r3 = add_map(b.h_list + k.first, k.n, BK_BLOCK_LISTS);
        k.wg_reduce32 = r1.second;
        k.wg_zero_fix = r2.second;
        k.wg_lists = r3.second;
        at_cls.push_back(r1);
        at_cls.push_back(r2);
        at_cls.push_back(r3);
    }
    for (int kind : {BK_COPY_G, BK_PART_AB, BK_SUM_AB, BK_PHIW, BK_ROWNORM, BK_BETA_PART})
        whole(b.h_list + n, n, kind);
    whole(b.h_list + n, b.nlds, BK_TRSM_LDS);
    DG_HIP(nullptr, hipMalloc((void **)&b.d_map, sizeof(int2) * std::max<size_t>(1, hmap.size())));
    DG_HIP(nullptr, hipMemcpy(b.d_map, hmap.data(), sizeof(int2) * hmap.size(),
                              hipMemcpyHostToDevice));
    for (int kind = 0; kind < BK_COUNT; ++kind) b.gr.map[kind] = b.d_map + at[kind];
    for (size_t i = 0; i < b.cls.size(); ++i) {
        b.cls[i].map_reduce32 = b.d_map + at_cls[3 * i].first;
        b.cls[i].map_zero_fix = b.d_map + at_cls[3 * i + 1].first;
        b.cls[i].map_lists = b.d_map + at_cls[3 * i + 2].first;
    }
    DG_HIP(nullptr, hipMalloc((void **)&b.d_counters, sizeof(int) * 4 * b.cls.size()));
    DG_HIP(nullptr, hipMemcpy(b.d_list, b.h_list, sizeof(int) * 2 * n, hipMemcpyHostToDevice));
    for (dg_ctx *c : ctx) DG_HIP(c, hipStreamSynchronize(c->stream));

    std::vector<dgc::AnnealMachine> mach;
    mach.reserve(n);
    for (int i = 0; i < n; ++i) {
        status[i] = DG_OK;
        mach.emplace_back(params[i], nullptr);
    }
    // one pass of each phase with every instance switched off (every kernel
    // returns at once): loads the code objects and sizes the persistent grids
    // before any capture
    int rc;
    if ((rc = run_phase(b, false, e_phase, b.ge)) || (rc = run_phase(b, false, m_phase, b.gm)))
        return rc;
    if (graphs && ((rc = capture(b, e_phase, b.ge)) || (rc = capture(b, m_phase, b.gm))))
        return rc;

    for (;;) {
        // critical-temperature searches (AUTO schedule) run per instance
        for (int i = 0; i < n; ++i)
            while (status[i] == DG_OK && mach[i].next() == dgc::AnnealMachine::CTEMP) {
                double ct = 0.0;
                if ((status[i] = dg_next_ctemp(ctx[i], mach[i].temp(), params[i].eps, nullptr,
                                               &ct, nullptr)))
                    break;
                mach[i].ctemp_done(ct);
            }
        // ---- E-phase
        int ne = 0;
        for (int i = 0; i < n; ++i) {
            BInst &I = b.h_tab[i];
            I.do_e = (status[i] == DG_OK && mach[i].next() == dgc::AnnealMachine::ESTEP) ? 1 : 0;
            I.do_m = 0;
            if (!I.do_e) continue;
            ++ne;
            const double temp = mach[i].temp();
            I.isDA = params[i].isDA ? 1 : 0;
            I.it = I.isDA ? (float)(1.0 / temp) : 1.0f;
            I.beta_f = (float)ctx[i]->beta;
        }
        if (ne == 0) break;
        if ((rc = run_phase(b, graphs, e_phase, b.ge))) return rc;
        int nm = 0;
        for (int i = 0; i < n; ++i) {
            if (!b.h_tab[i].do_e) continue;
            dg_ctx *c = ctx[i];
            const double *s = b.h_scal + (size_t)kScalCount * i;
            const double temp = mach[i].temp();
            const int isDA = params[i].isDA ? 1 : 0;
            const double llh = batch_quality(c->N, c->K, c->D, c->beta, s[kScalSumLogG], 1.0);
            const double nfree =
                isDA ? batch_quality(c->N, c->K, c->D, c->beta, s[kScalSumLogH], temp) : -llh;
            c->have_estep = true;
            c->last_isDA = isDA != 0;
            c->last_temp = isDA ? temp : 1.0;
            mach[i].estep_done(llh, nfree);
            mach[i].decide();
            if (mach[i].next() == dgc::AnnealMachine::MSTEP) {
                BInst &I = b.h_tab[i];
                I.do_m = 1;
                I.ridge = (c->cfg.lambda > 0.0) ? c->cfg.lambda / c->beta : 0.0;
                ++nm;
            }
        }
        for (int i = 0; i < n; ++i) b.h_tab[i].do_e = 0;
        // ---- M-phase
        if (nm == 0) continue;
        if ((rc = run_phase(b, graphs, m_phase, b.gm))) return rc;
        for (int i = 0; i < n; ++i) {
            BInst &I = b.h_tab[i];
            if (!I.do_m) continue;
            I.do_m = 0;
            dg_ctx *c = ctx[i];
            const double *s = b.h_scal + (size_t)kScalCount * i;
            c->have_estep = false;
            if (s[kScalSolveOK] != 1.0) {
                status[i] = dg_fail(c, DG_ESOLVE, "M-step Cholesky failed after jitter retries");
                continue;
            }
            c->have_w = true;
            c->beta = ((double)c->N * c->D) / (c->sum_tt + s[kScalBetaSum]);
            mach[i].mstep_done();
        }
    }
    for (int i = 0; i < n; ++i) {
        mach[i].result(&res[i]);
        res[i].beta = ctx[i]->beta;
    }
    return DG_OK;
}
