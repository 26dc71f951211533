// dagtm -- command-line driver of the MI355X DA-GTM engine.
//
// Keeps the reference program's options and HDF5 layout (dagtm/dagtm.cxx:
// options 719-891, fresh start 1104-1159, restart 1161-1213, iteration body
// 1531-1795, cooling 1909-2206, output 2342-2552, checkpoint 1609-1687,
// SIGCHLD 2628-2633) and runs every iteration through the C ABI of
// libdagtm_gpu (include/dagtm_gpu.h) on one GPU.  The MPI process grid of the
// reference collapses to one process; -z is accepted and ignored.
#include <getopt.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <ctime>
#include <functional>
#include <string>
#include <vector>

#include "../../include/dagtm_gpu.h"
#include "dg_h5.hpp"
#include "dg_init.hpp"

using dgh::H5File;
using dgh::Mat;

namespace {

volatile std::sig_atomic_t g_force_ckp = 0;
void on_sigchld(int) { g_force_ckp = 1; }

struct Opts {
    int L = 3, K = 125, M = 8, maxloop = 0, ckp_every = 100, verbose = 0, device = 0;
    double alpha = 0.99, delta = 0.1, eps = 1e-6, start_temp = 1.0, lambda = 0.0, s = 2.0;
    bool isDA = true, need_ctemp = true, checkpointing = true, log_progress = false;
    bool save_R = false, init_only = false, interp = false, hda = false;
    int precision = DG_FP64;
    int Kmax = 32768;
    double eps2 = 1e-3;
    int schedule = DG_SCHED_EXP;
    unsigned long seed = 0;
    std::string in = "INPUT.h5", out, restart, ckp_prefix = "ckp.";
    std::string dT = "T", dX = "X", dFI = "FI", dW = "W", dY = "Y", dR = "R", dbeta = "beta",
                dM = "M", did = "id", dlb = "lb", dtemp = "temp", diter = "iter", dseed = "seed",
                dqual = "qual";
};

void usage(const char *prog) {
    std::fprintf(stderr,
                 "usage: %s [OPTIONS]\n"
                 "-a number    : alpha (default: 0.99)\n"
                 "-b name      : dataset name for beta (default: beta)\n"
                 "-B name      : dataset name for label (default: lb)\n"
                 "-c schedule  : cooling schedule [auto|exp|linear] (default: exp)\n"
                 "-e number    : precision (convergence threshold, default 1e-6)\n"
                 "-E number    : HDA: least llh gain over a temperature to expand (default 1e-3)\n"
                 "-f dataset   : dataset name for FI (default: FI)\n"
                 "-h           : show help\n"
                 "-H           : hierarchical DA: double the grid side at critical points\n"
                 "-i filename  : input HDF5 filename (default: INPUT.h5)\n"
                 "-I           : interpolation: MAP of the -i data under the -r model\n"
                 "-j number    : maximum number of iterations\n"
                 "-k number    : HDA: largest K to expand from (default 32768)\n"
                 "-K number    : number of latent points K (eg, 100, 1k, 1M)\n"
                 "-L number    : latent dimension\n"
                 "-m           : EM (toggle deterministic annealing off)\n"
                 "-M number    : number of basis functions M\n"
                 "-n           : disable checkpointing\n"
                 "-o filename  : output HDF5 filename\n"
                 "-p           : log progress (Temp/LLH/NFree histories)\n"
                 "-P number    : checkpointing per num. loop (default 100)\n"
                 "-r filename  : restart from a checkpoint\n"
                 "-R           : save R\n"
                 "-s number    : seed\n"
                 "-S number    : start temp\n"
                 "-t dataset   : dataset name for T (default: T)\n"
                 "-v number    : set verbose level\n"
                 "-w prefix    : checkpointing file prefix (default: ckp.)\n"
                 "-W dataset   : dataset name for W (default: W)\n"
                 "-x dataset   : dataset name for X (default: X)\n"
                 "-z PxQ       : accepted for compatibility; the engine uses one GPU\n"
                 "--precision fp64|mixed|mixed_r1  arithmetic of the E-step (default fp64)\n"
                 "--lambda x               M-step ridge lambda/beta (default 0)\n"
                 "--rbf-width s            RBF width in centre spacings (default 2)\n"
                 "--device n               HIP device (default 0)\n"
                 "--init-only              write the fresh-start state to -o and exit\n",
                 prog);
}

int parse(int argc, char **argv, Opts &o) {
    static const option longopts[] = {{"precision", required_argument, nullptr, 1},
                                      {"lambda", required_argument, nullptr, 2},
                                      {"rbf-width", required_argument, nullptr, 3},
                                      {"device", required_argument, nullptr, 4},
                                      {"init-only", no_argument, nullptr, 5},
                                      {nullptr, 0, nullptr, 0}};
    int c;
    opterr = 0;
    while ((c = getopt_long(argc, argv, "a:b:B:c:e:E:f:g:hHi:Ij:k:K:lL:mM:no:pP:r:Rs:S:t:v:w:W:x:z:",
                            longopts, nullptr)) != -1) {
        switch (c) {
            case 'a': o.alpha = std::atof(optarg); break;
            case 'b': o.dbeta = optarg; break;
            case 'B': o.dlb = optarg; break;
            case 'c':
                if (!strcasecmp(optarg, "exp")) o.schedule = DG_SCHED_EXP;
                else if (!strcasecmp(optarg, "linear")) o.schedule = DG_SCHED_LINEAR;
                else if (!strcasecmp(optarg, "auto")) o.schedule = DG_SCHED_AUTO;
                else {
                    std::fprintf(stderr, "Unknown schedule option : %s\n", optarg);
                    return 1;
                }
                break;
            case 'e': o.eps = std::atof(optarg); break;
            case 'E': o.eps2 = std::atof(optarg); break;
            case 'f': o.dFI = optarg; break;
            case 'g':
                if (strcasecmp(optarg, "sweep")) {
                    std::fprintf(stderr, "Only the sweep grid is supported\n");
                    return 1;
                }
                break;
            case 'h': usage(argv[0]); std::exit(0);
            case 'H': o.hda = !o.hda; break;
            case 'i': o.in = optarg; break;
            case 'I': o.interp = true; break;
            case 'j': o.maxloop = std::atoi(optarg); break;
            case 'k': o.Kmax = std::atoi(optarg); break;
            case 'K': {
                float num = 0, unit = 1.0f;
                char ch = 0;
                std::sscanf(optarg, "%f%c", &num, &ch);
                if (ch == 'k' || ch == 'K') unit = 1000;
                else if (ch == 'm' || ch == 'M') unit = 1000 * 1000;
                else if (ch == 'g' || ch == 'G') unit = 1000 * 1000 * 1000;
                o.K = (int)(num * unit);
            } break;
            case 'l': break;                      // per-rank log files: one process here
            case 'L': o.L = std::atoi(optarg); break;
            case 'm': o.isDA = !o.isDA; break;
            case 'M': o.M = std::atoi(optarg); break;
            case 'n': o.checkpointing = false; break;
            case 'o': o.out = optarg; break;
            case 'p': o.log_progress = !o.log_progress; break;
            case 'P': o.ckp_every = std::atoi(optarg); break;
            case 'r': o.restart = optarg; break;
            case 'R': o.save_R = !o.save_R; break;
            case 's': o.seed = (unsigned long)std::atoi(optarg); break;
            case 'S': o.start_temp = std::atof(optarg); o.need_ctemp = false; break;
            case 't': o.dT = optarg; break;
            case 'v': o.verbose = std::atoi(optarg); break;
            case 'w': o.ckp_prefix = optarg; break;
            case 'W': o.dW = optarg; break;
            case 'x': o.dX = optarg; break;
            case 'z': break;
            case 1:
                if (!strcasecmp(optarg, "mixed")) o.precision = DG_MIXED;
                else if (!strcasecmp(optarg, "mixed_r1")) o.precision = DG_MIXED_R1;
                else if (!strcasecmp(optarg, "fp64")) o.precision = DG_FP64;
                else { std::fprintf(stderr, "Unknown precision : %s\n", optarg); return 1; }
                break;
            case 2: o.lambda = std::atof(optarg); break;
            case 3: o.s = std::atof(optarg); break;
            case 4: o.device = std::atoi(optarg); break;
            case 5: o.init_only = true; break;
            case '?':
                if (isprint(optopt)) std::fprintf(stderr, "Unknown option `-%c'.\n", optopt);
                else std::fprintf(stderr, "Unknown option.\n");
                return 1;
            default: return 1;
        }
    }
    return 0;
}

#define CHECK_DG(call)                                                              \
    do {                                                                            \
        int rc_ = (call);                                                           \
        if (rc_ != DG_OK) {                                                         \
            std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, dg_last_error(ctx)); \
            dg_destroy(ctx);                                                        \
            return 2;                                                               \
        }                                                                           \
    } while (0)

// The state dagtm.cxx:1983-2095 appends to the output file at an HDA
// expansion (with -p): M.<idx> = R^T X, X, Y, FI, W, the per-centre critical
// temperatures of the last search and beta, idx = progress entries - 1.  W0
// is the host's initial W, written while no M-step has run.
bool write_expansion_snapshot(dg_ctx *ctx, const Opts &o, long idx, long N, int D, const Mat &X,
                              const Mat &FI, const Mat *W0, const std::vector<double> &ctemp_k,
                              double beta) {
    const int K = (int)FI.rows;
    Mat Mo(N, X.cols), Yc(K, D), Wc(FI.cols, D);
    if (dg_download(ctx, W0 ? nullptr : Wc.data(), D, Yc.data(), D, nullptr) != DG_OK) return false;
    if (dg_map(ctx, X.data(), (int)X.cols, X.cols, Mo.data(), X.cols) != DG_OK) return false;
    const std::string sfx = "." + std::to_string(idx);
    H5File f = H5File::open_append(o.out);
    return f.ok() && f.write_matrix("M" + sfx, Mo) && f.write_matrix("X" + sfx, X) &&
           f.write_matrix("Y" + sfx, Yc) && f.write_matrix("FI" + sfx, FI) &&
           f.write_matrix("W" + sfx, W0 ? *W0 : Wc) && f.write_vector("ctemp" + sfx, ctemp_k) &&
           f.write_scalar("beta" + sfx, beta);
}

}  // namespace

int main(int argc, char **argv) {
    Opts o;
    if (parse(argc, argv, o)) return 1;
    if (!dgh::have_hdf5()) {
        std::fprintf(stderr, "this dagtm was built without HDF5 support\n");
        return 1;
    }
    std::signal(SIGCHLD, on_sigchld);

    // ---- input data (dagtm.cxx:919-930, 1076-1102)
    Mat T;
    std::vector<int> id, lb;
    {
        H5File in = H5File::open_read(o.in);
        if (!in.ok() || !in.read_matrix(o.dT, T)) {
            std::fprintf(stderr, "cannot read dataset '%s' from %s\n", o.dT.c_str(), o.in.c_str());
            return 1;
        }
        in.read_ints(o.did, id);
        in.read_ints(o.dlb, lb);
    }
    const long N = T.rows;
    const int D = (int)T.cols;

    // ---- fresh start or restart
    Mat X, FI, Y, W;
    double beta = 0.0, temp = 1.0;
    int nloop = 0;
    int K = o.K;
    if (o.restart.empty()) {
        auto t0 = std::chrono::steady_clock::now();
        auto lap = [&t0]() {
            const auto t = std::chrono::steady_clock::now();
            const double s = std::chrono::duration<double>(t - t0).count();
            t0 = t;
            return s;
        };
        std::vector<double> mu, sd;
        dgh::mean_sd(T, mu, sd);
        if (!dgh::setup_basis(o.L, o.K, o.M, o.s, X, FI)) {
            std::fprintf(stderr, "K and M must be perfect L-th powers (K=%d M=%d L=%d)\n", o.K, o.M, o.L);
            return 1;
        }
        if (o.seed == 0) o.seed = (unsigned long)std::time(nullptr);
        Mat Yrnd = dgh::random_centres(mu, sd, o.K, o.seed);
        const double t_basis = lap();
        W = dgh::least_squares(FI, Yrnd);
        if (W.rows == 0) {
            std::fprintf(stderr, "least-squares W0: eigen-decomposition did not converge\n");
            return 1;
        }
        Y = dgh::product(FI, W);
        const double t_w0 = lap();
        beta = (o.K > 1000) ? 0.1 : dgh::initial_beta(Yrnd);
        const double t_beta = lap();
        if (o.isDA) {
            if (o.need_ctemp) o.start_temp = dgh::first_critical_temp(T) / std::pow(o.alpha, 10);
            temp = o.start_temp;
        }
        const double t_temp = lap();
        if (o.verbose > 0)
            std::fprintf(stderr,
                         "init (%d host threads): statistics and basis %.3f s, W0 and Y0 %.3f s, "
                         "beta0 %.3f s, first critical temperature %.3f s\n",
                         dgh::pool().size(), t_basis, t_w0, t_beta, t_temp);
    } else {
        H5File ck = H5File::open_read(o.restart);
        int seed = 0;
        if (!ck.ok() || !ck.read_matrix(o.dX, X) || !ck.read_matrix(o.dFI, FI) ||
            !ck.read_matrix(o.dY, Y) || !ck.read_scalar(o.dbeta, beta) ||
            !ck.read_scalar(o.dtemp, temp) || !ck.read_scalar(o.diter, nloop)) {
            std::fprintf(stderr, "cannot read checkpoint %s\n", o.restart.c_str());
            return 1;
        }
        if (ck.read_scalar(o.dseed, seed)) o.seed = (unsigned long)seed;
        K = (int)FI.rows;
        o.M = (int)FI.cols - 1;
        o.L = (int)X.cols;
        o.start_temp = temp;
        if (o.maxloop) o.maxloop = nloop + o.maxloop;
        W = Mat(FI.cols, D);
    }
    if (!o.isDA) temp = 1.0;

    if (o.init_only) {
        if (o.out.empty()) { std::fprintf(stderr, "--init-only needs -o\n"); return 1; }
        H5File f = H5File::create(o.out);
        bool ok = f.ok() && f.write_matrix(o.dX, X) && f.write_matrix(o.dFI, FI) &&
                  f.write_matrix(o.dW, W) && f.write_matrix(o.dY, Y) &&
                  f.write_scalar(o.dbeta, beta) && f.write_scalar(o.dtemp, temp) &&
                  f.write_scalar(o.dseed, (int)o.seed);
        return ok ? 0 : 1;
    }

    // ---- the output file is created before any work (dagtm.cxx:1216-1220);
    // HDA snapshots, the -I projection and the final state are appended
    if (!o.out.empty()) {
        H5File f = H5File::create(o.out);
        if (!f.ok()) { std::fprintf(stderr, "cannot create %s\n", o.out.c_str()); return 1; }
    }

    // ---- the engine
    dg_ctx *ctx = nullptr;
    dg_config cfg{o.device, o.precision, o.lambda};
    int rc = dg_create(&cfg, &ctx);
    if (rc != DG_OK) {
        std::fprintf(stderr, "dg_create failed (%d): no usable HIP device\n", rc);
        return 2;
    }
    CHECK_DG(dg_upload(ctx, T.data(), N, D, D, FI.data(), K, (int)FI.cols, FI.cols, Y.data(), D, beta));

    auto write_ckp = [&](double qual) -> bool {
        char name[512];
        std::snprintf(name, sizeof name, "%s%d.h5", o.ckp_prefix.c_str(), nloop);
        Mat Yc(K, D);
        double b = 0.0;
        if (dg_download(ctx, nullptr, 0, Yc.data(), D, &b) != DG_OK) return false;
        H5File f = H5File::create(name);
        return f.ok() && f.write_matrix(o.dX, X) && f.write_matrix(o.dFI, FI) &&
               f.write_matrix(o.dY, Yc) && f.write_scalar(o.dbeta, b) &&
               f.write_scalar(o.dtemp, temp) && f.write_scalar(o.diter, nloop) &&
               f.write_scalar(o.dseed, (int)o.seed) && f.write_scalar(o.dqual, qual);
    };

    // ---- interpolation (dagtm.cxx:1247-1295): responsibilities of the new
    // data under the model at its beta and temperature, MAP appended to -o
    if (o.interp) {
        if (o.restart.empty()) std::fprintf(stderr, "No checkpointing file.. \n");
        if (o.out.empty()) { std::fprintf(stderr, "-I needs -o\n"); dg_destroy(ctx); return 1; }
        dg_estep_out eo;
        CHECK_DG(dg_estep(ctx, o.isDA, temp, &eo));
        Mat Mo(N, X.cols);
        CHECK_DG(dg_map(ctx, X.data(), (int)X.cols, X.cols, Mo.data(), X.cols));
        H5File f = H5File::open_append(o.out);
        const bool ok = f.ok() && f.write_matrix(o.dM, Mo);
        dg_destroy(ctx);
        return ok ? 0 : 1;
    }

    // ---- annealing loop (dagtm.cxx:1508-2206, AUTO / EXP / LINEAR / EM, HDA):
    // the library's control (dg_anneal_ctl) calls back into the steps below
    struct Run {
        dg_ctx *ctx;
        Opts *o;
        long N;
        int D, K;
        int nloop;
        double temp = 0.0, beta = 0.0, qual = 0.0;
        bool w_in_engine = false, fresh = true;
        Mat *X, *FI, *W;
        std::vector<double> hT, hL, hF, ctemp_k;
        std::function<bool(double)> write_ckp;
    };
    Run run{ctx, &o, N, D, K, nloop};
    run.beta = beta;
    run.fresh = o.restart.empty();
    run.X = &X; run.FI = &FI; run.W = &W;
    run.write_ckp = [&](double q) { nloop = run.nloop; temp = run.temp; return write_ckp(q); };
    dg_anneal_ops ops{};
    ops.estep = [](void *u, int isDA, double t, double *llh, double *nfree) {
        Run *r = static_cast<Run *>(u);
        dg_estep_out eo;
        const int rc = dg_estep(r->ctx, isDA, t, &eo);
        *llh = eo.llh; *nfree = eo.nfree;
        return rc;
    };
    ops.mstep = [](void *u) {
        Run *r = static_cast<Run *>(u);
        const int rc = dg_mstep(r->ctx, &r->beta);
        if (rc == DG_OK) r->w_in_engine = true;
        return rc;
    };
    ops.next_ctemp = [](void *u, double t, double eps, double *ct) {
        Run *r = static_cast<Run *>(u);
        r->ctemp_k.assign(r->K, 0.0);
        return dg_next_ctemp(r->ctx, t, eps, nullptr, ct, r->ctemp_k.data());
    };
    ops.on_iter = [](void *u, int nl, int lloop, double t, double nfree, double llh, double diff) {
        Run *r = static_cast<Run *>(u);
        const Opts &o = *r->o;
        r->nloop = nl;
        r->temp = t;
        r->qual = o.isDA ? nfree : llh;
        if (lloop == 1) {
            std::printf("\nT= %.10g\n", t);
            std::printf(">>> %15s %15s %15s %15s %15s %15s\n", "GSEQ", "SEQ", "K", "TEMP", "NFREE", "LLH");
        }
        std::printf("::: %15d %15d %15d %15.7f %15.7f %15.7f %15.7f\n", r->nloop, lloop, r->K, t,
                    o.isDA ? nfree : -llh, llh, diff);
        std::fflush(stdout);
        if (o.log_progress) {
            r->hT.push_back(t); r->hL.push_back(llh);
            if (o.isDA) r->hF.push_back(nfree);
        }
        if ((o.checkpointing && o.ckp_every > 0 && r->nloop % o.ckp_every == 0) || g_force_ckp) {
            if (!r->write_ckp(r->qual)) std::fprintf(stderr, "checkpoint write failed\n");
            if (g_force_ckp) {
                std::fprintf(stderr, "Force checkpointing done\n");
                return DG_ESTOP;
            }
        }
        return DG_OK;
    };
    ops.expand = [](void *u, int, double, int *Kio) {
        // dagtm.cxx:1983-2160: expansion snapshot, then double the grid side,
        // new grid and basis, Y = FI W with the current W
        Run *r = static_cast<Run *>(u);
        const Opts &o = *r->o;
        if (!o.out.empty() && o.log_progress &&
            !write_expansion_snapshot(r->ctx, o, (long)r->hT.size() - 1, r->N, r->D, *r->X, *r->FI,
                                      r->w_in_engine ? nullptr : r->W, r->ctemp_k, r->beta))
            std::fprintf(stderr, "HDA snapshot write failed\n");
        const int K = (int)std::lround(std::pow(std::lround(std::pow(*Kio, 1.0 / o.L)) * 2.0, o.L));
        if (!dgh::setup_basis(o.L, K, o.M, o.s, *r->X, *r->FI)) {
            std::fprintf(stderr, "HDA: no grid for K=%d M=%d L=%d\n", K, o.M, o.L);
            return (int)DG_EINVAL;
        }
        if (!r->w_in_engine && !r->fresh) {
            std::fprintf(stderr, "HDA: expansion before any M-step after a restart\n");
            return (int)DG_ESTATE;
        }
        const int rc = dg_set_basis(r->ctx, r->FI->data(), K, (int)r->FI->cols, r->FI->cols,
                                    r->w_in_engine ? nullptr : r->W->data(), r->D);
        if (rc != DG_OK) return rc;
        r->w_in_engine = true;
        r->K = K;
        *Kio = K;
        std::printf("new K : %d\n", K);
        return (int)DG_OK;
    };
    dg_anneal_params ap{o.isDA ? 1 : 0, o.schedule, o.maxloop, o.alpha, o.delta, o.eps, temp,
                        nloop};
    dg_hda_state hda{o.hda ? 1 : 0, K, o.Kmax, o.eps2, 0};
    dg_anneal_result ar{};
    rc = dg_anneal_ctl(&ap, &hda, &ops, &run, &ar);
    if (rc == DG_ESTOP) {
        dg_destroy(ctx);
        return 0;
    }
    if (rc != DG_OK) {
        std::fprintf(stderr, "annealing failed (%d): %s\n", rc, dg_last_error(ctx));
        dg_destroy(ctx);
        return 2;
    }
    nloop = ar.nloop;
    temp = ar.temp;
    beta = run.beta;
    K = run.K;
    const double llh = ar.llh, nfree = ar.nfree, qual = ar.qual;
    const int nexpanded = hda.nexpanded;
    std::vector<double> &hT = run.hT, &hL = run.hL, &hF = run.hF;

    std::printf("\nSummary: iter %d temp %.10g beta %.10g llh %.10g nfree %.10g\n", nloop, temp,
                beta, llh, nfree);
    if (o.hda) std::printf("HDA: %d expansions, K = %d\n", nexpanded, K);

    // ---- output (dagtm.cxx:2342-2552)
    if (!o.out.empty()) {
        Mat Mo(N, X.cols), Yo(K, D), Wo(FI.cols, D);
        double b = 0;
        CHECK_DG(dg_download(ctx, Wo.data(), D, Yo.data(), D, &b));
        CHECK_DG(dg_map(ctx, X.data(), (int)X.cols, X.cols, Mo.data(), X.cols));
        H5File f = H5File::open_append(o.out);
        bool ok = f.ok() && f.write_matrix(o.dT, T) && f.write_matrix(o.dM, Mo) &&
                  f.write_matrix(o.dX, X) && f.write_matrix(o.dFI, FI) && f.write_matrix(o.dY, Yo) &&
                  f.write_matrix(o.dW, Wo) && f.write_scalar(o.dbeta, b) &&
                  f.write_scalar(o.dtemp, temp) && f.write_scalar(o.diter, nloop) &&
                  f.write_scalar(o.dseed, (int)o.seed) && f.write_scalar(o.dqual, qual);
        if (ok && o.save_R) {
            Mat R(K, N);
            CHECK_DG(dg_resp_dense(ctx, R.data(), N));
            ok = f.write_matrix(o.dR, R);
        }
        if (ok && !id.empty()) ok = f.write_ints(o.did, id);
        if (ok && !lb.empty()) ok = f.write_ints(o.dlb, lb);
        if (ok && o.log_progress) {
            ok = f.write_vector("Temp", hT) && f.write_vector("LLH", hL);
            if (ok && o.isDA) ok = f.write_vector("NFree", hF);
        }
        if (!ok) {
            std::fprintf(stderr, "writing %s failed\n", o.out.c_str());
            dg_destroy(ctx);
            return 1;
        }
    }
    dg_destroy(ctx);
    return 0;
}
