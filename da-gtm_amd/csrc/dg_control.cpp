// The annealing control of dagtm.cxx:1508-2206 over caller-supplied steps
// (dg_anneal_ctl in dagtm_gpu.h).  Host code only: no device calls here, so
// the control can be exercised with scripted E-steps on any machine.
#include <cfloat>

#include "../../include/dagtm_gpu.h"

namespace {

// The reference's 50-entry moving-average queue (dagtm.cxx:45, 1244, 1521;
// comm.c:1704-1746): refilled with DBL_MAX per temperature, so its mean is
// +inf (or NaN after a difference) until at most one refill value is left,
// and the sum runs in index order, as dagtm_vector_sum does (comm.c:208-218).
struct QualQueue {
    static constexpr int kSize = 50;
    double v[kSize];
    int next = 0;
    void reset() {
        for (double &x : v) x = DBL_MAX;
        next = 0;
    }
    void push(double x) {
        v[next] = x;
        next = (next + 1) % kSize;
    }
    double mean() const {
        double s = 0.0;
        for (double x : v) s += x;
        return s / kSize;
    }
};

inline double absv(double a) { return (a > 0) ? a : -a; }

}  // namespace

extern "C" int dg_anneal_ctl(const dg_anneal_params *p, dg_hda_state *hda,
                             const dg_anneal_ops *ops, void *user, dg_anneal_result *res) {
    if (!p || !ops || !res || !ops->estep || !ops->mstep) return DG_EINVAL;
    if (p->schedule != DG_SCHED_EXP && p->schedule != DG_SCHED_LINEAR &&
        p->schedule != DG_SCHED_AUTO)
        return DG_EINVAL;
    const bool use_hda = hda && hda->enable && p->isDA && p->schedule == DG_SCHED_EXP;
    if ((p->isDA && p->schedule == DG_SCHED_AUTO && !ops->next_ctemp) ||
        (use_hda && (!ops->next_ctemp || !ops->expand)))
        return DG_EINVAL;
    if (hda) hda->nexpanded = 0;

    QualQueue Q;
    double temp = p->isDA ? p->start_temp : 1.0;
    int isLast = 0, nloop = p->nloop0, ntemps = 0, exit_maxloop = 0, rc = DG_OK;
    double qual = 0.0, qual_old = 0.0, llh = 0.0, nfree = 0.0;
    bool ready = false;                      // bReadyToExpand, dagtm.cxx:1504
    auto finish = [&](int status) {
        res->nloop = nloop; res->ntemps = ntemps; res->exit_maxloop = exit_maxloop;
        res->temp = temp; res->llh = llh; res->nfree = nfree; res->qual = qual;
        res->beta = 0.0;                     // the step provider owns beta
        return status;
    };
    do {
        if (temp < 1.0) { temp = 1.0; isLast = 1; }
        Q.reset();
        ntemps++;
        const double qual_before_temp = llh;
        int lloop = 0;
        while (true) {
            if ((rc = ops->estep(user, p->isDA, temp, &llh, &nfree))) return finish(rc);
            qual_old = qual;
            lloop++; nloop++;
            qual = p->isDA ? nfree : llh;
            const double diff = qual - qual_old;
            if (ops->on_iter && (rc = ops->on_iter(user, nloop, lloop, temp, nfree, llh, diff)))
                return finish(rc);
            if (absv(diff) < p->eps) break;
            if (nloop > QualQueue::kSize) {
                const double before = Q.mean();
                Q.push(qual);
                if (absv(Q.mean() - before) < p->eps) break;
            }
            if ((rc = ops->mstep(user))) return finish(rc);
            if (p->maxloop && nloop >= p->maxloop) { isLast = 1; exit_maxloop = 1; break; }
        }
        if (!p->isDA || isLast) break;
        if (p->schedule == DG_SCHED_LINEAR) {
            temp -= p->delta;
        } else if (p->schedule == DG_SCHED_AUTO) {
            double ct = 0.0;
            if ((rc = ops->next_ctemp(user, temp, p->eps, &ct))) return finish(rc);
            temp = (ct >= temp) ? temp * p->alpha : ct;
        } else {
            if (use_hda && hda->K < hda->Kmax) {
                const double gain = llh - qual_before_temp;
                bool fire = false;
                if (ready) {
                    if (gain < hda->eps2) ready = false;
                    else fire = true;
                }
                if (!ready) {
                    double ct = 0.0;
                    if ((rc = ops->next_ctemp(user, temp, p->eps, &ct))) return finish(rc);
                    if (ct >= temp * p->alpha) ready = true;
                }
                if (fire) {
                    ready = false;
                    int K = hda->K;
                    if ((rc = ops->expand(user, nloop, temp, &K))) return finish(rc);
                    hda->K = K;
                    hda->nexpanded++;
                }
            }
            temp *= p->alpha;
        }
    } while (!isLast);
    return finish(DG_OK);
}
