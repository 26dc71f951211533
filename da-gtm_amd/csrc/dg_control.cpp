// The annealing control of dagtm.cxx:1508-2206: the resumable machine of
// dg_control.h and dg_anneal_ctl (dagtm_gpu.h), which drives one machine with
// blocking callbacks.  Host code only: no device calls here, so the control
// can be exercised with scripted E-steps on any machine.
#include "dg_control.h"

#include <cfloat>

namespace dgc {

namespace {
inline double absv(double a) { return (a > 0) ? a : -a; }
}  // namespace

AnnealMachine::AnnealMachine(const dg_anneal_params &p, dg_hda_state *hda)
    : p_(p), hda_(hda),
      use_hda_(hda && hda->enable && p.isDA && p.schedule == DG_SCHED_EXP) {
    if (hda_) hda_->nexpanded = 0;
    temp_ = p_.isDA ? p_.start_temp : 1.0;
    nloop_ = p_.nloop0;
    begin_temperature();
}

// The reference's 50-entry moving-average queue (dagtm.cxx:1244, 1521;
// comm.c:1704-1746): refilled with DBL_MAX per temperature, so its mean is
// +inf (or NaN after a difference) until at most one refill value is left,
// and the sum runs in index order, as dagtm_vector_sum does (comm.c:208-218).
void AnnealMachine::queue_reset() {
    for (double &x : q_) x = DBL_MAX;
    qnext_ = 0;
}

double AnnealMachine::queue_mean() const {
    double s = 0.0;
    for (double x : q_) s += x;
    return s / kQ;
}

// dagtm.cxx:1510-1530: clamp, refill the queue, remember the llh, run.
void AnnealMachine::begin_temperature() {
    if (temp_ < 1.0) { temp_ = 1.0; isLast_ = 1; }
    queue_reset();
    ntemps_++;
    qual_before_temp_ = llh_;
    lloop_ = 0;
    act_ = ESTEP;
}

// dagtm.cxx:1553-1603: the qualities of this iteration
void AnnealMachine::estep_done(double llh, double nfree) {
    llh_ = llh;
    nfree_ = nfree;
    qual_old_ = qual_;
    lloop_++;
    nloop_++;
    qual_ = p_.isDA ? nfree_ : llh_;
    diff_ = qual_ - qual_old_;
}

// dagtm.cxx:1689-1709: both convergence tests, before the M-step
void AnnealMachine::decide() {
    if (absv(diff_) < p_.eps) { end_temperature(); return; }
    if (nloop_ > kQ) {
        const double before = queue_mean();
        q_[qnext_] = qual_;
        qnext_ = (qnext_ + 1) % kQ;
        if (absv(queue_mean() - before) < p_.eps) { end_temperature(); return; }
    }
    act_ = MSTEP;
}

// dagtm.cxx:1790-1795
void AnnealMachine::mstep_done() {
    if (p_.maxloop && nloop_ >= p_.maxloop) {
        isLast_ = 1;
        exit_maxloop_ = 1;
        end_temperature();
        return;
    }
    act_ = ESTEP;
}

// dagtm.cxx:1909-2179: the next temperature, or the end of the run.
void AnnealMachine::end_temperature() {
    if (!p_.isDA || isLast_) { act_ = DONE; return; }
    if (p_.schedule == DG_SCHED_LINEAR) {
        temp_ -= p_.delta;
        begin_temperature();
    } else if (p_.schedule == DG_SCHED_AUTO) {
        ctemp_for_ = 0;
        act_ = CTEMP;
    } else if (use_hda_ && hda_->K < hda_->Kmax) {
        // HDA (1938-1977): an armed expansion fires on enough llh gain and
        // is cancelled otherwise; an unarmed state searches the next critical
        // temperature and arms when it is within one cooling step
        const double gain = llh_ - qual_before_temp_;
        fire_ = false;
        if (ready_) {
            if (gain < hda_->eps2) ready_ = false;
            else fire_ = true;
        }
        if (!ready_) {
            ctemp_for_ = 1;
            act_ = CTEMP;
            return;
        }
        cool_exp();
    } else {
        temp_ *= p_.alpha;
        begin_temperature();
    }
}

// the rest of the EXP branch after the arming search: fire, then cool
void AnnealMachine::cool_exp() {
    if (fire_) {
        act_ = EXPAND;
        return;
    }
    temp_ *= p_.alpha;
    begin_temperature();
}

void AnnealMachine::ctemp_done(double ct) {
    if (ctemp_for_ == 0) {                 // AUTO, dagtm.cxx:1918-1934
        temp_ = (ct >= temp_) ? temp_ * p_.alpha : ct;
        begin_temperature();
        return;
    }
    if (ct >= temp_ * p_.alpha) ready_ = true;
    cool_exp();
}

void AnnealMachine::expand_done(int K) {
    ready_ = false;
    fire_ = false;
    hda_->K = K;
    hda_->nexpanded++;
    temp_ *= p_.alpha;
    begin_temperature();
}

void AnnealMachine::result(dg_anneal_result *r) const {
    r->nloop = nloop_; r->ntemps = ntemps_; r->exit_maxloop = exit_maxloop_;
    r->temp = temp_; r->llh = llh_; r->nfree = nfree_; r->qual = qual_;
    r->beta = 0.0;                         // the step provider owns beta
}

}  // namespace dgc

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
    dgc::AnnealMachine m(*p, hda);
    int rc = DG_OK;
    while (rc == DG_OK && m.next() != dgc::AnnealMachine::DONE) {
        switch (m.next()) {
            case dgc::AnnealMachine::ESTEP: {
                double llh = 0.0, nfree = 0.0;
                const double t = m.temp();
                if ((rc = ops->estep(user, p->isDA, t, &llh, &nfree))) break;
                m.estep_done(llh, nfree);
                // the progress row and checkpoint come before the tests
                if (ops->on_iter &&
                    (rc = ops->on_iter(user, m.nloop(), m.lloop(), t, nfree, llh, m.diff())))
                    break;
                m.decide();
                break;
            }
            case dgc::AnnealMachine::MSTEP:
                if ((rc = ops->mstep(user))) break;
                m.mstep_done();
                break;
            case dgc::AnnealMachine::CTEMP: {
                double ct = 0.0;
                if ((rc = ops->next_ctemp(user, m.temp(), p->eps, &ct))) break;
                m.ctemp_done(ct);
                break;
            }
            case dgc::AnnealMachine::EXPAND: {
                int K = hda->K;
                if ((rc = ops->expand(user, m.nloop(), m.temp(), &K))) break;
                m.expand_done(K);
                break;
            }
            case dgc::AnnealMachine::DONE:
                break;
        }
    }
    m.result(res);
    return rc;
}
