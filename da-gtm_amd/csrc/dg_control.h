// The annealing control of dagtm.cxx:1508-2206 as a resumable state machine:
// the caller asks next() which step to run, runs it, and reports its outcome.
// dg_anneal_ctl drives one machine with blocking callbacks; the batched
// runtime drives one machine per instance and runs the requested steps of all
// instances together.  Host code only.
#ifndef DG_CONTROL_H
#define DG_CONTROL_H

#include "../../include/dagtm_gpu.h"

namespace dgc {

class AnnealMachine {
public:
    enum Act { ESTEP, MSTEP, CTEMP, EXPAND, DONE };

    // p and hda (may be NULL) must stay valid while the machine runs
    AnnealMachine(const dg_anneal_params &p, dg_hda_state *hda);

    // The step to run now: ESTEP at temp(), MSTEP, CTEMP (the largest
    // critical temperature below temp()), EXPAND (HDA grid doubling) or DONE.
    Act next() const { return act_; }

    // outcomes, in the order next() asked for them.  After an E-step,
    // estep_done records the qualities (nloop, lloop and diff then describe
    // this iteration: the progress row) and decide() runs the convergence
    // tests and moves on.
    void estep_done(double llh, double nfree);
    void decide();
    void mstep_done();
    void ctemp_done(double ctemp_max);
    void expand_done(int K);

    double temp() const { return temp_; }
    int nloop() const { return nloop_; }
    int lloop() const { return lloop_; }
    double diff() const { return diff_; }
    double llh() const { return llh_; }
    double nfree() const { return nfree_; }
    void result(dg_anneal_result *r) const;

private:
    static constexpr int kQ = 50;          // INIT_QUEUE_SIZE, dagtm.cxx:45
    void begin_temperature();
    void end_temperature();
    void cool_exp();
    void queue_reset();
    double queue_mean() const;

    dg_anneal_params p_;
    dg_hda_state *hda_;
    bool use_hda_;
    Act act_ = DONE;
    int ctemp_for_ = 0;                    // 0: AUTO schedule, 1: HDA arming
    double temp_ = 1.0, qual_ = 0.0, qual_old_ = 0.0, llh_ = 0.0, nfree_ = 0.0, diff_ = 0.0;
    double qual_before_temp_ = 0.0;
    int isLast_ = 0, nloop_ = 0, lloop_ = 0, ntemps_ = 0, exit_maxloop_ = 0;
    bool ready_ = false, fire_ = false;
    double q_[kQ];
    int qnext_ = 0;
};

}  // namespace dgc

#endif
