"""The annealing control (dagtm.cxx:1508-2206) driven by scripted qualities.

Both restatements run here on the CPU: the product's dg_anneal_ctl (in
libdagtm_gpu.so; it makes no device call) and the oracle's orc_anneal_ctl.
Each case feeds a closed-form quality sequence and checks the step log
against an expected log worked out by hand from the reference's rules:

  - qual_old starts at 0 (Q7); |diff| < eps breaks before the M-step
    (dagtm.cxx:1689-1693);
  - once the global nloop exceeds 50, qual is pushed into a 50-entry queue
    refilled with DBL_MAX per temperature, and |mean after - mean before| <
    eps breaks, also before the M-step (1696-1709, comm.c:1704-1746).  While
    two or more DBL_MAX remain the mean is +inf and the difference NaN or
    infinite, so the window can first fire on the 51st push of a temperature;
  - the M-step runs, then maxloop ends the run (1790-1795);
  - temperature: clamp below 1 to 1 and make that pass the last (1510-1514);
    EXP multiplies by alpha, LINEAR subtracts delta, AUTO moves to the
    critical temperature or by alpha (1918-1934, 2166-2171);
  - HDA arm / cancel / fire (1938-1977).

A wrong placement of the break, a queue that is not refilled per
temperature, or lloop in place of nloop in the window test each give a
different break index in these cases (noted per test).
"""
import math

import pytest

import dagtm_gpu as dg
import oracle as O

IMPLS = ["product", "oracle"]


def run(impl, steps, **kw):
    f = dg.anneal_ctl if impl == "product" else O.anneal_ctl
    return f(steps, **kw)


class Script:
    """Scripted steps: quality (or (llh, nfree)) per global E-step number;
    logs every step as ('E', nloop, temp), ('M',), ('C', temp), ('X', nloop,
    temp, K)."""

    def __init__(self, qual, ctemp=None, nfree=None):
        self.qual, self.ctemp, self.nfree = qual, ctemp, nfree
        self.n = 0
        self.log = []

    def estep(self, isDA, temp):
        self.n += 1
        q = self.qual(self.n, temp)
        f = q if self.nfree is None else self.nfree(self.n, temp)
        self.log.append(("E", self.n, temp))
        return q, f

    def mstep(self):
        self.log.append(("M",))

    def next_ctemp(self, temp, eps):
        self.log.append(("C", temp))
        return self.ctemp(temp)

    def expand(self, nloop, temp, K):
        L = 2
        K2 = (2 * int(round(K ** (1.0 / L)))) ** L
        self.log.append(("X", nloop, temp, K2))
        return K2

    def count(self, kind):
        return sum(1 for e in self.log if e[0] == kind)


def alternating(A=1.0):
    return lambda n, t: A if n % 2 else -A


@pytest.mark.parametrize("impl", IMPLS)
def test_window_fires_on_the_51st_push(impl):
    """EM mode, qual = +1, -1, +1, ... (|diff| = 2 > eps after the first).
    Pushes start at nloop 51; after the 50th push (nloop 100) the queue holds
    q51..q100 with mean 0, compared with (DBL_MAX + 1)/50 before it: no
    break.  At nloop 101 q101 = +1 replaces q51 = +1: both means are exactly
    0, the window breaks before the 101st M-step.  (M-step before the tests:
    101 M-steps.)"""
    s = Script(alternating())
    rc, r, _, _ = run(impl, s, isDA=False, eps=1e-3)
    assert rc == 0
    assert r.nloop == 101 and r.ntemps == 1 and r.exit_maxloop == 0
    assert s.count("E") == 101 and s.count("M") == 100
    assert s.log[-1] == ("E", 101, 1.0)        # the run ends on an E-step
    assert r.qual == 1.0 and r.llh == 1.0


@pytest.mark.parametrize("impl", IMPLS)
def test_queue_refilled_per_temperature_and_global_nloop(impl):
    """DA, EXP from 1.5 with alpha 0.5: temperature 1.5 ends on the window at
    nloop 101 (as above); 0.75 is clamped to 1 and is the last.  At the second
    temperature nloop > 50 already, so pushes start at its first iteration;
    with the queue refilled, the window fires on the 51st push, lloop 51 =
    nloop 152.  (Queue not refilled: nloop 102; lloop in place of nloop:
    nloop 202.)"""
    s = Script(alternating())
    rc, r, _, _ = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=0.5, eps=1e-3,
                      start_temp=1.5)
    assert rc == 0
    assert r.nloop == 152 and r.ntemps == 2 and r.temp == 1.0
    E = [e for e in s.log if e[0] == "E"]
    assert [e[2] for e in E[:101]] == [1.5] * 101
    assert [e[2] for e in E[101:]] == [1.0] * 51
    assert s.count("M") == 100 + 50


@pytest.mark.parametrize("impl", IMPLS)
def test_first_diff_is_against_zero(impl):
    """Q7: qual_old is 0 before the first E-step, so a first quality below eps
    in magnitude ends the temperature at once.  With qual = 1e-9 everywhere
    every temperature takes one E-step and no M-step; EXP from 8 with alpha
    0.5 visits 8, 4, 2, 1 (exactly 1 is not below 1: not the last) and then
    0.5 -> 1 (the last): five temperatures, 1 twice."""
    s = Script(lambda n, t: 1e-9)
    rc, r, _, _ = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=0.5, eps=1e-6,
                      start_temp=8.0)
    assert rc == 0
    assert [e[2] for e in s.log] == [8.0, 4.0, 2.0, 1.0, 1.0]
    assert s.count("M") == 0 and r.nloop == 5 and r.ntemps == 5


@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("t0,alpha", [(3.7, 0.9), (100.0, 0.99), (1.0, 0.5), (0.3, 0.9)])
def test_exp_schedule_count(impl, t0, alpha):
    """Closed form of the EXP schedule with one E-step per temperature: the
    temperatures t0 alpha^k for k = 0 .. k* - 1 with t0 alpha^k >= 1, then
    one pass at 1 (k* = the first k with t0 alpha^k < 1, computed by the same
    repeated multiplication)."""
    expect = []
    t = t0
    while t >= 1.0:
        expect.append(t)
        t *= alpha
    expect.append(1.0)
    s = Script(lambda n, t_: 0.0)
    rc, r, _, _ = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=alpha, eps=1e-6,
                      start_temp=t0)
    assert rc == 0
    assert [e[2] for e in s.log] == expect
    assert r.ntemps == len(expect) == r.nloop
    n_closed = (math.floor(math.log(t0) / -math.log(alpha)) + 2) if t0 >= 1 else 1
    assert abs(len(expect) - n_closed) <= 1        # rounding at an exact power only


@pytest.mark.parametrize("impl", IMPLS)
def test_linear_schedule(impl):
    """LINEAR from 3 by 0.5: 3, 2.5, 2, 1.5, 1 (not below 1), then 0.5 -> 1."""
    s = Script(lambda n, t: 0.0)
    rc, r, _, _ = run(impl, s, isDA=True, schedule=dg.SCHED_LINEAR, delta=0.5, eps=1e-6,
                      start_temp=3.0)
    assert rc == 0
    assert [e[2] for e in s.log] == [3.0, 2.5, 2.0, 1.5, 1.0, 1.0]


@pytest.mark.parametrize("impl", IMPLS)
def test_auto_schedule(impl):
    """AUTO from 4 (alpha 0.9): a critical temperature at or above the current
    one is replaced by one EXP step (4 -> 3.6), a lower one is taken
    (3.6 -> 2), and one below 1 is clamped (2 -> 0.5 -> 1, the last)."""
    cts = iter([5.0, 2.0, 0.5])
    s = Script(lambda n, t: 0.0, ctemp=lambda t: next(cts))
    rc, r, _, _ = run(impl, s, isDA=True, schedule=dg.SCHED_AUTO, alpha=0.9, eps=1e-6,
                      start_temp=4.0)
    assert rc == 0
    assert [e for e in s.log] == [("E", 1, 4.0), ("C", 4.0), ("E", 2, 3.6), ("C", 3.6),
                                  ("E", 3, 2.0), ("C", 2.0), ("E", 4, 1.0)]


@pytest.mark.parametrize("impl", IMPLS)
def test_maxloop_counts(impl):
    """eps = 0 never breaks (|x| < 0 is false, NaN included): exactly maxloop
    E-steps and maxloop M-steps, the M-step before the stop, one temperature,
    exit_maxloop set."""
    s = Script(lambda n, t: float(n))
    rc, r, _, _ = run(impl, s, isDA=True, eps=0.0, maxloop=37, start_temp=5.0)
    assert rc == 0
    assert s.count("E") == 37 and s.count("M") == 37 and s.log[-1] == ("M",)
    assert r.nloop == 37 and r.ntemps == 1 and r.exit_maxloop == 1


@pytest.mark.parametrize("impl", IMPLS)
def test_restart_counts_on(impl):
    """A restart at iteration 60 (maxloop absolute, 70): nloop is already past
    50, so the window pushes from the first iteration; ten iterations run."""
    s = Script(lambda n, t: float(n))
    rc, r, _, _ = run(impl, s, isDA=True, eps=0.5, maxloop=70, start_temp=2.0, nloop0=60)
    assert rc == 0
    assert s.count("E") == 10 and s.count("M") == 10 and r.nloop == 70


@pytest.mark.parametrize("impl", IMPLS)
def test_hda_arm_fire_cancel(impl):
    """HDA on EXP from 16 by 0.5 (temperatures 16, 8, 4, 2, 1, 1), one E-step
    per temperature (nfree 0), llh per temperature 0, 5, 5.0005, 10, 10.0001,
    20; eps2 = 1e-3; K = 100 (L = 2), Kmax 10000.
      after 16: not armed -> ctemp 9 >= 16 * 0.5: armed;
      after 8:  armed, gain 5 >= eps2 -> fire (no ctemp search): K 100 -> 400;
      after 4:  not armed -> ctemp 2.5 >= 2: armed;
      after 2:  armed, gain 4.9995 -> fire: K 400 -> 1600;
      after 1 (not last): not armed -> ctemp 0.4 < 0.5: stays unarmed;
      the second 1 is the last: no HDA step.
    A cancel: the same with llh 5.0006 at temperature 2 (gain 1e-4 < eps2):
    disarm, then the search runs in the same pass (ctemp 0.1 < 1)."""
    llh_by_temp = {16.0: 0.0, 8.0: 5.0, 4.0: 5.0005, 2.0: 10.0}
    seen = []

    def llh(n, t):
        seen.append(t)
        if t == 1.0:
            return 10.0001 if seen.count(1.0) == 1 else 20.0
        return llh_by_temp[t]

    cts = iter([9.0, 2.5, 0.4])
    s = Script(llh, ctemp=lambda t: next(cts), nfree=lambda n, t: 0.0)
    rc, r, nexp, K = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=0.5, eps=1e-6,
                         start_temp=16.0, hda=dict(K=100, Kmax=10000, eps2=1e-3))
    assert rc == 0
    assert [e for e in s.log if e[0] != "E"] == [("C", 16.0), ("X", 2, 8.0, 400), ("C", 4.0),
                                                 ("X", 4, 2.0, 1600), ("C", 1.0)]
    assert nexp == 2 and K == 1600 and r.ntemps == 6

    llh_by_temp[2.0] = 5.0006
    seen.clear()
    cts = iter([9.0, 2.5, 0.1, 0.4])
    s = Script(llh, ctemp=lambda t: next(cts), nfree=lambda n, t: 0.0)
    rc, r, nexp, K = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=0.5, eps=1e-6,
                         start_temp=16.0, hda=dict(K=100, Kmax=10000, eps2=1e-3))
    assert rc == 0
    assert [e for e in s.log if e[0] != "E"] == [("C", 16.0), ("X", 2, 8.0, 400), ("C", 4.0),
                                                 ("C", 2.0), ("C", 1.0)]
    assert nexp == 1 and K == 400


@pytest.mark.parametrize("impl", IMPLS)
def test_hda_stops_at_kmax(impl):
    """K >= Kmax: no search and no expansion (dagtm.cxx:1938)."""
    s = Script(lambda n, t: 0.0, ctemp=lambda t: 1e9)
    rc, r, nexp, K = run(impl, s, isDA=True, schedule=dg.SCHED_EXP, alpha=0.5, eps=1e-6,
                         start_temp=8.0, hda=dict(K=400, Kmax=400, eps2=1e-3))
    assert rc == 0 and nexp == 0 and K == 400
    assert s.count("C") == 0 and s.count("X") == 0


@pytest.mark.parametrize("impl", IMPLS)
def test_hook_status_stops_the_loop(impl):
    """A nonzero on_iter status ends the run with that status (the driver's
    forced checkpoint returns DG_ESTOP), before the convergence tests."""
    class Stop(Script):
        def on_iter(self, nloop, lloop, temp, nfree, llh, diff):
            return dg.DG_ESTOP if nloop == 3 else 0

    s = Stop(lambda n, t: float(n))
    rc, r, _, _ = run(impl, s, isDA=True, eps=1e-9, start_temp=2.0)
    assert rc == dg.DG_ESTOP and r.nloop == 3
    assert s.count("E") == 3 and s.count("M") == 2
