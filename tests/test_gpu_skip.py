"""Pass 2 of the mixed E-step leaves out (centre group, 32-point block) pairs
whose pass-1 bound shows that every responsibility there rounds to zero, and
the mixed path orders the points by their nearest centre so that such pairs
are common.  Both are exact in the following sense, checked here:

  - skipping: with the same point order, an E-step and an M-step with
    skipping (default) and without it (DG_MIXED_SKIP=0) give bit-identical
    llh, G, R T, W, Y and beta -- a skipped pair would only have added exact
    zeros -- while the skipping run visits fewer pairs once the
    responsibilities concentrate;
  - ordering: with the points in input order (DG_MIXED_ORDER=0) the results
    from the same state differ only by the order of the FP32/FP64 sums (pass
    1 also starts each point tile's sweep at its home centre group in the
    ordered layout, and its screen leaves out other terms below 2^-48 of a
    point's sums: llh 1e-8; G, R T, W, Y and beta 1e-6), and per-point
    outputs, which use the input order, agree;
  - several annealing steps along a trajectory where the skipped fraction
    grows stay bit-identical to the no-skip run;
  - pass 1's screen of tiles far below each point's running maximum
    (DG_MIXED_SCREEN=0 turns it off) changes the results at most in their
    last bits.
The workload: binary fingerprints (50 prototypes, flip 0.05) at D = 166 and
K = 1728 (L = 3), the same with empty and two-bit points mixed in (points
whose pass-2 offset is positive), and Gaussian-mixture data at D = 32 (no
spare d row, so G comes from the FP32 sums and the stricter bound applies)."""
import os
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(kind):
    if kind in ("fp", "fp_sparse"):
        T = datagen.fingerprints(20000, 166, 7, nproto=50).astype(np.float64)
        if kind == "fp_sparse":
            # every 16th point empty or with two bits set: their largest
            # exponent without the point's own norm term is negative, so their
            # pass-2 offset pB is positive (a bound that dropped pB there would
            # fall below the exponents it bounds)
            rng = np.random.default_rng(3)
            T[::16] = 0.0
            T[8::32, rng.integers(0, 166, 2)] = 1.0
        X, FI = datagen.latent_basis(3, 1728, 64, 1.0)
    else:
        T = datagen.gmm(12000, 32, 5)
        X, FI = datagen.latent_basis(2, 1600, 25, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, FI, Y0, b0


def _run(prob, env, steps, temp):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with dg.Engine(precision=dg.MIXED, lam=1e-3) as e:
            e.upload(*prob)
            out = []
            for _ in range(steps):
                llh, nfree, nz = e.estep(True, temp)
                G, RT = e.stats()
                vis = e.pass2_blocks()
                beta = e.mstep()
                W, Y, _ = e.download()
                out.append((llh, nfree, G, RT, beta, W, Y, vis))
            return out
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("kind", ["fp", "gmm", "fp_sparse"])
def test_skipping_is_bit_identical(kind):
    _say(f"case: pass-2 skipping, {kind}")
    prob = _problem(kind)
    steps = 6
    a = _run(prob, {}, steps, 2.0)
    b = _run(prob, {"DG_MIXED_SKIP": "0"}, steps, 2.0)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x[0] == y[0] and x[1] == y[1], i
        for k in (2, 3, 5, 6):
            assert np.array_equal(x[k], y[k]), (i, k)
        assert x[4] == y[4], i
        v, t = x[7]
        assert y[7] == (t, t) and 0 < v <= t, (i, x[7], y[7])
        _say(f"  step {i}: llh {x[0]:.10g}, pass 2 visited {v} of {t} pairs ({v / t:.3f})")
    if kind == "fp":
        # the responsibilities concentrate as beta grows: pairs drop out
        # (measured: 0.93 of the pairs visited at the first step, 0.62 at the sixth)
        fr = [x[7][0] / x[7][1] for x in a]
        assert fr[-1] < 0.8 and fr[-1] < fr[0], fr


@pytest.mark.parametrize("kind", ["fp", "fp_sparse"])
def test_pass1_screen_leaves_out_only_negligible_terms(kind):
    """Pass 1 leaves out tiles whose every exponent is more than 48 / it below
    its point's running maximum (a bound from the hi-term product): each
    omitted term weighs less than 2^-48 of its point's sums, so E- and
    M-steps with the screen (default) and without it (DG_MIXED_SCREEN=0)
    agree to the last bits -- mostly bit for bit; where a compensated sum
    rounds the other way, llh within 1e-12 and G, R T, W, Y and beta within
    1e-9 of their maxima over five steps."""
    _say(f"case: pass-1 screen, {kind}")
    prob = _problem(kind)
    a = _run(prob, {}, 5, 2.0)
    b = _run(prob, {"DG_MIXED_SCREEN": "0"}, 5, 2.0)
    for i, (x, y) in enumerate(zip(a, b)):
        rel = lambda u, v: float(np.abs(u - v).max() / max(np.abs(v).max(), 1e-300))
        d = [rel(x[k], y[k]) for k in (2, 3, 5, 6)]
        dl, db = abs(x[0] - y[0]) / abs(y[0]), abs(x[4] - y[4]) / y[4]
        _say(f"  step {i}: llh {dl:.1e} G {d[0]:.1e} RT {d[1]:.1e} W {d[2]:.1e} Y {d[3]:.1e} "
             f"beta {db:.1e}")
        assert x[1] == y[1] and dl <= 1e-12 and db <= 1e-9, (i, dl, db)
        assert max(d) <= 1e-9, (i, d)


def test_point_order_changes_only_the_sum_order():
    """From the same state: llh to 1e-8 (each point's FP32 sums over centres
    run in another order, and the FP64 sum over points is grouped
    differently), G and R T to 1e-6 of their maxima (FP32 chunk sums over
    other points), and the M-step's W, Y and beta to 1e-6."""
    _say("case: point ordering vs input order")
    prob = _problem("fp")
    x = _run(prob, {}, 1, 2.0)[0]
    y = _run(prob, {"DG_MIXED_ORDER": "0"}, 1, 2.0)[0]
    _say(f"  llh {x[0]:.15g} vs {y[0]:.15g}; beta {x[4]:.15g} vs {y[4]:.15g}")
    assert x[0] == pytest.approx(y[0], rel=1e-8)
    for k in (2, 3, 5, 6):
        assert np.abs(x[k] - y[k]).max() <= 1e-6 * np.abs(y[k]).max(), k
    assert x[4] == pytest.approx(y[4], rel=1e-6)


def test_per_point_outputs_keep_input_order():
    """MAP projections read the per-point state in input order: identical
    with and without the ordering after an E-step and an M-step."""
    _say("case: per-point outputs in input order")
    prob = _problem("fp")
    X = datagen.latent_basis(3, 1728, 64, 1.0)[0]
    res = []
    for env in ({}, {"DG_MIXED_ORDER": "0"}):
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            with dg.Engine(precision=dg.MIXED, lam=1e-3) as e:
                e.upload(*prob)
                e.estep(True, 2.0)
                e.mstep()
                e.estep(True, 2.0)
                res.append(e.map(X))
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
    # one M-step from states that differ by the order of the sums (W, Y about
    # 5e-7): the posterior means move by about 1e-6 (measured 1.1e-6)
    assert np.abs(res[0] - res[1]).max() <= 1e-5, np.abs(res[0] - res[1]).max()
