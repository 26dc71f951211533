"""Many independent instances through dg_batch_anneal: every kernel is
deterministic, so each instance's result must equal, bit for bit, the same
instance annealed alone in one context -- in the batched mode (all
instances' steps as batched launches over a device work queue, replayed as
hipGraphs, or launched directly) and with host workers, whatever order the
work items or the instances are taken in."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# (N, D, L, K, M, binary, maxloop, start temp)
CASES = [(3000, 12, 2, 400, 25, False, 40, 4.0), (1200, 32, 2, 1600, 25, False, 25, 3.0),
         (5000, 64, 2, 400, 25, False, 30, 2.5), (2000, 166, 3, 1000, 27, True, 20, 2.0),
         (800, 12, 2, 400, 25, False, 60, 1.5), (4000, 166, 2, 400, 25, True, 15, 3.5),
         (1500, 64, 3, 343, 27, False, 35, 2.0)]


def _instances():
    out = []
    for i, (N, D, L, K, M, binary, maxloop, t0) in enumerate(CASES):
        T = (datagen.fingerprints(N, D, 100 + i).astype(np.float64) if binary
             else datagen.gmm(N, D, 100 + i))
        X, FI = datagen.latent_basis(L, K, M, 1.0)
        W0, Y0, b0 = datagen.pca_init(T, FI, X)
        out.append((T, FI, Y0, b0, dict(maxloop=maxloop, start_temp=t0)))
    return out


@pytest.mark.parametrize("prec,workers", [(dg.FP64, 3), (dg.MIXED, 4), (dg.MIXED, 0),
                                          (dg.MIXED, -1), (dg.MIXED_R1, 0)])
def test_batch_equals_sequential(prec, workers):
    _say(f"case: batch of {len(CASES)} instances, precision {prec}, {workers} workers")
    probs = _instances()
    kw = dict(alpha=0.9, eps=1e-6)
    out, ph = dg.batch_anneal(probs, precision=prec, lam=1e-3, workers=workers, **kw)
    assert ph[0] > 0 and ph[1] > 0
    for i, (p, (st, r, W, Y)) in enumerate(zip(probs, out)):
        assert st == dg.DG_OK, (i, st)
        with dg.Engine(precision=prec, lam=1e-3) as e:
            e.upload(*p[:4])
            rs, rows = e.anneal(isDA=True, **kw, **p[4])
            Ws, Ys, bs = e.download()
        assert (r.nloop, r.ntemps, r.exit_maxloop) == (rs.nloop, rs.ntemps, rs.exit_maxloop), i
        assert (r.llh, r.nfree, r.beta, r.temp) == (rs.llh, rs.nfree, rs.beta, rs.temp), i
        assert np.array_equal(W, Ws) and np.array_equal(Y, Ys), i
        _say(f"  instance {i}: nloop {r.nloop} llh {r.llh:.9g} identical")


def test_batch_item_errors_are_per_item():
    _say("case: batch with one invalid instance")
    probs = _instances()[:3]
    T, FI, Y0, b0, kw = probs[1]
    Tbad = np.zeros((50, 200))                   # D > 192 is rejected by dg_upload
    probs.insert(1, (Tbad, FI, np.zeros((FI.shape[0], 200)), 1.0, kw))
    out, _ = dg.batch_anneal(probs, precision=dg.FP64, workers=2, alpha=0.9, eps=1e-6)
    assert [o[0] for o in out] == [dg.DG_OK, dg.DG_EINVAL, dg.DG_OK, dg.DG_OK]
    assert all(o[1].nloop > 0 for i, o in enumerate(out) if i != 1)


def test_tuned_chunk_count_is_used_from_the_table(tmp_path, monkeypatch):
    """dg_tune keeps the fastest pass-2 chunk count and records it; a context
    uploaded with DG_TUNE_FILE naming that table uses the same count, so its
    E-step is bit-identical to the tuned context's."""
    _say("case: pass-2 chunk tuning table")
    T, FI, Y0, b0, _ = _instances()[2]
    table = str(tmp_path / "tune.txt")
    with dg.Engine(precision=dg.MIXED) as e:
        e.upload(T, FI, Y0, b0)
        nc = e.tune(table)
        ref = e.estep(True, 2.0)
    line = open(table).read().split()
    assert [int(x) for x in line] == [dg.MIXED, T.shape[1], FI.shape[0],
                                      int(np.ceil(np.log2(T.shape[0]))), nc]
    assert nc > 0 and nc % 8 == 0
    monkeypatch.setenv("DG_TUNE_FILE", table)
    with dg.Engine(precision=dg.MIXED) as e:
        e.upload(T, FI, Y0, b0)
        got = e.estep(True, 2.0)
    assert got == ref
    with dg.Engine(precision=dg.FP64) as e:       # no tuned choice in FP64
        e.upload(T, FI, Y0, b0)
        assert e.tune(None) == 0
        l64 = e.estep(True, 2.0)[0]
    assert got[0] == pytest.approx(l64, rel=1e-5)


@pytest.fixture(scope="module")
def c4():
    _say("generating the 64-instance batched workload")
    return datagen.batch_instances()


def test_c4_batch_one_iteration(c4):
    """The 64-instance workload (N log-uniform in [1e3, 2e5], D in {12, 32, 64,
    166}, K in {400, 1600, 8000}), one DA E+M iteration per instance at its
    start temperature in one batched step.  Each instance must equal its own
    single-context run bit for bit.  Against FP64 -- the oracle where it can
    hold the instance (N K <= 4e7), else the FP64 GPU path:
      - binary data (exact in FP16): beta 1e-6, Y 1e-4 of max |Y|, and
        sum_k G_k = N to 1e-6;
      - Gaussian-mixture float data (both operands split): beta 5e-4, Y 1e-4,
        sum_k G_k = N to 1e-4.  Their exponents reach a few thousand (log2
        units: beta |t|^2 / 2 with |t|^2 up to ~300), so the FP32 roundings of
        pass 1's normaliser and of pass 2's recomputed exponents differ by up
        to ~1e-4 relative per element; measured on this workload: sum G 2e-5,
        beta 1.7e-4, Y 2.7e-5 at worst."""
    import oracle as O
    O.set_threads(16)
    out, ph = dg.batch_anneal(c4, precision=dg.MIXED, lam=1e-3, workers=0, alpha=0.99, eps=0.0,
                              maxloop=1)
    errs = []
    for i, (p, (st, r, W, Y)) in enumerate(zip(c4, out)):
        T, FI, Y0, b0, kw = p
        t0 = kw["start_temp"]
        assert st == dg.DG_OK and r.nloop == 1, (i, st)
        with dg.Engine(precision=dg.MIXED, lam=1e-3) as e:
            e.upload(T, FI, Y0, b0)
            e.estep(True, t0)
            G, _ = e.stats()
            e.mstep()
            Ws, Ys, bs = e.download()
        assert np.array_equal(W, Ws) and np.array_equal(Y, Ys) and r.beta == bs, i
        N, D, K = T.shape[0], T.shape[1], FI.shape[0]
        if N * K <= 4e7:
            m = O.Model(T, FI, Y0, b0, lam=1e-3, solver=O.SOLVE_CHOL)
            m.estep(1, t0)
            m.mstep()
            bref, Yref = m.beta, m.Y
            del m
        else:
            with dg.Engine(precision=dg.FP64, lam=1e-3) as e:
                e.upload(T, FI, Y0, b0)
                e.estep(True, t0)
                bref = e.mstep()
                _, Yref, _ = e.download()
        eg = abs(G.sum() - N) / N
        eb = abs(r.beta - bref) / bref
        ey = float(np.abs(Y - Yref).max() / np.abs(Yref).max())
        errs.append((i, N, D, K, eg, eb, ey))
        _say(f"  instance {i}: N={N} D={D} K={K} sum G rel {eg:.1e} beta {eb:.1e} Y {ey:.1e}")
    _say(f"  64 instances in one batched step ({ph[1]:.1f} ms)")
    for (i, N, D, K, eg, eb, ey) in errs:
        if D == 166:
            assert eg <= 1e-6 and eb <= 1e-6 and ey <= 1e-4, (i, eg, eb, ey)
        else:
            assert eg <= 1e-4 and eb <= 5e-4 and ey <= 1e-4, (i, eg, eb, ey)
