"""Many independent instances through dg_batch_anneal: every kernel is
deterministic, so each instance's result must equal, bit for bit, the same
instance annealed alone in one context -- whatever the worker count and the
order in which the workers picked the instances up."""
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


@pytest.mark.parametrize("prec,workers", [(dg.FP64, 3), (dg.MIXED, 4)])
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
