"""Critical temperatures on the GPU (dg_next_ctemp) and the AUTO schedule,
against the FP64 oracle (mpicomm.c:561-686 restated in oracle/dagtm_oracle.c).

dg_next_ctemp differs from the oracle only in summation order and in the
eigensolver (two D x D problems instead of one 2D x 2D one), so per-centre
values agree to 1e-9 relative.  A full AUTO run in FP64 must visit the same
temperatures and take the same number of iterations."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen
import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(seed, N, D, K, M, L, fp=False):
    T = (datagen.fingerprints(N, D, seed).astype(np.float64) if fp else datagen.gmm(N, D, seed))
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, X, FI, Y0, b0


def _oracle_at(e, T, FI, temp):
    """Oracle model at the engine's current state, with its E-step done."""
    _, Y, beta = e.download()
    m = O.Model(T, FI, Y, beta)
    m.estep(1, temp)
    return m, beta


@pytest.mark.parametrize("prec", [dg.FP64, dg.MIXED])
@pytest.mark.parametrize("N,D,K,M,L,fp,sel", [
    (1000, 12, 100, 16, 2, False, None),           # several N chunks per centre
    (333, 5, 27, 8, 3, False, None),               # ragged shapes
    (130, 166, 900, 25, 2, True, [0, 1, 301, 607, 608, 899]),   # 6 tile pairs, 2 batches
])
def test_next_ctemp_matches_oracle(prec, N, D, K, M, L, fp, sel):
    _say(f"case: next_ctemp prec={prec} N={N} D={D} K={K}")
    T, X, FI, Y0, b0 = _problem(11, N, D, K, M, L, fp)
    temp = 2.5
    with dg.Engine(precision=prec) as e:
        e.upload(T, FI, Y0, b0)
        for _ in range(3):
            e.estep(isDA=True, temp=temp)
            e.mstep()
        e.estep(isDA=True, temp=temp)
        m, beta = _oracle_at(e, T, FI, temp)
        g = m.R.sum(1)
        if sel is not None:                    # the oracle's 2D x 2D solve is slow at D=166
            g = np.where(np.isin(np.arange(K), sel), g, 0.0)
        for t in (temp, 1.3):
            cg, ckg = e.next_ctemp(t, 1e-5, g)
            co, cko = m.next_ctemp(t, 1e-5, g)
            fin = np.isfinite(cko)
            assert np.array_equal(fin, np.isfinite(ckg))
            assert np.allclose(ckg[fin], cko[fin], rtol=1e-9, atol=1e-12 * t), t
            assert cg == ckg.max()
            _say(f"  temp {t}: ctemp_max gpu {cg!r} oracle {co!r}")


def test_next_ctemp_uses_last_mstep_g():
    _say("case: next_ctemp default weights")
    T, X, FI, Y0, b0 = _problem(3, 600, 8, 36, 9, 2)
    with dg.Engine(precision=dg.FP64) as e:
        e.upload(T, FI, Y0, b0)
        e.estep(isDA=True, temp=3.0)
        c0, ck0 = e.next_ctemp(3.0, 1e-5)
        assert c0 == np.inf and np.all(ck0 == np.inf)      # no M-step yet: G = 0
        G1, _ = e.stats()
        e.mstep()
        e.estep(isDA=True, temp=3.0)
        G2, _ = e.stats()
        assert not np.array_equal(G1, G2)
        c1, ck1 = e.next_ctemp(3.0, 1e-5)
        c2, ck2 = e.next_ctemp(3.0, 1e-5, G1)
        assert np.array_equal(ck1, ck2) and c1 == c2


def test_next_ctemp_needs_estep():
    T, X, FI, Y0, b0 = _problem(3, 200, 4, 9, 4, 2)
    with dg.Engine(precision=dg.FP64) as e:
        e.upload(T, FI, Y0, b0)
        with pytest.raises(dg.DgError) as ex:
            e.next_ctemp(2.0, 1e-5)
        assert ex.value.code == dg.DG_ESTATE
        e.estep(isDA=True, temp=2.0)
        e.mstep()
        with pytest.raises(dg.DgError) as ex:
            e.next_ctemp(2.0, 1e-5)
        assert ex.value.code == dg.DG_ESTATE


def test_anneal_auto_matches_oracle():
    _say("case: AUTO anneal N=1000 D=12 K=100 M=16")
    T, X, FI, Y0, b0 = _problem(42, 1000, 12, 100, 16, 2)
    lam = 1e-3
    # after a jump to a critical temperature the marginal centre keeps that
    # critical value, so AUTO then creeps down by little more than eps per
    # temperature (the reference's behaviour too): bound the run by maxloop
    kw = dict(isDA=True, schedule=O.SCHED_AUTO, alpha=0.9, eps=1e-6,
              start_temp=O.ctemp(T) / 0.9 ** 10, maxloop=200)
    m = O.Model(T, FI, Y0, b0, lam=lam, solver=O.SOLVE_CHOL)
    ro, rows_o = m.anneal(**kw)
    with dg.Engine(precision=dg.FP64, lam=lam) as e:
        e.upload(T, FI, Y0, b0)
        rg, rows_g = e.anneal(**kw)
    _say(f"oracle nloop={ro.nloop} ntemps={ro.ntemps}; gpu nloop={rg.nloop} ntemps={rg.ntemps}")
    assert rg.nloop == ro.nloop and rg.ntemps == ro.ntemps
    a = np.array(rows_g); b = np.array(rows_o)
    for col in (2, 3, 4, 6):        # temp, nfree, llh, beta
        assert np.allclose(a[:, col], b[:, col], rtol=1e-8, atol=1e-10), col
    temps = sorted(set(a[:, 2]), reverse=True)
    assert len(temps) > 2
    assert any(abs(y - 0.9 * x) > 1e-9 * x for x, y in zip(temps, temps[1:]) if y != 1.0)


def test_anneal_auto_mixed_follows_schedule():
    _say("case: AUTO anneal mixed N=1000 D=12 K=100 M=16")
    T, X, FI, Y0, b0 = _problem(42, 1000, 12, 100, 16, 2)
    with dg.Engine(precision=dg.MIXED, lam=1e-3) as e:
        e.upload(T, FI, Y0, b0)
        r, rows = e.anneal(isDA=True, schedule=dg.SCHED_AUTO, alpha=0.9, eps=1e-6,
                           start_temp=O.ctemp(T) / 0.9 ** 10, maxloop=200)
    temps = []
    for row in rows:
        if not temps or row[2] != temps[-1]:
            temps.append(row[2])
    assert r.nloop == len(rows) and r.ntemps == len(temps) and len(temps) > 2
    assert all(x - y > 0.99e-6 for x, y in zip(temps, temps[1:]))    # each step is at least eps
    assert any(abs(y - 0.9 * x) > 1e-9 * x for x, y in zip(temps, temps[1:]) if y != 1.0)
