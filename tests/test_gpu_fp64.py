"""Parity of the FP64 HIP path with the CPU oracle and the reference fixture.

Every call goes through the C ABI (libdagtm_gpu.so).  Bars: the reference
fixture at 1e-12..1e-13; oracle parity at 1e-10 relative for per-step
quantities (the two differ only in summation order); full annealing runs
must take the same number of iterations with llh/-F/beta within 1e-8."""
import math
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen
import oracle as O

pytestmark = pytest.mark.gpu

OIL_LLH = -0.1692069355836593


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(**kw):
    return dg.Engine(precision=dg.FP64, **kw)


def test_oil_fixture_estep(oil_state):
    _say("case: oil fixture E-step (N=1000 D=12 K=125 M1=9)")
    s = oil_state
    with _engine() as e:
        e.upload(s["T"], s["FI"], s["Y"], s["beta"][0])
        llh, nfree, nz = e.estep(isDA=True, temp=1.0)
        assert nz == 0
        assert llh == pytest.approx(OIL_LLH, rel=1e-12)
        assert nfree == pytest.approx(OIL_LLH, rel=1e-12)
        R = e.resp_dense()
        assert np.abs(R - s["R"]).max() < 1e-13
        G, RT = e.stats()
        assert np.allclose(G, s["R"].sum(1), rtol=1e-12, atol=1e-12)
        assert np.allclose(RT, s["R"] @ s["T"], rtol=1e-11, atol=1e-11)
        assert np.abs(e.map(s["X"]) - s["M"]).max() < 1e-13


def _rand_problem(seed, N, D, K, M, L, s=1.0):
    T = datagen.gmm(N, D, seed)
    X, FI = datagen.latent_basis(L, K, M, s)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, X, FI, Y0, b0


@pytest.mark.parametrize("N,D,K,M,L,temp,isDA", [
    (1000, 12, 400, 25, 2, 3.7, 1),      # the N=1000, D=12, K=400 case
    (333, 5, 27, 8, 3, 1.0, 1),          # ragged N, K, D
    (2049, 40, 100, 16, 2, 12.5, 1),     # several pass-2 chunks
    (700, 17, 64, 4, 2, 1.0, 0),         # EM mode
    (130, 166, 900, 25, 2, 2.0, 1),      # fingerprint width
])
def test_estep_matches_oracle(N, D, K, M, L, temp, isDA):
    _say(f"case: estep N={N} D={D} K={K} M={M} temp={temp} isDA={isDA}")
    T, X, FI, Y0, b0 = _rand_problem(N + D + K, N, D, K, M, L)
    m = O.Model(T, FI, Y0, b0)
    llh_o, nf_o, nz_o = m.estep(isDA, temp)
    R = m.R
    with _engine() as e:
        e.upload(T, FI, Y0, b0)
        llh, nf, nz = e.estep(isDA=isDA, temp=temp)
        assert nz == nz_o == 0
        assert llh == pytest.approx(llh_o, rel=1e-11, abs=1e-12)
        assert nf == pytest.approx(nf_o, rel=1e-11, abs=1e-12)
        G, RT = e.stats()
        assert np.allclose(G, R.sum(1), rtol=1e-10, atol=1e-12)
        assert np.allclose(RT, R @ T, rtol=1e-10, atol=1e-10)
        assert np.abs(e.resp_dense() - R).max() < 1e-12


def test_underflow_and_zero_columns():
    """Elements whose exp underflows contribute 0 and all-underflow columns
    fall back to R = 1/K, exactly as the reference's exp-then-pow order."""
    _say("case: underflow / zero columns")
    rng = np.random.default_rng(9)
    K, N, D = 16, 200, 3
    Y = rng.normal(size=(K, D))
    T = rng.normal(size=(N, D))
    T[:7] += 40.0                      # far points: every exp underflows at beta=8
    FI = np.ones((K, 1))
    beta = 8.0
    m = O.Model(T, FI, Y, beta)
    for isDA, temp in ((1, 30.0), (0, 1.0)):
        llh_o, nf_o, nz_o = m.estep(isDA, temp)
        with _engine() as e:
            e.upload(T, FI, Y, beta)
            llh, nf, nz = e.estep(isDA=isDA, temp=temp)
            assert nz == nz_o == 7
            if isDA:
                assert llh == -math.inf and llh_o == -math.inf
                assert nf == pytest.approx(nf_o, rel=1e-12)
            else:
                assert llh == pytest.approx(llh_o, rel=1e-12)
            assert np.abs(e.resp_dense() - m.R).max() < 1e-13


def test_one_mstep_matches_oracle():
    _say("case: one M-step N=1000 D=12 K=400 M=25 s=1 lambda=1e-3")
    T, X, FI, Y0, b0 = _rand_problem(42, 1000, 12, 400, 25, 2)
    lam = 1e-3
    m = O.Model(T, FI, Y0, b0, lam=lam, solver=O.SOLVE_CHOL)
    m.estep(1, 5.0)
    m.mstep()
    with _engine(lam=lam) as e:
        e.upload(T, FI, Y0, b0)
        e.estep(isDA=True, temp=5.0)
        beta = e.mstep()
        W, Y, b = e.download()
        assert b == beta
        assert beta == pytest.approx(m.beta, rel=1e-10)
        assert np.allclose(Y, m.Y, rtol=1e-9, atol=1e-9 * np.abs(m.Y).max())
        assert np.allclose(W, m.W, rtol=1e-6, atol=1e-6 * np.abs(m.W).max())


def test_anneal_matches_oracle():
    """Full DA run on the N=1000, D=12, K=400, M=25 (s=1, lambda=1e-3) case:
    identical iteration count and per-iteration llh, -F and beta."""
    _say("case: anneal N=1000 D=12 K=400 M=25")
    T, X, FI, Y0, b0 = _rand_problem(42, 1000, 12, 400, 25, 2)
    lam = 1e-3
    t0 = O.ctemp(T) / 0.9 ** 10
    kw = dict(isDA=True, alpha=0.9, eps=1e-6, start_temp=t0, maxloop=300)
    m = O.Model(T, FI, Y0, b0, lam=lam, solver=O.SOLVE_CHOL)
    ro, rows_o = m.anneal(**kw)
    with _engine(lam=lam) as e:
        e.upload(T, FI, Y0, b0)
        rg, rows_g = e.anneal(**kw)
    _say(f"oracle nloop={ro.nloop} gpu nloop={rg.nloop}")
    assert rg.nloop == ro.nloop and rg.ntemps == ro.ntemps
    a = np.array(rows_g); b = np.array(rows_o)
    for col in (2, 3, 4, 6):        # temp, nfree, llh, beta
        assert np.allclose(a[:, col], b[:, col], rtol=1e-8, atol=1e-10), col


@pytest.mark.parametrize("L,K,M", [(2, 900, 144), (3, 1000, 125)])
def test_mstep_large_basis_matches_oracle(L, K, M):
    """M+1 = 145 takes the global-memory factorisation, M+1 = 126 the LDS one."""
    _say(f"case: M-step K={K} M+1={M + 1}")
    T = datagen.fingerprints(3000, 40, 3).astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    rng = np.random.default_rng(1)
    Y0 = FI @ np.linalg.lstsq(FI, T.mean(0) + 0.2 * rng.standard_normal((K, 40)), rcond=None)[0]
    lam = 1e-2
    m = O.Model(T, FI, Y0, 5.0, lam=lam, solver=O.SOLVE_CHOL)
    m.estep(1, 2.0)
    m.mstep()
    with _engine(lam=lam) as e:
        e.upload(T, FI, Y0, 5.0)
        e.estep(isDA=True, temp=2.0)
        beta = e.mstep()
        W, Y, b = e.download()
    assert beta == pytest.approx(m.beta, rel=1e-9)
    assert np.allclose(Y, m.Y, rtol=1e-8, atol=1e-8 * np.abs(m.Y).max())
