"""HDA re-set-up on the GPU (dg_set_basis) against the FP64 oracle: after a new
basis the engine's state must be the oracle model built from FI' and the same
W and beta (dagtm.cxx:2101-2160)."""
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


@pytest.mark.parametrize("prec", [dg.FP64, dg.MIXED])
@pytest.mark.parametrize("N,D,L,K,M,fp", [
    (1000, 12, 2, 25, 9, False),       # 25 -> 100 centres
    (600, 166, 3, 27, 8, True),        # 27 -> 216 centres, fingerprint width
])
def test_set_basis_matches_oracle(prec, N, D, L, K, M, fp):
    _say(f"case: set_basis prec={prec} N={N} D={D} K={K}->{(2 * round(K ** (1 / L))) ** L}")
    T = datagen.fingerprints(N, D, 4).astype(np.float64) if fp else datagen.gmm(N, D, 4)
    X, FI = O.stp(L, K, M, 2.0)
    W0 = np.linalg.lstsq(FI, T.mean(0) + 0.1 * np.random.default_rng(2).standard_normal((K, D)),
                         rcond=None)[0]
    Y0 = FI @ W0
    b0, temp = 3.0, 2.0
    K2 = (2 * int(round(K ** (1.0 / L)))) ** L
    X2, FI2 = O.stp(L, K2, M, 2.0)
    with dg.Engine(precision=prec) as e:
        e.upload(T, FI, Y0, b0)
        with pytest.raises(dg.DgError) as ex:
            e.set_basis(FI2)                   # no M-step yet and no W given
        assert ex.value.code == dg.DG_ESTATE
        with pytest.raises(dg.DgError) as ex:
            e.set_basis(FI2[:, :-1], W0[:-1])   # M+1 may not change
        assert ex.value.code == dg.DG_EINVAL
        e.estep(isDA=True, temp=temp)
        beta = e.mstep()
        W, Y, _ = e.download()
        e.set_basis(FI2)
        assert e.K == K2
        W2, Y2, b2 = e.download()
        assert b2 == beta and np.array_equal(W2, W)
        assert np.allclose(Y2, FI2 @ W, rtol=1e-12, atol=1e-12 * np.abs(Y2).max())
        llh, nfree, _ = e.estep(isDA=True, temp=temp * 0.9)
        m = O.Model(T, FI2, FI2 @ W, beta, solver=O.SOLVE_CHOL)
        lo, fo, _ = m.estep(1, temp * 0.9)
        tol = 1e-10 if prec == dg.FP64 else 1e-5     # mixed: the E-step tolerance for float data
        assert llh == pytest.approx(lo, rel=tol) and nfree == pytest.approx(fo, rel=tol)
        # the M-step runs at the new K, and the critical-temperature weights restart at zero
        c, _ = e.next_ctemp(temp, 1e-6)
        assert c == np.inf
        e.mstep()
        m.mstep()
        _, Y3, b3 = e.download()
        assert b3 == pytest.approx(m.beta, rel=1e-9 if prec == dg.FP64 else 1e-4)
        # explicit W: back to the first basis
        e.set_basis(FI, W0)
        _, Y4, _ = e.download()
        assert e.K == K and np.allclose(Y4, Y0, rtol=1e-12, atol=1e-12 * np.abs(Y0).max())
