"""The mid-size configuration against the oracle: binary fingerprints
(fingerprints(100000, 166, seed 7): 50 prototypes, bit p ~ Beta(0.5, 2),
flip 0.05), L = 3, K = 8000 (20^3 grid), M + 1 = 126 (5^3 RBFs + bias),
s = 1, lambda = 1e-3, one DA E+M iteration at temperature 2.

The oracle materialises two K x N FP64 matrices, so N is the first 16000
points (2 GB of host memory, about a second of oracle time with 16 threads);
K, D, M + 1 and the generator are the configuration's own.

  - FP64 GPU vs the oracle: llh and -F to 1e-11 relative, G and R T to 1e-10
    of their maxima, beta and Y after the M-step to 1e-9;
  - mixed vs the oracle: llh and -F 2e-6, G and R T 1e-5, beta and Y 1e-5
    (the mixed-path tolerances stated in tests/test_gpu_mixed.py)."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen
import oracle as O

pytestmark = pytest.mark.gpu

L, K, M, D, NS = 3, 8000, 125, 166, 16000
LAM, TEMP = 1e-3, 2.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2():
    T = datagen.fingerprints(100_000, D, 7)[:NS].astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    O.set_threads(16)
    m = O.Model(T, FI, Y0, b0, lam=LAM, solver=O.SOLVE_CHOL)
    llh, nfree, nz = m.estep(1, TEMP)
    R = m.R
    G, RT = R.sum(1), R @ T
    del R
    m.mstep()
    ref = dict(llh=llh, nfree=nfree, nz=nz, G=G, RT=RT, beta=m.beta, Y=m.Y.copy())
    del m
    return T, FI, Y0, b0, ref


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.mark.parametrize("prec,tl,ts,tm", [(dg.FP64, 1e-11, 1e-10, 1e-9),
                                           (dg.MIXED, 2e-6, 1e-5, 1e-5)])
def test_c2_one_iteration_vs_oracle(c2, prec, tl, ts, tm):
    T, FI, Y0, b0, ref = c2
    _say(f"case: c2 K=8000 M+1=126 D=166 N={NS} precision {prec}")
    with dg.Engine(precision=prec, lam=LAM) as e:
        e.upload(T, FI, Y0, b0)
        llh, nfree, nz = e.estep(True, TEMP)
        G, RT = e.stats()
        beta = e.mstep()
        _, Y, _ = e.download()
    _say(f"  llh {llh!r} oracle {ref['llh']!r}; beta {beta!r} oracle {ref['beta']!r}")
    assert nz == ref["nz"] == 0
    assert llh == pytest.approx(ref["llh"], rel=tl)
    assert nfree == pytest.approx(ref["nfree"], rel=tl)
    assert _rel(G, ref["G"]) <= ts and _rel(RT, ref["RT"]) <= ts
    assert beta == pytest.approx(ref["beta"], rel=tm)
    assert _rel(Y, ref["Y"]) <= tm
