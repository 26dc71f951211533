"""The shapes of the largest configuration: K = 64000 centres (L = 3, 40^3 grid),
M + 1 = 1001 basis functions (10^3 RBFs + bias), D = 166 binary fingerprints
from 200 prototypes.  N is cut so the checkers finish in seconds; the
centre-side sizes, which set every buffer, tile count and the M-step
factorisation (global-memory Cholesky, M + 1 > 126), are the real ones.

  - FP64 E-step vs the oracle (N = 2000): G and R T to 1e-10;
  - M-step vs numpy: from the engine's own G and R T, A = FI^T diag(G) FI +
    (lambda / beta) I and B = FI^T (R T) solved by numpy must give the
    engine's W, and Y = FI W;
  - mixed vs FP64 GPU (N = 60000): llh to 2e-6 and the centres after one
    M-step to the mixed tolerance."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen
import oracle as O

pytestmark = pytest.mark.gpu

L, K, M, D = 3, 64000, 1000, 166
LAM = 1e-3


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def basis():
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    return X, FI


def _problem(N, seed, basis):
    X, FI = basis
    T = datagen.fingerprints(N, D, seed, nproto=200).astype(np.float64)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, W0, Y0, b0


def test_large_shape_fp64_estep_vs_oracle(basis):
    _say("case: K=64000 M+1=1001 D=166 N=2000 FP64 E-step vs oracle")
    T, W0, Y0, b0 = _problem(2000, 3, basis)
    X, FI = basis
    O.set_threads(16)
    m = O.Model(T, FI, Y0, b0, lam=LAM, solver=O.SOLVE_CHOL)
    lo, fo, _ = m.estep(1, 2.0)
    R = m.R
    with dg.Engine(precision=dg.FP64, lam=LAM) as e:
        e.upload(T, FI, Y0, b0)
        llh, nfree, _ = e.estep(isDA=True, temp=2.0)
        G, RT = e.stats()
    assert llh == pytest.approx(lo, rel=1e-11) and nfree == pytest.approx(fo, rel=1e-11)
    Gr = R.sum(1)
    assert np.abs(G - Gr).max() <= 1e-10 * Gr.max()
    RTr = R @ T
    assert np.abs(RT - RTr).max() <= 1e-10 * np.abs(RTr).max()


def test_large_shape_mstep_vs_numpy(basis):
    _say("case: K=64000 M+1=1001 M-step (global-memory Cholesky) vs numpy")
    T, W0, Y0, b0 = _problem(20000, 4, basis)
    X, FI = basis
    with dg.Engine(precision=dg.FP64, lam=LAM) as e:
        e.upload(T, FI, Y0, b0)
        e.estep(isDA=True, temp=2.0)
        G, RT = e.stats()
        beta0 = b0
        beta = e.mstep()
        W, Y, _ = e.download()
    A = FI.T @ (G[:, None] * FI) + (LAM / beta0) * np.eye(M + 1)
    B = FI.T @ RT
    Wr = np.linalg.solve(A, B)
    cond = np.linalg.cond(A)
    _say(f"  cond(A) = {cond:.3g}")
    assert np.abs(W - Wr).max() <= 1e-13 * cond * np.abs(Wr).max()
    Yr = FI @ W
    assert np.abs(Y - Yr).max() <= 1e-12 * np.abs(Yr).max()
    # beta from its closed form over the new centres (mpicomm.c:276-302)
    s = (T * T).sum() - 2 * (Y * RT).sum() + (G * (Y * Y).sum(1)).sum()
    assert beta == pytest.approx(T.size / s, rel=1e-9)


def test_large_shape_mixed_vs_fp64(basis):
    _say("case: K=64000 M+1=1001 D=166 N=60000 mixed vs FP64 GPU")
    T, W0, Y0, b0 = _problem(60000, 5, basis)
    X, FI = basis
    out = {}
    for prec in (dg.FP64, dg.MIXED):
        with dg.Engine(precision=prec, lam=LAM) as e:
            e.upload(T, FI, Y0, b0)
            llh, nfree, _ = e.estep(isDA=True, temp=2.0)
            G, RT = e.stats()
            e.mstep()
            _, Y, beta = e.download()
            out[prec] = (llh, nfree, G, RT, Y, beta)
    a, b = out[dg.MIXED], out[dg.FP64]
    assert a[0] == pytest.approx(b[0], rel=2e-6) and a[1] == pytest.approx(b[1], rel=2e-6)
    assert np.abs(a[2] - b[2]).max() <= 2e-5 * b[2].max()
    assert np.abs(a[3] - b[3]).max() <= 2e-5 * np.abs(b[3]).max()
    assert np.abs(a[4] - b[4]).max() <= 1e-4 * np.abs(b[4]).max()
    assert a[5] == pytest.approx(b[5], rel=1e-5)
