"""The largest configuration at full size: N = 2,000,000 binary fingerprints
(200 prototypes, flip 0.05, seed 7), D = 166, L = 3, K = 64000 (40^3 grid),
M + 1 = 1001 (10^3 RBFs + bias), s = 1, lambda = 1e-3, one DA iteration at
temperature 2.  No K x N buffer exists (it would be 1 TB in FP64); the oracle
cannot run this size, so the checks are size-independent properties and the
FP64 GPU path from the same state:

  - sum_k G_k = N and sum_k (R T)_kd = sum_n T_nd (every point's
    responsibilities sum to 1): FP64 to 1e-10, mixed to 1e-5 relative;
  - a repeated mixed E-step is bit-identical;
  - mixed vs FP64 GPU: llh and -F 2e-6, G and R T 1e-5 of their maxima, beta
    and Y after the M-step 1e-5 (the tolerances of tests/test_gpu_mixed.py);
  - the M-step at M + 1 = 1001 (global-memory Cholesky) finishes and its
    device time is reported."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu

N, D, L, K, M = 2_000_000, 166, 3, 64000, 1000
LAM, TEMP = 1e-3, 2.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


def test_c5_full_size():
    _say("case: c5 full size N=2e6 K=64000 M+1=1001: generating")
    T8 = datagen.fingerprints(N, D, 7, nproto=200)
    colsum = T8.sum(0, dtype=np.int64).astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T8[:200_000].astype(np.float64), FI, X)
    T = T8.astype(np.float64)
    del T8
    out = {}
    for prec in (dg.MIXED, dg.FP64):
        _say(f"  precision {prec}: upload + E-step")
        with dg.Engine(precision=prec, lam=LAM) as e:
            e.upload(T, FI, Y0, b0)
            e.set_timing(True)
            llh, nfree, nz = e.estep(True, TEMP)
            G, RT = e.stats()
            ph = e.timing()
            if prec == dg.MIXED:
                llh2, nfree2, _ = e.estep(True, TEMP)
                G2, RT2 = e.stats()
                assert llh2 == llh and nfree2 == nfree
                assert np.array_equal(G2, G) and np.array_equal(RT2, RT)
            beta = e.mstep()
            tm = e.timing()[3]
            _, Y, _ = e.download()
        _say(f"  precision {prec}: llh {llh!r} nfree {nfree!r} beta {beta!r}; E-passes "
             f"{ph[0]:.1f} + {ph[1]:.1f} ms, M-step (Cholesky of 1001 x 1001) {tm:.2f} ms")
        tol = 1e-10 if prec == dg.FP64 else 1e-5
        assert nz == 0
        assert abs(G.sum() - N) <= tol * N
        assert _rel(RT.sum(0), colsum) <= tol
        out[prec] = (llh, nfree, G, RT, beta, Y)
    l16, f16, G16, RT16, b16, Y16 = out[dg.MIXED]
    l64, f64, G64, RT64, b64, Y64 = out[dg.FP64]
    _say(f"  mixed vs fp64: llh {abs(l16 - l64) / abs(l64):.2e} G {_rel(G16, G64):.2e} "
         f"RT {_rel(RT16, RT64):.2e} beta {abs(b16 - b64) / b64:.2e} Y {_rel(Y16, Y64):.2e}")
    assert l16 == pytest.approx(l64, rel=2e-6) and f16 == pytest.approx(f64, rel=2e-6)
    assert _rel(G16, G64) <= 1e-5 and _rel(RT16, RT64) <= 1e-5
    assert b16 == pytest.approx(b64, rel=1e-5) and _rel(Y16, Y64) <= 1e-5
