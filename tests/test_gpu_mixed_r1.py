"""DG_MIXED_R1: the mixed path with R carried as one FP16 term in the R T
product, against the FP64 HIP path.

One FP16 term rounds each R (scaled by 2^14) with relative error <= 2^-12;
the errors are not biased, so over the points of a centre they average out to
~2e-5 of G and R T at the largest entries (measured with numpy on the
fingerprint workload: 1.4e-5 in G, 1.8e-5 in R T, a signed mean of -2.4e-7).
Stated tolerances (relative, FP64 path as the reference value):
  llh / -F of one E-step .......... 5e-6 (binary data), 1e-5 (float data)
  G, RT ........................... 6e-5 (relative to max |.|)
  beta, Y after one M-step ........ 1e-4 (beta = N D / (sum|t|^2 - 2 sum y.RT
                                    + sum G|y|^2) amplifies the G, RT errors by
                                    the cancellation: 3.1e-5 measured at 12
                                    points per centre)
The hi/lo path's tolerances are in test_gpu_mixed.py."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _binary(seed, N, D, K, M, L):
    T = datagen.fingerprints(N, D, seed).astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    rng = np.random.default_rng(seed)
    W = np.linalg.lstsq(FI, T.mean(0) + 0.3 * rng.standard_normal((K, D)) * T.std(0),
                        rcond=None)[0]
    return T, FI, FI @ W, 8.0


def _compare(T, FI, Y, beta, temps, tol_llh):
    out = {}
    for prec in (dg.FP64, dg.MIXED_R1):
        with dg.Engine(precision=prec, lam=1e-3) as e:
            e.upload(T, FI, Y, beta)
            rows = []
            for temp in temps:
                llh, nfree, nz = e.estep(True, temp)
                G, RT = e.stats()
                rows.append((llh, nfree, nz, G, RT))
            b = e.mstep()
            _, Yn, _ = e.download()
            out[prec] = (rows, b, Yn)
    (r64, b64, Y64), (r16, b16, Y16) = out[dg.FP64], out[dg.MIXED_R1]
    for a, c in zip(r16, r64):
        _say(f"  llh rel {abs(a[0] - c[0]) / abs(c[0]):.2e}  G {_rel(a[3], c[3]):.2e}  "
             f"RT {_rel(a[4], c[4]):.2e}")
        assert a[2] == c[2]
        assert a[0] == pytest.approx(c[0], rel=tol_llh)
        assert a[1] == pytest.approx(c[1], rel=tol_llh)
        assert _rel(a[3], c[3]) < 6e-5 and _rel(a[4], c[4]) < 6e-5
    assert b16 == pytest.approx(b64, rel=1e-4)
    assert _rel(Y16, Y64) < 1e-4


@pytest.mark.parametrize("N,D,K,M,L,temp", [
    (20000, 166, 1000, 64, 3, 3.0),
    (5000, 166, 1000, 64, 3, 1.0),
    (3001, 64, 343, 27, 3, 2.0),
])
def test_r1_binary_vs_fp64(N, D, K, M, L, temp):
    _say(f"case: R1 binary vs fp64 N={N} D={D} K={K} temp={temp}")
    _compare(*_binary(N + K, N, D, K, M, L), (temp,), 5e-6)


@pytest.mark.parametrize("N,D,K,M,L", [(1000, 12, 400, 25, 2), (2500, 40, 216, 27, 3)])
def test_r1_float_vs_fp64(N, D, K, M, L):
    _say(f"case: R1 float vs fp64 N={N} D={D} K={K}")
    T = datagen.gmm(N, D, 42)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    _compare(T, FI, Y0, b0, (4.0, 1.0), 1e-5)
