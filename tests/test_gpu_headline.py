"""The headline workload itself (N = 10^6 binary fingerprints from 50
prototypes, D = 166, L = 3 with K = 8000 centres, M + 1 = 126, lambda = 1e-3,
DA at temperature 2), checked through properties that hold at any size, since
the oracle cannot run it in seconds:

  - every point's responsibilities sum to 1, so sum_k G_k = N and
    sum_k (R T)_kd = sum_n T_nd: FP64 path to 1e-10 relative, mixed path to
    1e-5 (FP32 partial sums);
  - two E-steps from one state give bit-identical llh, G and R T (every
    reduction has a fixed order);
  - mixed against FP64 GPU from one state: llh to 2e-6 relative, G and R T to
    1e-5 of their largest entry, and after one M-step beta to 1e-5 and Y to
    1e-5 of max |Y| (the tolerances the small-size mixed tests state)."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu

N, D, L, K, M, LAM, TEMP = 1_000_000, 166, 3, 8000, 125, 1e-3, 2.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem():
    T = datagen.fingerprints(N, D, 7).astype(np.float64)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, FI, Y0, b0


@pytest.fixture(scope="module")
def runs(problem):
    """One E-step (twice) and one M-step per precision from the same state."""
    T, FI, Y0, b0 = problem
    out = {}
    for prec in (dg.FP64, dg.MIXED):
        _say(f"[headline] precision {prec}: upload N={N} D={D} K={K}")
        with dg.Engine(precision=prec, lam=LAM) as e:
            e.upload(T, FI, Y0, b0)
            r1 = e.estep(True, TEMP)
            G1, RT1 = e.stats()
            r2 = e.estep(True, TEMP)
            G2, RT2 = e.stats()
            beta = e.mstep()
            _, Y, _ = e.download()
        out[prec] = dict(r1=r1, r2=r2, G1=G1, RT1=RT1, G2=G2, RT2=RT2, beta=beta, Y=Y)
        _say(f"[headline] precision {prec}: llh {r1[0]:.12g} beta {beta:.12g}")
    return out


@pytest.mark.parametrize("prec,tol", [(dg.FP64, 1e-10), (dg.MIXED, 1e-5)])
def test_responsibilities_sum_to_one(problem, runs, prec, tol):
    T = problem[0]
    r = runs[prec]
    assert abs(r["G1"].sum() - N) <= tol * N, (prec, r["G1"].sum())
    colsum = T.sum(axis=0)
    err = np.abs(r["RT1"].sum(axis=0) - colsum).max() / colsum.max()
    assert err <= tol, (prec, err)


@pytest.mark.parametrize("prec", [dg.FP64, dg.MIXED])
def test_repeat_estep_is_bit_identical(runs, prec):
    r = runs[prec]
    assert r["r1"] == r["r2"]
    assert np.array_equal(r["G1"], r["G2"])
    assert np.array_equal(r["RT1"], r["RT2"])


def test_mixed_matches_fp64(runs):
    f, m = runs[dg.FP64], runs[dg.MIXED]
    llh64, llh16 = f["r1"][0], m["r1"][0]
    assert abs(llh16 - llh64) <= 2e-6 * abs(llh64), (llh16, llh64)
    assert abs(m["r1"][1] - f["r1"][1]) <= 2e-6 * abs(f["r1"][1])     # -F (DA quality)
    assert m["r1"][2] == f["r1"][2]                                    # zero columns
    gerr = np.abs(m["G1"] - f["G1"]).max() / np.abs(f["G1"]).max()
    rterr = np.abs(m["RT1"] - f["RT1"]).max() / np.abs(f["RT1"]).max()
    assert gerr <= 1e-5 and rterr <= 1e-5, (gerr, rterr)
    assert abs(m["beta"] - f["beta"]) <= 1e-5 * f["beta"], (m["beta"], f["beta"])
    yerr = np.abs(m["Y"] - f["Y"]).max() / np.abs(f["Y"]).max()
    assert yerr <= 1e-5, yerr
