"""dg_map_noalloc (the MAP projection staged through the library's fixed
device scratch instead of per-call buffers) against dg_map, bit for bit, on
two calls in a row -- the second with a smaller problem, so stale scratch
contents from the first would show -- in both precisions, plus its size limit."""
import sys

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(N, D, K, M, L, seed):
    T = datagen.gmm(N, D, seed)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, X, FI, Y0, b0


@pytest.mark.parametrize("prec", [dg.FP64, dg.MIXED])
def test_map_noalloc_matches_map(prec):
    for (N, D, K, M, L, seed) in [(20000, 12, 400, 25, 2, 1), (3000, 12, 125, 27, 3, 2)]:
        _say(f"case: map_noalloc vs map, precision {prec}, N={N} K={K} L={L}")
        T, X, FI, Y0, b0 = _problem(N, D, K, M, L, seed)
        with dg.Engine(precision=prec, lam=1e-3) as e:
            e.upload(T, FI, Y0, b0)
            e.estep(True, 2.0)
            e.mstep()
            e.estep(True, 1.5)
            m0 = e.map_noalloc(X)             # first: right after the E-step
            m1 = e.map(X)
            m2 = e.map_noalloc(X)             # second: after dg_map used its own buffers
            assert np.array_equal(m0, m1) and np.array_equal(m2, m1)
            assert np.all(np.isfinite(m1))
            e.mstep()                          # the model moved: both rerun the E-step
            m3 = e.map_noalloc(X)
            m4 = e.map(X)
            assert np.array_equal(m3, m4)


def test_map_noalloc_size_limit():
    """(K + N) L above 2^21 doubles is refused on the host, before any
    device work; dg_map handles the same problem."""
    _say("case: map_noalloc size limit")
    N, K, L = 700000, 400, 3                  # (K + N) * 3 > 2^21
    T = datagen.fingerprints(N, 16, 5).astype(np.float64)
    X, FI = datagen.latent_basis(2, K, 25, 1.0)
    X = np.hstack([X, np.zeros((K, 1))])
    Y0 = np.tile(T[:K].mean(0), (K, 1)) + 0.01 * np.arange(K)[:, None]
    with dg.Engine(precision=dg.FP64) as e:
        e.upload(T, FI, Y0, 4.0)
        e.estep(True, 3.0)
        with pytest.raises(dg.DgError) as ei:
            e.map_noalloc(X)
        assert ei.value.code == dg.DG_EINVAL
        Mo = e.map(X)
        assert Mo.shape == (N, L) and np.all(np.isfinite(Mo))
        Mn = e.map_noalloc(X[:, :2])          # (K + N) * 2 fits
        assert np.array_equal(Mn, e.map(X[:, :2]))
