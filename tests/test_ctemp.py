"""Critical temperatures for AUTO cooling (mpicomm.c:561-686, comm.c:2149-2161).

Pins, none of which shares code with what it checks:
  - the library's host eigenvalue routine (da-gtm_amd/csrc/dg_eig.cpp, built
    here with g++ alone) against numpy's eigvalsh;
  - the oracle's orc_next_ctemp against H assembled in numpy and numpy's
    eigvalsh with the reference's limit / max rule;
  - the block identity eig([[A, B], [B, A]]) = eig(A + B) u eig(A - B), which
    the library uses to split the 2D x 2D problem into two D x D ones;
  - a closed form: one centre holding all points (K = 1, R = 1, y = mean)
    has S2 = S1, so its critical temperature is beta * lambda_max(C^T C / N),
    which at beta = N D / sum C^2 is the first critical temperature orc_ctemp
    (itself pinned against numpy in test_oracle.py).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EIG_SRC = os.path.join(ROOT, "da-gtm_amd", "csrc", "dg_eig.cpp")


@pytest.fixture(scope="module")
def eig(tmp_path_factory):
    so = tmp_path_factory.mktemp("eig") / "libdgeig.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", EIG_SRC, "-o", str(so)],
                   check=True, capture_output=True, timeout=300)
    L = C.CDLL(str(so))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.dg_sym_eigvals.argtypes = [C.c_int, dp, C.c_long, dp]
    L.dg_sym_eigvals.restype = C.c_int

    def f(A, n=None):
        A = np.ascontiguousarray(A, dtype=np.float64)
        n = A.shape[0] if n is None else n
        w = np.zeros(n)
        assert L.dg_sym_eigvals(n, A, A.shape[1], w) == 0
        return np.sort(w)
    return f


def _sym(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return A + A.T


@pytest.mark.parametrize("n", [1, 2, 3, 12, 37, 166])
def test_eigvals_random(eig, n):
    A = _sym(n, n)
    ref = np.linalg.eigvalsh(A)
    assert np.abs(eig(A) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_eigvals_structured(eig):
    rng = np.random.default_rng(5)
    n = 40
    # repeated eigenvalues, exact zeros off the diagonal, a split tridiagonal
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.repeat([3.0, -1.0, 0.0, 7.5], n // 4)
    cases = [np.eye(n), np.diag(np.arange(n, dtype=float)), (Q * lam) @ Q.T]
    Tm = np.diag(rng.standard_normal(n)) + np.diag(np.r_[rng.standard_normal(9), 0.0,
                                                         rng.standard_normal(n - 11)], 1)
    cases.append(Tm + np.triu(Tm, 1).T)
    # low-rank PSD (the shape of a weighted scatter matrix) and wide range
    X = rng.standard_normal((5, n))
    cases.append(X.T @ X)
    cases.append((Q * np.logspace(-12, 6, n)) @ Q.T)
    for A in cases:
        A = (A + A.T) / 2
        ref = np.linalg.eigvalsh(A)
        assert np.abs(eig(A) - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("seed", [0, 2, 25, 36])
def test_eigvals_rank_deficient_scatter(eig, seed):
    """Weighted scatter matrices of a few binary rows, n = 166: most eigenvalues
    are zero up to rounding, which a deflation test relative to the diagonal
    neighbours alone never accepts."""
    rng = np.random.default_rng(seed)
    n, r = 166, int(rng.integers(2, 12))
    X = (rng.random((r, n)) < 0.3).astype(float) - rng.random(n) * 0.5
    A = (X * rng.random(r)[:, None]).T @ X
    ref = np.linalg.eigvalsh(A)
    assert np.abs(eig(A) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_eigvals_reads_lower_triangle_with_stride(eig):
    A = _sym(9, 1)
    P = np.full((9, 13), np.nan)          # NaN above the diagonal and in the padding
    P[:, :9] = np.where(np.tri(9, dtype=bool), A, np.nan)
    ref = np.linalg.eigvalsh(A)
    assert np.abs(eig(P, 9) - ref).max() < 1e-12 * np.abs(ref).max()


def _numpy_ctemp(T, Y, R, gg, temp, beta, eps):
    K, D = Y.shape
    out = np.zeros(K)
    for k in range(K):
        if gg[k] == 0:
            out[k] = np.inf
            continue
        X = T - Y[k]
        r = R[k][:, None]
        S1 = 2 * (X * r).T @ X
        S2 = (X * r * r).T @ X
        H = np.block([[S1 - S2, -S2], [-S2, S1 - S2]])
        lam = np.linalg.eigvalsh(H)
        limit = (temp - eps) * gg[k] * 2 / beta
        lam = np.where(lam >= limit, 0.0, lam)
        out[k] = beta / 2 / gg[k] * lam.max()
    return out


def test_oracle_next_ctemp_against_numpy(oil_state):
    T, Y, R, beta = oil_state["T"], oil_state["Y"], oil_state["R"], oil_state["beta"][0]
    m = O.Model(T, oil_state["FI"], Y, beta)
    m.estep(1, 1.0)
    gg = R.sum(1)
    for temp in (3.5, 2.0, 1.2):
        c, ck = m.next_ctemp(temp, 1e-5, gg)
        ref = _numpy_ctemp(T, Y, m.R, gg, temp, beta, 1e-5)
        assert np.allclose(ck, ref, rtol=1e-10, atol=1e-12)
        assert c == ck.max()
        assert np.all(ck[np.isfinite(ck)] < temp)


def test_oracle_next_ctemp_zero_weight(oil_state):
    T, Y, beta = oil_state["T"], oil_state["Y"], oil_state["beta"][0]
    m = O.Model(T, oil_state["FI"], Y, beta)
    m.estep(1, 1.0)
    gg = m.R.sum(1)
    gg[7] = 0.0
    c, ck = m.next_ctemp(2.0, 1e-5, gg)
    assert ck[7] == np.inf and c == np.inf     # any empty centre forces the EXP step
    # below 1e-250 the reference scales H before the eigensolver; the result is
    # the same function of g up to rounding
    g2 = m.R.sum(1)
    g2[3] = 1e-260
    _, ck2 = m.next_ctemp(2.0, 1e-5, g2)
    ref = _numpy_ctemp(T, Y, m.R, g2, 2.0, beta, 1e-5)
    assert np.allclose(ck2, ref, rtol=1e-9)


def test_block_identity():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((50, 6))
    r = rng.uniform(0, 1, 50)[:, None]
    S1 = 2 * (X * r).T @ X
    S2 = (X * r * r).T @ X
    H = np.block([[S1 - S2, -S2], [-S2, S1 - S2]])
    both = np.sort(np.r_[np.linalg.eigvalsh(S1), np.linalg.eigvalsh(S1 - 2 * S2)])
    assert np.abs(np.linalg.eigvalsh(H) - both).max() < 1e-12 * np.abs(both).max()


def test_single_centre_is_first_critical_temperature():
    rng = np.random.default_rng(9)
    T = rng.normal(size=(400, 5)) * [1, 2, 3, 0.5, 1.5] + 1.0
    N, D = T.shape
    C0 = T - T.mean(0)
    beta = N * D / (C0 * C0).sum()
    Y = T.mean(0)[None, :]
    FI = np.ones((1, 1))
    m = O.Model(T, FI, Y, beta)
    m.estep(1, 1.0)
    assert np.all(m.R == 1.0)
    c, ck = m.next_ctemp(1e6, 0.0, np.array([float(N)]))
    assert c == pytest.approx(O.ctemp(T), rel=1e-12)
    lmax = np.linalg.eigvalsh(C0.T @ C0 / N).max()
    assert c == pytest.approx(beta * lmax, rel=1e-12)


def test_oracle_anneal_auto():
    """AUTO visits strictly decreasing temperatures, each either the next
    critical temperature or the EXP step, and ends at 1."""
    rng = np.random.default_rng(0)
    T = np.concatenate([rng.normal(c, 0.4, size=(100, 4)) for c in (-2, 0, 2)])
    X, FI, W, Y, b0 = O.stp3_rnd(T, 2, 16, 4, 1.0, 7)
    m = O.Model(T, FI, Y, b0, solver=O.SOLVE_CHOL)
    res, rows = m.anneal(isDA=True, schedule=O.SCHED_AUTO, alpha=0.9, eps=1e-5,
                         start_temp=O.ctemp(T) / 0.9 ** 10, maxloop=5000)
    assert not res.exit_maxloop
    temps = []
    for r in rows:
        if not temps or r[2] != temps[-1]:
            temps.append(r[2])
    assert all(a > b for a, b in zip(temps, temps[1:]))
    assert temps[-1] == 1.0 and res.ntemps == len(temps)
    jumps = [b for a, b in zip(temps, temps[1:]) if b != 1.0 and abs(b - 0.9 * a) > 1e-12 * a]
    assert jumps, "no temperature came from a critical temperature"
    # AUTO differs from EXP on this problem (it jumps to critical points)
    m2 = O.Model(T, FI, Y, b0, solver=O.SOLVE_CHOL)
    res2, _ = m2.anneal(isDA=True, schedule=O.SCHED_EXP, alpha=0.9, eps=1e-5,
                        start_temp=O.ctemp(T) / 0.9 ** 10, maxloop=5000)
    assert res2.ntemps != res.ntemps
