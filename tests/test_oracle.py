"""Pins the CPU FP64 oracle before it is trusted as the checker.

Each test compares the oracle with something it does not share code with:
the reference's own output fixture (data/Oil.out.h5, exported to
tests/golden/oil_state.npz), a textbook routine (scipy softmax/logsumexp,
numpy solvers, numpy's legacy MT19937), a closed form, or brute force.
"""
import math

import numpy as np
import pytest
from scipy.spatial.distance import cdist
from scipy.special import logsumexp, softmax

import oracle as O

OIL_LLH = -0.1692069355836593          # normalised llh of the Oil state (tau=1)
OIL_BETA_ID = 32.335774000459786       # N*D / sum(R o dist(Y,T)) on the Oil state


# --------------------------------------------------------------- fixture pins
def test_grid_and_basis_match_oil(oil_state):
    X, FI = O.stp(3, 125, 8, 2.0)
    assert np.abs(X - oil_state["X"]).max() <= 1e-15
    assert np.abs(FI - oil_state["FI"]).max() <= 1e-15


def test_oil_centres_from_basis(oil_state):
    Y = oil_state["FI"] @ oil_state["W"]
    assert np.abs(Y - oil_state["Y"]).max() / np.abs(oil_state["Y"]).max() < 1e-12


@pytest.mark.parametrize("isDA", [0, 1])
def test_oil_responsibilities(oil_state, isDA):
    Dm = O.dist(oil_state["Y"], oil_state["T"])
    R, g, h, nz = O.resp(Dm, oil_state["beta"][0], isDA, 1.0)
    assert nz == 0
    assert np.abs(R - oil_state["R"]).max() < 1e-13
    assert np.abs(R.sum(0) - 1).max() < 1e-14


def test_oil_llh(oil_state):
    Dm = O.dist(oil_state["Y"], oil_state["T"])
    beta = oil_state["beta"][0]
    R, g, h, _ = O.resp(Dm, beta, 1, 1.0)
    N, D = oil_state["T"].shape
    llh = O.qual(g, N, 125, D, beta, 0, 1.0)
    nfree = O.qual(h, N, 125, D, beta, 1, 1.0)
    assert llh == pytest.approx(OIL_LLH, rel=1e-12)
    assert nfree == pytest.approx(OIL_LLH, rel=1e-12)   # tau=1: -F = llh


def test_oil_map(oil_state):
    M = oil_state["R"].T @ oil_state["X"]
    assert np.abs(M - oil_state["M"]).max() < 1e-14
    m = O.Model(oil_state["T"], oil_state["FI"], oil_state["Y"], oil_state["beta"][0])
    m.estep(1, 1.0)
    assert np.abs(m.map(oil_state["X"]) - oil_state["M"]).max() < 1e-13


def test_oil_beta_identity(oil_state):
    T, Y, R = oil_state["T"], oil_state["Y"], oil_state["R"]
    Dm = O.dist(Y, T)
    N, D = T.shape
    b = O.lib().orc_beta(np.ascontiguousarray(Y), np.ascontiguousarray(T), R.copy(),
                         N, Y.shape[0], D, np.zeros_like(Dm))
    assert b == pytest.approx(OIL_BETA_ID, rel=1e-12)
    # closed form used by the GPU: sum R o d = sum|t|^2 - 2 sum y.RT + sum G|y|^2
    G = R.sum(1); RT = R @ T
    s = (T * T).sum() - 2 * (Y * RT).sum() + (G * (Y * Y).sum(1)).sum()
    assert N * D / s == pytest.approx(OIL_BETA_ID, rel=1e-11)


# --------------------------------------------------------------- distances
def test_dist_against_cdist():
    rng = np.random.default_rng(1)
    Y = rng.normal(size=(37, 13)); T = rng.normal(size=(53, 13))
    assert np.allclose(O.dist(Y, T), cdist(Y, T, "sqeuclidean"), rtol=0, atol=1e-12)


# --------------------------------------------------------------- E-step
@pytest.mark.parametrize("temp", [1.0, 1.7, 25.0])
def test_resp_is_tempered_softmax(temp):
    rng = np.random.default_rng(2)
    Dm = cdist(rng.normal(size=(41, 5)), rng.normal(size=(29, 5)), "sqeuclidean")
    beta = 3.3
    R, g, h, nz = O.resp(Dm, beta, 1, temp)
    assert nz == 0
    assert np.allclose(R, softmax(-beta * Dm / (2 * temp), axis=0), rtol=1e-12, atol=1e-15)
    assert np.allclose(np.log(g), logsumexp(-beta * Dm / 2, axis=0), rtol=1e-13)
    assert np.allclose(np.log(h), logsumexp(-beta * Dm / (2 * temp), axis=0), rtol=1e-13)


def test_resp_underflow_semantics():
    """exp before pow (Q1): an element with beta*d/2 past exp's underflow is
    0 at any temperature; an all-underflow column falls back to 1/K (Q2)."""
    K, N, beta = 4, 3, 2.0
    Dm = np.full((K, N), 1.0)
    Dm[1, 0] = 800.0                    # underflows: beta*d/2 = 800 > 745.13
    Dm[:, 2] = 760.0                    # whole column underflows
    R, g, h, nz = O.resp(Dm, beta, 1, 50.0)
    assert nz == 1
    assert R[1, 0] == 0.0
    assert np.allclose(R[[0, 2, 3], 0], 1 / 3)
    assert np.all(R[:, 2] == 1 / K) and h[2] == K and g[2] == 0.0
    # EM mode: v aliases g, so the patch lands in g
    R, g, h, nz = O.resp(Dm, beta, 0, 1.0)
    assert nz == 1 and g[2] == K
    # DA mode: log g = -inf makes llh -inf
    _, g, _, _ = O.resp(Dm, beta, 1, 50.0)
    assert O.qual(g, N, K, 7, beta, 0, 1.0) == -math.inf


def test_qual_closed_form():
    """All centres equidistant: g = K exp(-beta d/2), so the normalised
    llh is (D/2) log(beta/2pi) - beta d/2; h = K exp(-beta d/(2T)), so
    -F = (D/2) log(beta/2pi) + (T-1) log K - beta d/2."""
    K, N, D, beta, d = 9, 11, 6, 1.7, 2.5
    Dm = np.full((K, N), d)
    for temp in (1.0, 3.0):
        R, g, h, _ = O.resp(Dm, beta, 1, temp)
        ref = (D / 2) * math.log(beta / (2 * math.pi)) - beta * d / 2
        assert O.qual(g, N, K, D, beta, 0, temp) == pytest.approx(ref, rel=1e-14)
        refF = ref + (temp - 1) * math.log(K)
        assert O.qual(h, N, K, D, beta, 1, temp) == pytest.approx(refF, rel=1e-14)


# --------------------------------------------------------------- solvers
def _spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("n", [9, 26, 126])
def test_solvers_match_numpy(n):
    A = _spd(n, 1e6, n)
    B = np.random.default_rng(n).normal(size=(n, 7))
    ref = np.linalg.solve(A, B)
    assert np.allclose(O.solve2(A, B), ref, rtol=1e-7, atol=1e-9)
    Xc, jit = O.chol_solve(A, B)
    assert jit == 0.0
    assert np.allclose(Xc, ref, rtol=1e-7, atol=1e-9)
    ev, V = O.eig_sym(A)
    assert np.allclose(np.sort(ev), np.linalg.eigvalsh(A), rtol=1e-10)


def test_ridge_and_least_squares():
    rng = np.random.default_rng(5)
    A = _spd(20, 1e3, 1); B = rng.normal(size=(20, 3))
    X, _ = O.chol_solve(A, B, ridge=0.5)
    assert np.allclose(X, np.linalg.solve(A + 0.5 * np.eye(20), B), rtol=1e-10)
    F = rng.normal(size=(60, 8)); Z = rng.normal(size=(60, 4))
    assert np.allclose(O.solve_ls(F, Z), np.linalg.lstsq(F, Z, rcond=None)[0], rtol=1e-9)


def test_chol_jitter_on_singular():
    v = np.arange(1.0, 6.0)
    A = np.outer(v, v)                  # rank 1
    X, jit = O.chol_solve(A, np.ones((5, 1)))
    assert jit > 0 and np.all(np.isfinite(X))


# --------------------------------------------------------------- set-up chain
def test_mt19937_known_answers():
    m = O.MT19937(5489)
    first = [m.next_u32() for _ in range(3)]
    assert first == [3499211612, 581869302, 3890346734]
    for _ in range(9996):
        m.next_u32()
    assert m.next_u32() == 4123659995          # C++ std::mt19937 10000th output
    for seed in (7, 23, 42, 123456789):
        bg = np.random.MT19937(); bg._legacy_seeding(seed)   # init_genrand
        ref = bg.random_raw(1000)
        m = O.MT19937(seed)
        assert [m.next_u32() for _ in range(1000)] == list(ref)


def test_yrnd2_transform():
    mu = np.array([0.5, -1.0, 3.0]); sd = np.array([1.0, 0.25, 2.0])
    Y = O.yrnd2(mu, sd, 5, 23)
    bg = np.random.MT19937(); bg._legacy_seeding(23)
    u = bg.random_raw(15).astype(np.float64) / 2.0**32
    up = np.ceil(u / 1e-15) * 1e-15
    ref = (2 * (6 * sd) * up.reshape(5, 3) - 6 * sd) + mu
    assert np.allclose(Y, ref, rtol=0, atol=1e-15)
    assert np.all(np.abs(Y - mu) <= 6 * sd)


def test_mean_sd_bi_ctemp():
    rng = np.random.default_rng(3)
    T = rng.normal(size=(200, 5)) * [1, 2, 3, 0.5, 1.5] + 1.0
    mu, sd = O.mean_sd(T)
    assert np.allclose(mu, T.mean(0), rtol=1e-13)
    assert np.allclose(sd, T.std(0, ddof=1), rtol=1e-12)
    Y = rng.normal(size=(30, 5))
    Dm = cdist(Y, Y, "sqeuclidean"); np.fill_diagonal(Dm, np.inf)
    assert O.bi(Y) == pytest.approx(2 / Dm.min(0).mean(), rel=1e-12)
    C = T - T.mean(0)
    lmax = np.linalg.eigvalsh(C.T @ C / len(T)).max()
    assert O.ctemp(T) == pytest.approx(lmax * T.size / (C * C).sum(), rel=1e-10)


def test_fresh_start_chain():
    rng = np.random.default_rng(4)
    T = rng.normal(size=(300, 4))
    X, FI, W, Y, b0 = O.stp3_rnd(T, 2, 16, 4, 1.0, 11)
    mu, sd = T.mean(0), T.std(0, ddof=1)
    Yr = O.yrnd2(mu, sd, 16, 11)
    # W0 is the least-squares fit of the random Y (comm.c:1133-1139)
    assert np.allclose(W, np.linalg.lstsq(FI, Yr, rcond=None)[0], rtol=1e-8, atol=1e-10)
    assert b0 == pytest.approx(O.bi(Yr), rel=1e-14)     # Q10: from Yrnd


# --------------------------------------------------------------- iteration
def _toy(seed=0, N=300, D=4, K=16, M=4, L=2, s=1.0):
    rng = np.random.default_rng(seed)
    T = np.concatenate([rng.normal(c, 0.4, size=(N // 3, D)) for c in (-2, 0, 2)])
    X, FI, W, Y, b0 = O.stp3_rnd(T, L, K, M, s, 7)
    return T, X, FI, Y, b0


def test_one_iteration_against_numpy():
    T, X, FI, Y, b0 = _toy()
    m = O.Model(T, FI, Y, b0)
    llh, nfree, _ = m.estep(1, 2.0)
    R = m.R
    Dm = cdist(Y, T, "sqeuclidean")
    assert np.allclose(R, softmax(-b0 * Dm / 4.0, axis=0), rtol=1e-10, atol=1e-14)
    m.mstep()
    G = R.sum(1); A = FI.T @ (G[:, None] * FI); B = FI.T @ (R @ T)
    W = np.linalg.solve(A, B); Ynew = FI @ W
    assert np.allclose(m.Y, Ynew, rtol=1e-8, atol=1e-10)
    s = (R * cdist(Ynew, T, "sqeuclidean")).sum()
    assert m.beta == pytest.approx(T.size / s, rel=1e-9)


def test_em_increases_llh():
    T, X, FI, Y, b0 = _toy(1)
    m = O.Model(T, FI, Y, b0)
    prev = -np.inf
    for _ in range(15):
        llh, _, _ = m.estep(0, 1.0)
        assert llh >= prev - 1e-12
        prev = llh
        m.mstep()


def test_anneal_loop_semantics():
    T, X, FI, Y, b0 = _toy(2)
    m = O.Model(T, FI, Y, b0, solver=O.SOLVE_CHOL)
    res, rows = m.anneal(isDA=True, alpha=0.8, eps=1e-4, start_temp=4.0, maxloop=3000)
    assert not res.exit_maxloop
    assert rows[0][5] == pytest.approx(rows[0][3])      # Q7: first diff = qual
    temps = [r[2] for r in rows]
    assert temps[0] == 4.0 and temps[-1] == 1.0          # clamp to 1 on the last pass
    assert all(a >= b for a, b in zip(temps, temps[1:]))
    assert res.nloop == len(rows)
    # EM mode runs one pass at temp 1
    m2 = O.Model(T, FI, Y, b0, solver=O.SOLVE_CHOL)
    res2, rows2 = m2.anneal(isDA=False, maxloop=30)
    assert all(r[2] == 1.0 for r in rows2) and res2.ntemps == 1


def _c1_model(lam=1e-3, K=400, M=25):
    import datagen
    T = datagen.gmm(1000, 12, 42)
    X, FI = datagen.latent_basis(2, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, FI, Y0, b0, O.Model(T, FI, Y0, b0, lam=lam, solver=O.SOLVE_CHOL)


def test_anneal_huge_eps_takes_no_mstep():
    """eps = 1e300: every temperature ends on its first |diff| test, so each
    runs one E-step and no M-step, Y and beta stay bit-identical, and EXP from
    8 by 0.5 visits 8, 4, 2, 1, 1 (1 is not below 1; 0.5 is clamped to 1 and
    is the last): five temperatures."""
    T, FI, Y0, b0, m = _c1_model()
    res, rows = m.anneal(isDA=True, alpha=0.5, eps=1e300, start_temp=8.0)
    assert [r[2] for r in rows] == [8.0, 4.0, 2.0, 1.0, 1.0]
    assert res.nloop == 5 and res.ntemps == 5 and not res.exit_maxloop
    assert np.array_equal(m.Y, Y0) and m.beta == b0
    # LINEAR from 2 by 0.4: 2, 1.6, 1.2, 0.8 -> 1 (four temperatures)
    T, FI, Y0, b0, m = _c1_model()
    res, rows = m.anneal(isDA=True, schedule=O.SCHED_LINEAR, delta=0.4, eps=1e300,
                         start_temp=2.0)
    assert [round(r[2], 12) for r in rows] == [2.0, 1.6, 1.2, 1.0]
    assert np.array_equal(m.Y, Y0)


def test_anneal_eps_zero_runs_maxloop_iterations():
    """eps = 0 never breaks: exactly maxloop E-steps and maxloop M-steps (the
    state equals maxloop manual E+M iterations, bit for bit), exit_maxloop."""
    T, FI, Y0, b0, m = _c1_model()
    res, rows = m.anneal(isDA=True, alpha=0.9, eps=0.0, start_temp=3.0, maxloop=7)
    assert res.nloop == 7 and res.exit_maxloop == 1 and res.ntemps == 1 and len(rows) == 7
    m2 = O.Model(T, FI, Y0, b0, lam=1e-3, solver=O.SOLVE_CHOL)
    for _ in range(7):
        m2.estep(1, 3.0)
        m2.mstep()
    assert np.array_equal(m.Y, m2.Y) and m.beta == m2.beta


WINDOW_CASE = dict(lam=1.0, eps=1e-8, nloop=274)


def _window_replay(q, eps, QN=50):
    """The break rule applied to a recorded quality sequence (EM mode, one
    temperature): returns (break index, 'diff' | 'window')."""
    Q = [float(np.finfo(np.float64).max)] * QN
    nxt, prev = 0, 0.0
    with np.errstate(over="ignore", invalid="ignore"):
        for i, x in enumerate(q):
            n = i + 1
            if abs(x - prev) < eps:
                return n, "diff"
            prev = x
            if n > QN:
                old = np.float64(0.0)
                for v in Q:
                    old += v
                Q[nxt] = x
                nxt = (nxt + 1) % QN
                new = np.float64(0.0)
                for v in Q:
                    new += v
                if abs(new / QN - old / QN) < eps:
                    return n, "window"
    return None, None


def test_window_break_on_real_dynamics():
    """A run that ends on the 50-iteration window, not on |diff|: EM on the
    N=1000, D=12 mixture, K=400, M+1=26, s=1, lambda=1 (the ridge lets the
    llh peak near iteration 243 and then fall slowly, so |diff| stays above
    1.5e-8 while the 50-step mean change crosses zero).  With eps = 1e-8 the
    window fires at nloop 274 (|mean change| 4.8e-9 there, 5.0e-8 one step
    earlier); the GPU test runs the same case."""
    T, FI, Y0, b0, m = _c1_model(lam=WINDOW_CASE["lam"])
    res, rows = m.anneal(isDA=False, eps=WINDOW_CASE["eps"])
    assert res.nloop == WINDOW_CASE["nloop"] and not res.exit_maxloop
    q = [r[4] for r in rows]
    assert _window_replay(q, WINDOW_CASE["eps"]) == (WINDOW_CASE["nloop"], "window")
    d = np.abs(np.diff(np.concatenate([[0.0], q])))
    assert d.min() > 1.5 * WINDOW_CASE["eps"]
    assert int(np.argmax(q)) + 1 < WINDOW_CASE["nloop"]       # the llh peaked before


def test_window_replay_checker():
    """The replay rule itself on the scripted alternating sequence (breaks at
    101 on the window) and on a sequence whose |diff| drops below eps."""
    assert _window_replay([1.0 if n % 2 else -1.0 for n in range(1, 300)], 1e-3) == (101, "window")
    assert _window_replay([1.0 - 0.5 ** n for n in range(1, 60)], 1e-3) == (10, "diff")
