"""Hierarchical DA (-H): the oracle's HDA loop (oracle.anneal_hda, a restatement
of dagtm.cxx:1508-1795, 1909-2164) pinned against code it does not share.

  - with Kmax at or below the starting K it must reproduce the C oracle's EXP
    loop (orc_anneal) row for row: same queue, convergence and cooling logic;
  - each expansion doubles the grid side, K' = (2 round(K^(1/L)))^L, builds the
    basis with the same M (checked against orc_stp), and restarts from
    Y = FI' W with the last M-step's W: the first E-step after it must equal a
    fresh model built from that state;
  - an expansion fires only at the end of a temperature, never past Kmax.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _start(L=2, K=25, M=9, seed=5):
    T = np.load(os.path.join(GOLDEN, "oil_state.npz"))["T"]
    X, FI, W, Y, b0 = O.stp3_rnd(T, L, K, M, 2.0, seed)
    return T, X, FI, W, Y, b0


def test_hda_without_room_is_the_exp_loop():
    T, X, FI, W, Y, b0 = _start()
    t0 = O.ctemp(T) / 0.9 ** 10
    kw = dict(alpha=0.9, eps=1e-6, start_temp=t0, maxloop=200)
    res, rows = O.anneal_hda(T, FI, Y, b0, W, 2, 9, 2.0, Kmax=25, solver=O.SOLVE_CHOL, **kw)
    m = O.Model(T, FI, Y, b0, solver=O.SOLVE_CHOL)
    ro, orows = m.anneal(isDA=True, schedule=O.SCHED_EXP, **kw)
    assert res["nexp"] == 0 and len(rows) == len(orows) == ro.nloop
    for a, b in zip(rows, orows):
        assert a[0] == b[0] and a[1] == b[1] and a[2] == 25
        assert a[3:] == b[2:]                     # temp, nfree, llh, diff, beta: bit-identical


def test_hda_expansions():
    T, X, FI, W, Y, b0 = _start()
    t0 = O.ctemp(T) / 0.9 ** 10
    snaps = []
    res, rows = O.anneal_hda(T, FI, Y, b0, W, 2, 9, 2.0, Kmax=300, eps2=1e-3, alpha=0.9,
                             eps=1e-6, start_temp=t0, maxloop=300, solver=O.SOLVE_CHOL, X=X,
                             on_expand=lambda *a: snaps.append(a))
    assert res["nexp"] == len(snaps) == 2
    Ks = [25]
    for idx, Mo, Xs, Ys, FIs, Ws, ck, beta in snaps:
        K = FIs.shape[0]
        assert K == Ks[-1] and Xs.shape == (K, 2) and Mo.shape == (1000, 2)
        assert ck.shape == (K,) and np.all(np.abs(Mo) <= 1 + 1e-12)
        last = rows[idx]                          # row of the last E-step before expanding
        assert last[2] == K
        nxt = rows[idx + 1]                       # first row after it: new K, next temperature
        K2 = (2 * int(round(K ** 0.5))) ** 2
        assert nxt[1] == 1 and nxt[2] == K2 and nxt[3] == pytest.approx(last[3] * 0.9, rel=1e-15)
        X2, FI2 = O.stp(2, K2, 9, 2.0)
        m = O.Model(T, FI2, FI2 @ Ws, beta, solver=O.SOLVE_CHOL)
        llh, nfree, _ = m.estep(1, nxt[3])
        assert nxt[5] == pytest.approx(llh, rel=1e-12) and nxt[4] == pytest.approx(nfree, rel=1e-12)
        Ks.append(K2)
    assert Ks == [25, 100, 400] and res["K"] == 400
    ks = [r[2] for r in rows]
    assert all(a <= b for a, b in zip(ks, ks[1:]))
    # Kmax = 100 allows the first doubling only
    res2, _ = O.anneal_hda(T, FI, Y, b0, W, 2, 9, 2.0, Kmax=100, eps2=1e-3, alpha=0.9,
                           eps=1e-6, start_temp=t0, maxloop=300, solver=O.SOLVE_CHOL)
    assert res2["nexp"] == 1 and res2["K"] == 100


def test_k_doubling_rule():
    for L, K, K2 in ((1, 7, 14), (2, 25, 100), (3, 8, 64), (3, 125, 1000), (2, 16, 64)):
        assert (2 * int(round(K ** (1.0 / L)))) ** L == K2
