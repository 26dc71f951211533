"""The host set-up chain's symmetric eigensolver (Householder tridiagonal
form, implicit QL; da-gtm_amd/host/dg_init.cpp), which the least-squares W0
(comm.c:1247-1294) and the first critical temperature (mpicomm.c:511-559)
use, checked on the CPU against LAPACK through numpy: eigenvalues,
A v = lambda v residuals and orthonormal vectors, on random, rank-deficient
and RBF normal-equation matrices.  Every element of the chain is computed by
one worker in a fixed order, so the results must be the same bytes for any
number of host threads -- checked here for the eigensolver and, through the
dagtm driver, for the whole fresh-start chain."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import h5util

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "da-gtm_amd", "host")
DAGTM = os.path.join(ROOT, "da-gtm_amd", "dagtm")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("eig") / "eig_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", "-o", exe,
                    os.path.join(ROOT, "tests", "host", "eig_driver.cpp"),
                    os.path.join(HOST, "dg_init.cpp")], check=True)
    return exe


def _eig(driver, tmp_path, A, threads=None, vectors=True):
    n = A.shape[0]
    fin, fout = str(tmp_path / "A.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(np.int32(n).tobytes())
        f.write(np.ascontiguousarray(A, dtype=np.float64).tobytes())
    env = dict(os.environ)
    if threads:
        env["DG_HOST_THREADS"] = str(threads)
    subprocess.run([driver, fin, fout] + ([] if vectors else ["values"]), check=True, env=env)
    raw = open(fout, "rb").read()
    ev = np.frombuffer(raw[1:1 + 8 * n])
    Vt = np.frombuffer(raw[1 + 8 * n:]).reshape(n, n) if vectors else None
    return raw[0], ev, Vt, raw


def _rbf_normal_equations(K, M, L, s=2.0):
    g, gm = round(K ** (1 / L)), round(M ** (1 / L))
    axes = [np.linspace(-1, 1, g)] * L
    X = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, L)
    mu1 = np.linspace(-1, 1, gm) * gm / (gm - 1)
    MU = np.stack(np.meshgrid(*[mu1] * L, indexing="ij"), -1).reshape(-1, L)
    sig = s * (mu1[1] - mu1[0])
    d2 = ((X[:, None, :] - MU[None, :, :]) ** 2).sum(-1)
    FI = np.hstack([np.exp(-0.5 * d2 / sig ** 2), np.ones((X.shape[0], 1))])
    return FI.T @ FI


def _cases():
    rng = np.random.default_rng(11)
    out = []
    for n in (1, 2, 3, 8, 61, 166, 300):
        B = rng.standard_normal((n, n))
        out.append((f"random{n}", B + B.T))
    Z = rng.standard_normal((40, 200))
    out.append(("psd_rank40", Z.T @ Z))
    out.append(("rbf_1000x126", _rbf_normal_equations(1000, 125, 3)))
    out.append(("diagonal", np.diag(np.arange(12.0) - 5)))
    out.append(("zero", np.zeros((5, 5))))
    out.append(("repeated", np.eye(9) * 3.0))
    return out


@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_eigensolver_against_lapack(driver, tmp_path, name, A):
    A = (A + A.T) / 2
    n = A.shape[0]
    ok, ev, Vt, _ = _eig(driver, tmp_path, A)
    assert ok == 1
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), 1e-300)
    # backward-stable: eigenvalues within a few n eps |A| of LAPACK's
    assert np.abs(np.sort(ev) - ref).max() <= 1e-13 * scale * max(1, n / 50)
    V = Vt.T
    assert np.abs(A @ V - V * ev[None, :]).max() <= 1e-13 * scale * max(1, n / 50)
    assert np.abs(Vt @ V - np.eye(n)).max() <= 1e-13 * max(1, n / 50)
    _, ev2, _, _ = _eig(driver, tmp_path, A, vectors=False)
    assert np.array_equal(ev, ev2)          # values alone take the same QL steps


def test_eigensolver_same_bytes_for_any_thread_count(driver, tmp_path):
    rng = np.random.default_rng(5)
    B = rng.standard_normal((700, 700))
    outs = [_eig(driver, tmp_path, B + B.T, threads=t)[3] for t in (1, 3, 8)]
    assert outs[0] == outs[1] == outs[2]


@pytest.mark.skipif(not (os.path.exists(DAGTM) and h5util.available()),
                    reason="dagtm not built or HDF5 tools missing")
def test_fresh_start_chain_same_bytes_for_any_thread_count(tmp_path):
    rng = np.random.default_rng(3)
    T = (rng.random((30000, 166)) < 0.3).astype(np.float64)
    inp = str(tmp_path / "in.h5")
    h5util.write_matrix(inp, "T", T)
    res = []
    for t in (1, 4):
        out = str(tmp_path / f"init{t}.h5")
        env = dict(os.environ, DG_HOST_THREADS=str(t))
        subprocess.run([DAGTM, "--init-only", "-i", inp, "-o", out, "-K", "1000", "-M", "512",
                        "-L", "3", "-s", "9"], check=True, env=env)
        res.append({k: h5util.read(out, k) for k in ("FI", "W", "Y", "beta", "temp")})
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k]), k
