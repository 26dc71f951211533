"""The dagtm driver end to end on the GPU: progress rows against the oracle,
checkpoint/restart, the SIGCHLD forced checkpoint, and the output file."""
import os
import signal
import subprocess
import sys
import time

import numpy as np
import pytest

import h5util
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DAGTM = os.path.join(ROOT, "da-gtm_amd", "dagtm")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(DAGTM) and h5util.available()),
                                 reason="dagtm not built or HDF5 tools missing")]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(stdout):
    out = []
    for line in stdout.splitlines():
        if line.startswith(":::"):
            f = line.split()
            out.append((int(f[1]), int(f[2]), float(f[4]), float(f[5]), float(f[6]), float(f[7])))
    return out


@pytest.fixture()
def oil_input(tmp_path, oil_state):
    p = str(tmp_path / "in.h5")
    h5util.write_matrix(p, "T", oil_state["T"])
    return p


def test_cli_run_matches_oracle(tmp_path, oil_state, oil_input):
    _say("case: dagtm on the Oil data, K=125 M=8, 60 iterations, fp64")
    init = str(tmp_path / "init.h5")
    common = ["-i", oil_input, "-K", "125", "-M", "8", "-L", "3", "-s", "23", "-a", "0.9"]
    subprocess.run([DAGTM, "--init-only", "-o", init] + common, check=True)
    out = str(tmp_path / "out.h5")
    r = subprocess.run([DAGTM, "-n", "-j", "60", "-o", out, "-p"] + common, check=True,
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    rows = _rows(r.stdout)
    assert len(rows) == 60
    T = oil_state["T"]
    FI = h5util.read(init, "FI").reshape(125, 9)
    Y0 = h5util.read(init, "Y").reshape(125, 12)
    b0 = h5util.read(init, "beta")[0]
    t0 = h5util.read(init, "temp")[0]
    m = O.Model(T, FI, Y0, b0, solver=O.SOLVE_CHOL)
    res, orows = m.anneal(isDA=True, alpha=0.9, eps=1e-6, start_temp=t0, maxloop=60)
    # Past the first phase transition (temp ~1.46, iteration ~47) the DA dynamics
    # amplify any difference: the oracle's own trajectory moves by ~1e-5 in llh when
    # its initial Y is perturbed by 1e-13. The llh tolerance of each row is therefore
    # the printed precision plus 50x that measured sensitivity.
    mp = O.Model(T, FI, Y0 * (1.0 + 1e-13), b0, solver=O.SOLVE_CHOL)
    _, prows = mp.anneal(isDA=True, alpha=0.9, eps=1e-6, start_temp=t0, maxloop=60)
    assert len(orows) == len(rows) == len(prows)
    for g, o, p in zip(rows, orows, prows):
        assert g[0] == o[0] and g[1] == o[1]
        assert g[2] == pytest.approx(o[2], rel=1e-7)            # temp (printed %15.7f)
        tol = 2e-7 + 1e-7 * abs(o[4]) + 50.0 * abs(p[4] - o[4])  # llh (printed %15.7f)
        assert abs(g[4] - o[4]) <= tol, (g[0], g[4], o[4], tol)
    ysens = np.abs(mp.Y - m.Y).max()
    # output layout (dagtm.cxx:2342-2552)
    for name, shp in (("T", (1000, 12)), ("M", (1000, 3)), ("X", (125, 3)), ("FI", (125, 9)),
                      ("Y", (125, 12)), ("W", (9, 12)), ("beta", (1,)), ("temp", (1,)),
                      ("iter", (1,)), ("seed", (1,)), ("qual", (1,)), ("Temp", (60,)),
                      ("LLH", (60,)), ("NFree", (60,))):
        assert h5util.shape(out, name) == shp, name
    assert h5util.read(out, "iter", "<i4")[0] == 60
    Yo = h5util.read(out, "Y").reshape(125, 12)
    assert np.abs(Yo - m.Y).max() <= 1e-8 * np.abs(m.Y).max() + 50.0 * ysens
    assert abs(h5util.read(out, "beta")[0] - m.beta) <= 1e-8 * m.beta + 50.0 * abs(mp.beta - m.beta)
    Mo = h5util.read(out, "M").reshape(1000, 3)
    assert np.all(np.abs(Mo) <= 1.0 + 1e-12)                  # convex combinations of the grid


def test_cli_auto_schedule_matches_oracle(tmp_path, oil_state, oil_input):
    _say("case: dagtm -c auto on the Oil data, K=125 M=8, 40 iterations, fp64")
    init = str(tmp_path / "init.h5")
    common = ["-i", oil_input, "-K", "125", "-M", "8", "-L", "3", "-s", "23", "-a", "0.9"]
    subprocess.run([DAGTM, "--init-only", "-o", init] + common, check=True)
    r = subprocess.run([DAGTM, "-n", "-j", "40", "-c", "auto"] + common, check=True,
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    rows = _rows(r.stdout)
    T = oil_state["T"]
    FI = h5util.read(init, "FI").reshape(125, 9)
    Y0 = h5util.read(init, "Y").reshape(125, 12)
    b0 = h5util.read(init, "beta")[0]
    t0 = h5util.read(init, "temp")[0]
    kw = dict(isDA=True, schedule=O.SCHED_AUTO, alpha=0.9, eps=1e-6, start_temp=t0, maxloop=40)
    _, orows = O.Model(T, FI, Y0, b0, solver=O.SOLVE_CHOL).anneal(**kw)
    _, prows = O.Model(T, FI, Y0 * (1.0 + 1e-13), b0, solver=O.SOLVE_CHOL).anneal(**kw)
    assert len(orows) == len(rows) == len(prows) == 40
    assert len(set(o[2] for o in orows)) > 2               # AUTO moved the temperature
    for g, o, p in zip(rows, orows, prows):
        assert g[0] == o[0] and g[1] == o[1]
        assert abs(g[2] - o[2]) <= 1e-7 + 1e-7 * o[2] + 50.0 * abs(p[2] - o[2]), (g[0], g[2], o[2])
        tol = 2e-7 + 1e-7 * abs(o[4]) + 50.0 * abs(p[4] - o[4])
        assert abs(g[4] - o[4]) <= tol, (g[0], g[4], o[4], tol)


def test_cli_hda_matches_oracle(tmp_path, oil_state, oil_input):
    _say("case: dagtm -H on the Oil data, K=25 L=2 M=9, 150 iterations, fp64")
    init = str(tmp_path / "init.h5")
    common = ["-i", oil_input, "-K", "25", "-M", "9", "-L", "2", "-s", "5", "-a", "0.9"]
    subprocess.run([DAGTM, "--init-only", "-o", init] + common, check=True)
    out = str(tmp_path / "out.h5")
    r = subprocess.run([DAGTM, "-n", "-j", "150", "-H", "-k", "300", "-E", "1e-3", "-p", "-o", out]
                       + common, check=True, capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    rows = _rows(r.stdout)
    T = oil_state["T"]
    rd = lambda n, sh: h5util.read(init, n).reshape(sh)
    X0, FI, Y0, W0 = rd("X", (25, 2)), rd("FI", (25, 10)), rd("Y", (25, 12)), rd("W", (10, 12))
    b0, t0 = h5util.read(init, "beta")[0], h5util.read(init, "temp")[0]
    kw = dict(Kmax=300, eps2=1e-3, alpha=0.9, eps=1e-6, start_temp=t0, maxloop=150,
              solver=O.SOLVE_CHOL, X=X0)
    snaps = []
    res, orows = O.anneal_hda(T, FI, Y0, b0, W0, 2, 9, 2.0, on_expand=lambda *a: snaps.append(a),
                              **kw)
    # Past the phase transition near iteration 49 the run is chaotic: oracle runs
    # from Y0 (1 + 1e-15 .. 1e-13) differ from it by 1e-7 .. 1e-4 in llh, about
    # independently of the perturbation's size.  The envelope of three such runs
    # over a +-5 row window, times 10, bounds what the GPU may differ by.
    pert = []
    for dp in (1e-15, 1e-14, 1e-13):
        ps = []
        _, prows = O.anneal_hda(T, FI, Y0 * (1.0 + dp), b0, W0, 2, 9, 2.0,
                                on_expand=lambda *a: ps.append(a), **kw)
        assert [p[2] for p in prows] == [o[2] for o in orows]      # same expansion point
        pert.append((prows, ps[0]))
    ks = [int(line.split()[3]) for line in r.stdout.splitlines() if line.startswith(":::")]
    assert res["nexp"] == 1 and len(rows) == len(ks) == len(orows) == 150
    dev = np.array([[abs(p[i][5] - orows[i][5]) for i in range(150)] for p, _ in pert]).max(0)
    env = np.array([dev[max(0, i - 5):i + 6].max() for i in range(150)])
    # g: (nloop, lloop, temp, nfree, llh, diff); o: (nloop, lloop, K, temp, nfree, llh, ...)
    for i, (g, k, o) in enumerate(zip(rows, ks, orows)):
        assert g[0] == o[0] and g[1] == o[1] and k == o[2]                 # nloop, lloop, K
        assert abs(g[2] - o[3]) <= 1e-7 + 1e-7 * o[3], (g[0], g[2], o[3])  # temp
        tol = 2e-7 + 1e-7 * abs(o[5]) + 10.0 * env[i]                     # llh (printed %15.7f)
        assert abs(g[4] - o[5]) <= tol, (g[0], g[4], o[5], tol)
    assert "new K : 100" in r.stdout
    # the expansion snapshot (dagtm.cxx:1983-2095) and the final state at K = 100
    idx, Mo, Xs, Ys, FIs, Ws, ck, beta = snaps[0]
    assert all(ps[0] == idx for _, ps in pert)
    sfx = ".%d" % idx
    assert np.abs(h5util.read(out, "X" + sfx).reshape(25, 2) - Xs).max() < 1e-15
    assert np.abs(h5util.read(out, "FI" + sfx).reshape(25, 10) - FIs).max() < 1e-14
    for name, shp, pos in (("M", (1000, 2), 1), ("Y", (25, 12), 3), ("W", (10, 12), 5),
                           ("ctemp", (25,), 6), ("beta", (1,), 7)):
        ref = np.atleast_1d(np.asarray(snaps[0][pos], dtype=float))
        spread = max(np.abs(np.atleast_1d(np.asarray(ps[pos], dtype=float)) - ref).max()
                     for _, ps in pert)
        tol = 1e-8 * np.abs(ref).max() + 10.0 * spread
        assert np.abs(h5util.read(out, name + sfx).reshape(shp) - ref).max() <= tol, name
    for name, shp in (("X", (100, 2)), ("FI", (100, 10)), ("Y", (100, 12)), ("M", (1000, 2))):
        assert h5util.shape(out, name) == shp, name


def test_checkpoint_restart_reproduces_trajectory(tmp_path, oil_input):
    _say("case: dagtm checkpoint every 10, restart from iteration 20")
    common = ["-i", oil_input, "-K", "125", "-M", "8", "-L", "3", "-s", "5", "-e", "0",
              "-S", "3.0", "-w", str(tmp_path / "ckp.")]
    full = subprocess.run([DAGTM, "-P", "10", "-j", "30"] + common, check=True,
                          capture_output=True, text=True, timeout=120)
    assert os.path.exists(tmp_path / "ckp.20.h5")
    rest = subprocess.run([DAGTM, "-n", "-j", "10", "-r", str(tmp_path / "ckp.20.h5")] + common,
                          check=True, capture_output=True, text=True, timeout=120)
    a = {r[0]: r for r in _rows(full.stdout)}
    b = _rows(rest.stdout)
    assert [r[0] for r in b] == list(range(21, 31))
    # the checkpoint holds the state of iteration 20's E-step (written before
    # its M-step, dagtm.cxx:1609-1687), so the resumed run replays that E-step
    # as iteration 21 -- the reference's restart semantics
    for r in b:
        assert r[4] == a[r[0] - 1][4]      # llh, bitwise-identical state on restart


def test_sigchld_forces_checkpoint_and_exit(tmp_path, oil_input):
    _say("case: dagtm SIGCHLD forced checkpoint")
    cmd = [DAGTM, "-i", oil_input, "-K", "125", "-M", "8", "-L", "3", "-s", "5", "-e", "0",
           "-S", "3.0", "-j", "2000000", "-P", "0", "-w", str(tmp_path / "ckp.")]
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    seen = 0
    t0 = time.time()
    for line in p.stdout:
        if line.startswith(":::"):
            seen += 1
            if seen == 5:
                p.send_signal(signal.SIGCHLD)
                break
        if time.time() - t0 > 60:
            break
    try:
        p.communicate(timeout=60)
    except subprocess.TimeoutExpired:
        p.kill()
        p.communicate()
        raise
    rc = p.returncode
    assert seen == 5 and rc == 0
    ck = [f for f in os.listdir(tmp_path) if f.startswith("ckp.") and f.endswith(".h5")]
    assert len(ck) == 1
    n = int(ck[0].split(".")[1])
    assert n >= 5
    assert h5util.read(str(tmp_path / ck[0]), "iter", "<i4")[0] == n


def test_interpolation_mode_matches_oracle(tmp_path, oil_state, oil_input):
    """-I: responsibilities of the -i data under the -r model, M = R^T X appended
    to the -o file (dagtm.cxx:1247-1295)."""
    _say("case: dagtm -I interpolation")
    common = ["-i", oil_input, "-K", "125", "-M", "8", "-L", "3", "-s", "5", "-S", "2.5",
              "-w", str(tmp_path / "ckp.")]
    subprocess.run([DAGTM, "-P", "15", "-j", "15"] + common, check=True, capture_output=True,
                   text=True, timeout=120)
    ck = str(tmp_path / "ckp.15.h5")
    out = str(tmp_path / "interp.h5")
    subprocess.run([DAGTM, "-I", "-r", ck, "-o", out] + common, check=True, capture_output=True,
                   text=True, timeout=120)
    X = h5util.read(ck, "X").reshape(125, 3)
    FI = h5util.read(ck, "FI").reshape(125, 9)
    Y = h5util.read(ck, "Y").reshape(125, 12)
    beta = h5util.read(ck, "beta")[0]
    temp = h5util.read(ck, "temp")[0]
    m = O.Model(oil_state["T"], FI, Y, beta)
    m.estep(1, temp)
    Mref = m.map(X)
    Mo = h5util.read(out, "M").reshape(1000, 3)
    assert np.abs(Mo - Mref).max() < 1e-12
