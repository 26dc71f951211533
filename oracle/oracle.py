"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline).  The product library never imports it.

Every wrapper takes and returns C-contiguous float64 numpy arrays.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


class OrcModel(C.Structure):
    _fields_ = [
        ("N", C.c_long), ("D", C.c_int), ("K", C.c_int), ("M1", C.c_int),
        ("T", C.c_void_p), ("FI", C.c_void_p), ("Y", C.c_void_p), ("W", C.c_void_p),
        ("beta", C.c_double), ("lambda_", C.c_double), ("solver", C.c_int),
        ("Dm", C.c_void_p), ("R", C.c_void_p), ("g", C.c_void_p), ("h", C.c_void_p),
        ("G", C.c_void_p), ("A", C.c_void_p), ("B", C.c_void_p), ("FIR", C.c_void_p),
        ("jitter", C.c_double), ("nzero", C.c_long),
    ]


class OrcAnnealParams(C.Structure):
    _fields_ = [("isDA", C.c_int), ("schedule", C.c_int), ("maxloop", C.c_int),
                ("alpha", C.c_double), ("delta", C.c_double), ("eps", C.c_double),
                ("start_temp", C.c_double), ("nloop0", C.c_int)]


class OrcAnnealResult(C.Structure):
    _fields_ = [("nloop", C.c_int), ("ntemps", C.c_int), ("exit_maxloop", C.c_int),
                ("temp", C.c_double), ("llh", C.c_double), ("nfree", C.c_double),
                ("qual", C.c_double), ("beta", C.c_double)]


ITER_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                      C.c_double, C.c_double, C.c_double)

# step callbacks of orc_anneal_ctl
ESTEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double),
                       C.POINTER(C.c_double))
MSTEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)
CTEMP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double))
ONITER_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                        C.c_double, C.c_double)
EXPAND_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int))


class OrcSteps(C.Structure):
    _fields_ = [("estep", ESTEP_FN), ("mstep", MSTEP_FN), ("next_ctemp", CTEMP_FN),
                ("on_iter", ONITER_FN), ("expand", EXPAND_FN)]


class OrcHda(C.Structure):
    _fields_ = [("enable", C.c_int), ("K", C.c_int), ("Kmax", C.c_int), ("eps2", C.c_double),
                ("nexpanded", C.c_int)]

SOLVE_PINV, SOLVE_CHOL = 0, 1
SCHED_EXP, SCHED_LINEAR, SCHED_AUTO = 0, 1, 2


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_rctg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_stp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_mt_next.argtypes = [C.c_void_p]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_mt_uniform.argtypes = [C.c_void_p]
        L.orc_mt_uniform.restype = C.c_double
        L.orc_yrnd2.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_ulong, _dp]
        L.orc_bi.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_bi.restype = C.c_double
        L.orc_mean_sd.argtypes = [_dp, C.c_long, C.c_int, _dp, _dp]
        L.orc_ctemp.argtypes = [_dp, C.c_long, C.c_int]
        L.orc_ctemp.restype = C.c_double
        L.orc_stp3_rnd.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_ulong, _dp, _dp, _dp,
                                   C.POINTER(C.c_double)]
        L.orc_eig_sym.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_solve2.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_solve.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_chol_solve.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _dp, _dp,
                                     C.POINTER(C.c_double)]
        L.orc_dist.argtypes = [_dp, C.c_long, _dp, C.c_long, C.c_int, _dp]
        L.orc_resp.argtypes = [C.c_int, C.c_long, C.c_double, _dp, C.c_int, C.c_double,
                               _dp, _dp, _dp]
        L.orc_resp.restype = C.c_long
        L.orc_qual.argtypes = [C.c_long, C.c_int, C.c_int, C.c_double, _dp, C.c_int,
                               C.c_double]
        L.orc_qual.restype = C.c_double
        L.orc_beta.argtypes = [_dp, _dp, _dp, C.c_long, C.c_int, C.c_int, _dp]
        L.orc_beta.restype = C.c_double
        L.orc_model_alloc.argtypes = [C.POINTER(OrcModel)]
        L.orc_model_free.argtypes = [C.POINTER(OrcModel)]
        L.orc_estep.argtypes = [C.POINTER(OrcModel), C.c_int, C.c_double,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_estep.restype = C.c_long
        L.orc_mstep.argtypes = [C.POINTER(OrcModel)]
        L.orc_anneal.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcAnnealParams),
                                 C.POINTER(OrcAnnealResult), ITER_CB, C.c_void_p]
        L.orc_anneal_ctl.argtypes = [C.POINTER(OrcAnnealParams), C.POINTER(OrcHda),
                                     C.POINTER(OrcSteps), C.c_void_p,
                                     C.POINTER(OrcAnnealResult)]
        L.orc_map.argtypes = [C.POINTER(OrcModel), _dp, C.c_int, _dp]
        L.orc_next_ctemp.argtypes = [_dp, C.c_long, C.c_int, _dp, C.c_int, _dp, _dp, C.c_double,
                                     C.c_double, C.c_double, _dp]
        L.orc_next_ctemp.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def rctg(xdim, ydim, zdim, L):
    n = xdim * max(ydim, 1) * max(zdim, 1)
    X = np.zeros((n, L))
    lib().orc_rctg(xdim, ydim, zdim, L, X)
    return X


def stp(L, K, M, s):
    X = np.zeros((K, L))
    FI = np.zeros((K, M + 1))
    rc = lib().orc_stp(L, K, M, float(s), X, FI)
    if rc:
        raise ValueError("orc_stp failed: %d" % rc)
    return X, FI


class MT19937:
    """MT19937 state as the reference's GSL default generator seeds it."""

    def __init__(self, seed):
        self._buf = C.create_string_buffer(624 * 4 + 8)
        lib().orc_mt_seed(self._buf, seed)

    def next_u32(self):
        return lib().orc_mt_next(self._buf)

    def uniform(self):
        return lib().orc_mt_uniform(self._buf)


def yrnd2(mu, sd, K, seed):
    mu, sd = _f64(mu), _f64(sd)
    Y = np.zeros((K, mu.size))
    lib().orc_yrnd2(mu, sd, K, mu.size, seed, Y)
    return Y


def bi(Y):
    Y = _f64(Y)
    return lib().orc_bi(Y, Y.shape[0], Y.shape[1])


def mean_sd(T):
    T = _f64(T)
    mu = np.zeros(T.shape[1]); sd = np.zeros(T.shape[1])
    lib().orc_mean_sd(T, T.shape[0], T.shape[1], mu, sd)
    return mu, sd


def ctemp(T):
    T = _f64(T)
    return lib().orc_ctemp(T, T.shape[0], T.shape[1])


def stp3_rnd(T, L, K, M, s, seed):
    """Fresh-start chain (dagtm.cxx:1104-1143): returns X, FI, W, Y, beta0."""
    mu, sd = mean_sd(T)
    D = mu.size
    X = np.zeros((K, L)); FI = np.zeros((K, M + 1)); W = np.zeros((M + 1, D))
    b = C.c_double()
    rc = lib().orc_stp3_rnd(mu, sd, D, L, K, M, float(s), seed, X, FI, W, C.byref(b))
    if rc:
        raise ValueError("orc_stp3_rnd failed: %d" % rc)
    return X, FI, W, FI @ W, b.value


def eig_sym(A):
    A = _f64(A).copy()
    n = A.shape[0]
    ev = np.zeros(n); V = np.zeros((n, n))
    lib().orc_eig_sym(n, A, ev, V)
    return ev, V


def solve2(A, B):
    A, B = _f64(A), _f64(B)
    X = np.zeros_like(B)
    lib().orc_solve2(A.shape[0], A, B.shape[1], B, X)
    return X


def solve_ls(A, B):
    A, B = _f64(A), _f64(B)
    X = np.zeros((A.shape[1], B.shape[1]))
    lib().orc_solve(A.shape[0], A.shape[1], A, B.shape[1], B, X)
    return X


def chol_solve(A, B, ridge=0.0):
    A, B = _f64(A), _f64(B)
    X = np.zeros_like(B)
    j = C.c_double()
    rc = lib().orc_chol_solve(A.shape[0], A, float(ridge), B.shape[1], B, X, C.byref(j))
    if rc:
        raise np.linalg.LinAlgError("Cholesky failed")
    return X, j.value


def dist(Y, T):
    Y, T = _f64(Y), _f64(T)
    Dm = np.zeros((Y.shape[0], T.shape[0]))
    lib().orc_dist(Y, Y.shape[0], T, T.shape[0], Y.shape[1], Dm)
    return Dm


def resp(Dm, beta, isDA, temp):
    Dm = _f64(Dm)
    K, N = Dm.shape
    R = np.zeros_like(Dm); g = np.zeros(N); h = np.zeros(N)
    nz = lib().orc_resp(K, N, float(beta), Dm, int(isDA), float(temp), R, g, h)
    return R, g, h, nz


def qual(v, N, K, D, beta, isDA, temp):
    v = _f64(v).copy()
    return lib().orc_qual(N, K, D, float(beta), v, int(isDA), float(temp))


class Model:
    """Materialised FP64 model in the reference's step order."""

    def __init__(self, T, FI, Y, beta, lam=0.0, solver=SOLVE_PINV):
        self.T = _f64(T); self.FI = _f64(FI); self.Y = _f64(Y).copy()
        N, D = self.T.shape; K, M1 = self.FI.shape
        self.W = np.zeros((M1, D))
        self.m = OrcModel(N=N, D=D, K=K, M1=M1,
                          T=self.T.ctypes.data, FI=self.FI.ctypes.data,
                          Y=self.Y.ctypes.data, W=self.W.ctypes.data,
                          beta=float(beta), lambda_=float(lam), solver=int(solver))
        if lib().orc_model_alloc(C.byref(self.m)):
            raise MemoryError("orc_model_alloc")

    def __del__(self):
        try:
            lib().orc_model_free(C.byref(self.m))
        except Exception:
            pass

    @property
    def beta(self):
        return self.m.beta

    def _arr(self, name, shape):
        ptr = getattr(self.m, name)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=shape).copy()

    @property
    def R(self):
        return self._arr("R", (self.m.K, self.m.N))

    @property
    def G(self):
        return self._arr("G", (self.m.K,))

    @property
    def Dm(self):
        return self._arr("Dm", (self.m.K, self.m.N))

    def estep(self, isDA, temp):
        llh = C.c_double(); nf = C.c_double()
        nz = lib().orc_estep(C.byref(self.m), int(isDA), float(temp), C.byref(llh), C.byref(nf))
        return llh.value, nf.value, nz

    def mstep(self):
        return lib().orc_mstep(C.byref(self.m))

    def anneal(self, isDA=True, schedule=SCHED_EXP, maxloop=0, alpha=0.99,
               delta=0.1, eps=1e-6, start_temp=1.0):
        p = OrcAnnealParams(isDA=int(isDA), schedule=schedule, maxloop=maxloop,
                            alpha=alpha, delta=delta, eps=eps, start_temp=start_temp)
        r = OrcAnnealResult()
        rows = []

        def _cb(_u, nloop, lloop, temp, nfree, llh, diff, beta):
            rows.append((nloop, lloop, temp, nfree, llh, diff, beta))

        cb = ITER_CB(_cb)
        lib().orc_anneal(C.byref(self.m), C.byref(p), C.byref(r), cb, None)
        return r, rows

    def next_ctemp(self, temp, eps, gg=None):
        """Next critical temperature below temp from the last E-step's R and
        (by default) the last M-step's G, as the reference's loop passes them.
        Returns (max, per-centre array)."""
        g = self.G if gg is None else _f64(gg)
        ck = np.zeros(self.m.K)
        c = lib().orc_next_ctemp(self.T, self.m.N, self.m.D, self.Y, self.m.K, self.R, g,
                                 float(temp), float(self.m.beta), float(eps), ck)
        return c, ck

    def map(self, X):
        X = _f64(X)
        Mo = np.zeros((self.m.N, X.shape[1]))
        lib().orc_map(C.byref(self.m), X, X.shape[1], Mo)
        return Mo


def anneal_ctl(steps, isDA=True, schedule=SCHED_EXP, maxloop=0, alpha=0.99, delta=0.1,
               eps=1e-6, start_temp=1.0, nloop0=0, hda=None):
    """orc_anneal_ctl over a Python step object with estep(isDA, temp) ->
    (llh, nfree), mstep(), and optionally next_ctemp(temp, eps) -> float,
    on_iter(nloop, lloop, temp, nfree, llh, diff) -> int and expand(nloop,
    temp, K) -> new K.  hda: None or dict(K=, Kmax=, eps2=).  A Python
    exception in a step stops the loop (status 99) and is re-raised here.
    Returns (status, OrcAnnealResult, nexpanded, K)."""
    err = []

    def guard(f):
        def g(*a):
            try:
                return f(*a)
            except BaseException as e:       # noqa: BLE001 - re-raised below
                err.append(e)
                return 99
        return g

    def _e(_u, isDA_, temp, pl, pf):
        pl[0], pf[0] = steps.estep(isDA_, temp)
        return 0

    def _m(_u):
        steps.mstep()
        return 0

    def _c(_u, temp, e, pc):
        pc[0] = steps.next_ctemp(temp, e)
        return 0

    def _i(_u, nloop, lloop, temp, nfree, llh, diff):
        return int(steps.on_iter(nloop, lloop, temp, nfree, llh, diff) or 0)

    def _x(_u, nloop, temp, pK):
        pK[0] = int(steps.expand(nloop, temp, pK[0]))
        return 0

    st = OrcSteps(ESTEP_FN(guard(_e)), MSTEP_FN(guard(_m)),
                  CTEMP_FN(guard(_c)) if hasattr(steps, "next_ctemp") else CTEMP_FN(),
                  ONITER_FN(guard(_i)) if hasattr(steps, "on_iter") else ONITER_FN(),
                  EXPAND_FN(guard(_x)) if hasattr(steps, "expand") else EXPAND_FN())
    p = OrcAnnealParams(isDA=int(isDA), schedule=schedule, maxloop=maxloop, alpha=alpha,
                        delta=delta, eps=eps, start_temp=start_temp, nloop0=nloop0)
    h = OrcHda(enable=1, K=hda["K"], Kmax=hda["Kmax"], eps2=hda["eps2"]) if hda else None
    r = OrcAnnealResult()
    rc = lib().orc_anneal_ctl(C.byref(p), C.byref(h) if h is not None else None, C.byref(st),
                              None, C.byref(r))
    if err:
        raise err[0]
    return rc, r, (h.nexpanded if h is not None else 0), (h.K if h is not None else 0)


class _HdaSteps:
    """Steps of anneal_hda: the materialised model, replaced at an expansion."""

    def __init__(self, T, FI, Y, beta, W, L, M, s, lam, solver, X, on_expand):
        self.T, self.FI, self.L, self.M, self.s = T, FI, L, M, s
        self.lam, self.solver, self.X, self.on_expand = lam, solver, X, on_expand
        self.m = Model(T, FI, Y, beta, lam=lam, solver=solver)
        self.Wlast = _f64(W).copy()
        self.K = FI.shape[0]
        self.ctemp_k = None
        self.rows = []

    def estep(self, isDA, temp):
        llh, nfree, _ = self.m.estep(isDA, temp)
        self._beta_used = self.m.beta
        return llh, nfree

    def mstep(self):
        self.m.mstep()
        self.Wlast = self.m.W.copy()

    def next_ctemp(self, temp, eps):
        c, self.ctemp_k = self.m.next_ctemp(temp, eps)
        return c

    def on_iter(self, nloop, lloop, temp, nfree, llh, diff):
        self.rows.append((nloop, lloop, self.K, temp, nfree, llh, diff, self._beta_used))
        return 0

    def expand(self, nloop, temp, K):
        if self.on_expand is not None:
            self.on_expand(nloop - 1, self.m.map(self.X), self.X, self.m.Y.copy(), self.FI,
                           self.Wlast, self.ctemp_k, self.m.beta)
        g = int(round(K ** (1.0 / self.L))) * 2
        self.K = g ** self.L
        self.X, self.FI = stp(self.L, self.K, self.M, self.s)
        self.m = Model(self.T, self.FI, self.FI @ self.Wlast, self.m.beta, lam=self.lam,
                       solver=self.solver)
        return self.K


def anneal_hda(T, FI, Y, beta, W, L, M, s, Kmax=32768, eps2=1e-3, alpha=0.99, eps=1e-6,
               start_temp=1.0, maxloop=0, lam=0.0, solver=SOLVE_PINV, X=None, on_expand=None):
    """Hierarchical DA: the EXP loop of dagtm.cxx:1508-1795, 1909-2164 with -H,
    run by orc_anneal_ctl over the materialised model.

    After each temperature (not the last) and while K < Kmax: if an expansion
    was armed, it fires when the llh gained over the temperature
    (qual_diff_at_temp, 1912) is at least eps2 and is cancelled otherwise; an
    unarmed state computes the next critical temperature and arms when it is
    >= temp * alpha (1944-1970).  Firing doubles the grid side,
    K <- (2 round(K^(1/L)))^L, builds the new grid and basis with the same M and
    width s, and sets Y = FI W with the last M-step's W (W0 before any M-step,
    2101-2160).  Then temp *= alpha.  on_expand(idx, M, X, Y, FI, W, ctemp,
    beta) receives the state the reference logs at an expansion (1983-2095),
    idx = nloop - 1.  Returns (summary dict, rows), each row (nloop, lloop, K,
    temp, nfree, llh, diff, beta)."""
    T = _f64(T)
    K = FI.shape[0]
    if X is None:
        X, _ = stp(L, K, M, s)
    st = _HdaSteps(T, _f64(FI), Y, beta, W, L, M, s, lam, solver, X, on_expand)
    rc, r, nexp, Kout = anneal_ctl(st, isDA=True, schedule=SCHED_EXP, maxloop=maxloop,
                                   alpha=alpha, eps=eps, start_temp=start_temp,
                                   hda=dict(K=K, Kmax=Kmax, eps2=eps2))
    if rc:
        raise RuntimeError("orc_anneal_ctl failed: %d" % rc)
    return dict(nloop=r.nloop, ntemps=r.ntemps, nexp=nexp, K=Kout, temp=r.temp, llh=r.llh,
                nfree=r.nfree, beta=st.m.beta, Y=st.m.Y.copy(), X=st.X, FI=st.FI,
                W=st.Wlast), st.rows
