"""ctypes binding of libdagtm_gpu.so (the C ABI declared in include/dagtm_gpu.h).

The binding is thin on purpose: every call goes straight to the HIP library,
and a missing or unloadable library raises instead of falling back to any
CPU code.  Method names follow the reference's step names
(dagtm/dagtm.cxx:1531-1795): ``estep`` = dist + resp + qual, ``mstep`` =
rowsum + A/B + solve + FI W + beta, ``anneal`` = the temperature loop.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdagtm_gpu.so")

DG_OK, DG_EINVAL, DG_ENOMEM, DG_EHIP, DG_ESTATE, DG_ESOLVE, DG_ESTOP = 0, 1, 2, 3, 4, 5, 6
FP64, MIXED, MIXED_R1 = 0, 1, 2
SCHED_EXP, SCHED_LINEAR, SCHED_AUTO = 0, 1, 2

EXPORTS = ["dg_create", "dg_destroy", "dg_last_error", "dg_version", "dg_upload",
           "dg_set_state", "dg_set_basis", "dg_estep", "dg_mstep", "dg_anneal", "dg_next_ctemp", "dg_map",
           "dg_resp_dense", "dg_download", "dg_get_stats", "dg_set_timing",
           "dg_get_timing", "dg_batch_anneal", "dg_tune", "dg_anneal_ctl", "dg_map_noalloc",
           "dg_device_peaks", "dg_pass2_blocks"]


class DgConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("lambda_", C.c_double)]


class DgEstepOut(C.Structure):
    _fields_ = [("llh", C.c_double), ("nfree", C.c_double), ("nzero", C.c_long)]


class DgAnnealParams(C.Structure):
    _fields_ = [("isDA", C.c_int), ("schedule", C.c_int), ("maxloop", C.c_int),
                ("alpha", C.c_double), ("delta", C.c_double), ("eps", C.c_double),
                ("start_temp", C.c_double), ("nloop0", C.c_int)]


class DgAnnealResult(C.Structure):
    _fields_ = [("nloop", C.c_int), ("ntemps", C.c_int), ("exit_maxloop", C.c_int),
                ("temp", C.c_double), ("llh", C.c_double), ("nfree", C.c_double),
                ("qual", C.c_double), ("beta", C.c_double)]


class DgBatchItem(C.Structure):
    _fields_ = [("T", C.c_void_p), ("N", C.c_long), ("D", C.c_int), ("ldT", C.c_long),
                ("FI", C.c_void_p), ("K", C.c_int), ("M1", C.c_int), ("ldFI", C.c_long),
                ("Y0", C.c_void_p), ("ldY0", C.c_long), ("beta0", C.c_double),
                ("params", DgAnnealParams),
                ("W", C.c_void_p), ("ldW", C.c_long), ("Y", C.c_void_p), ("ldY", C.c_long)]


ITER_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                      C.c_double, C.c_double, C.c_double)

# step callbacks of dg_anneal_ctl (dg_anneal_ops)
ESTEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double),
                       C.POINTER(C.c_double))
MSTEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)
CTEMP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double))
ONITER_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                        C.c_double, C.c_double)
EXPAND_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int))


class DgAnnealOps(C.Structure):
    _fields_ = [("estep", ESTEP_FN), ("mstep", MSTEP_FN), ("next_ctemp", CTEMP_FN),
                ("on_iter", ONITER_FN), ("expand", EXPAND_FN)]


class DgHdaState(C.Structure):
    _fields_ = [("enable", C.c_int), ("K", C.c_int), ("Kmax", C.c_int), ("eps2", C.c_double),
                ("nexpanded", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


class DgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dagtm_gpu error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libdagtm_gpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdagtm_gpu.so not built: run `make -C da-gtm_amd`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.dg_create.argtypes = [C.POINTER(DgConfig), C.POINTER(vp)]
    L.dg_destroy.argtypes = [vp]
    L.dg_last_error.argtypes = [vp]
    L.dg_last_error.restype = C.c_char_p
    L.dg_version.restype = C.c_char_p
    L.dg_upload.argtypes = [vp, _dp, C.c_long, C.c_int, C.c_long, _dp, C.c_int, C.c_int,
                            C.c_long, _dp, C.c_long, C.c_double]
    L.dg_set_state.argtypes = [vp, _dp, C.c_long, C.c_double]
    L.dg_set_basis.argtypes = [vp, _dp, C.c_int, C.c_int, C.c_long, vp, C.c_long]
    L.dg_tune.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
    L.dg_estep.argtypes = [vp, C.c_int, C.c_double, C.POINTER(DgEstepOut)]
    L.dg_mstep.argtypes = [vp, C.POINTER(C.c_double)]
    L.dg_anneal.argtypes = [vp, C.POINTER(DgAnnealParams), C.POINTER(DgAnnealResult),
                            ITER_CB, vp]
    L.dg_next_ctemp.argtypes = [vp, C.c_double, C.c_double, vp, C.POINTER(C.c_double), vp]
    L.dg_map.argtypes = [vp, _dp, C.c_int, C.c_long, _dp, C.c_long]
    L.dg_map_noalloc.argtypes = [vp, _dp, C.c_int, C.c_long, _dp, C.c_long]
    L.dg_device_peaks.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.dg_pass2_blocks.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.dg_resp_dense.argtypes = [vp, _dp, C.c_long]
    L.dg_download.argtypes = [vp, vp, C.c_long, vp, C.c_long, C.POINTER(C.c_double)]
    L.dg_get_stats.argtypes = [vp, vp, vp, C.c_long]
    L.dg_set_timing.argtypes = [vp, C.c_int]
    L.dg_get_timing.argtypes = [vp, C.POINTER(C.c_double)]
    L.dg_batch_anneal.argtypes = [C.POINTER(DgConfig), C.c_int, C.POINTER(DgBatchItem), C.c_int,
                                  C.POINTER(DgAnnealResult), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
    L.dg_anneal_ctl.argtypes = [C.POINTER(DgAnnealParams), C.POINTER(DgHdaState),
                                C.POINTER(DgAnnealOps), vp, C.POINTER(DgAnnealResult)]
    _lib = L
    return L


def anneal_ctl(steps, isDA=True, schedule=SCHED_EXP, maxloop=0, alpha=0.99, delta=0.1,
               eps=1e-6, start_temp=1.0, nloop0=0, hda=None):
    """dg_anneal_ctl over a Python step object with estep(isDA, temp) ->
    (llh, nfree), mstep(), and optionally next_ctemp(temp, eps) -> float,
    on_iter(nloop, lloop, temp, nfree, llh, diff) -> int and expand(nloop,
    temp, K) -> new K.  hda: None or dict(K=, Kmax=, eps2=).  A Python
    exception in a step stops the loop (status 99) and is re-raised here.
    Returns (status, DgAnnealResult, nexpanded, K)."""
    L = load()
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

    ops = DgAnnealOps(ESTEP_FN(guard(_e)), MSTEP_FN(guard(_m)),
                      CTEMP_FN(guard(_c)) if hasattr(steps, "next_ctemp") else CTEMP_FN(),
                      ONITER_FN(guard(_i)) if hasattr(steps, "on_iter") else ONITER_FN(),
                      EXPAND_FN(guard(_x)) if hasattr(steps, "expand") else EXPAND_FN())
    p = DgAnnealParams(isDA=int(isDA), schedule=schedule, maxloop=maxloop, alpha=alpha,
                       delta=delta, eps=eps, start_temp=start_temp, nloop0=nloop0)
    h = DgHdaState(enable=1, K=hda["K"], Kmax=hda["Kmax"], eps2=hda["eps2"]) if hda else None
    r = DgAnnealResult()
    rc = L.dg_anneal_ctl(C.byref(p), C.byref(h) if h is not None else None, C.byref(ops), None,
                         C.byref(r))
    if err:
        raise err[0]
    return rc, r, (h.nexpanded if h is not None else 0), (h.K if h is not None else 0)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One DA-GTM problem resident on one GPU."""

    def __init__(self, precision=FP64, lam=0.0, device=0):
        self.L = load()
        cfg = DgConfig(device=device, precision=precision, lambda_=lam)
        h = C.c_void_p()
        rc = self.L.dg_create(C.byref(cfg), C.byref(h))
        if rc:
            raise DgError(rc, "dg_create failed (no usable HIP device?)")
        self.h = h
        self.precision = precision
        self.N = self.D = self.K = self.M1 = 0

    def _chk(self, rc):
        if rc:
            raise DgError(rc, self.L.dg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.dg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, T, FI, Y, beta):
        T, FI, Y = _f64(T), _f64(FI), _f64(Y)
        N, D = T.shape
        K, M1 = FI.shape
        self._chk(self.L.dg_upload(self.h, T, N, D, D, FI, K, M1, M1, Y, D, float(beta)))
        self.N, self.D, self.K, self.M1 = N, D, K, M1

    def upload_strided(self, Tp, D, FI, Y, beta):
        """upload() from data rows with a row stride: Tp is N x ldT with the
        data in its first D columns (a gsl_matrix's tda layout)."""
        Tp, FI, Y = _f64(Tp), _f64(FI), _f64(Y)
        N, ldT = Tp.shape
        K, M1 = FI.shape
        self._chk(self.L.dg_upload(self.h, Tp, N, D, ldT, FI, K, M1, M1, Y, D, float(beta)))
        self.N, self.D, self.K, self.M1 = N, D, K, M1

    def set_state(self, Y, beta):
        Y = _f64(Y)
        self._chk(self.L.dg_set_state(self.h, Y, Y.shape[1], float(beta)))

    def set_basis(self, FI, W=None):
        """New basis FI (K' x (M+1)); Y = FI W with W given or the last M-step's."""
        FI = _f64(FI)
        K, M1 = FI.shape
        w = None if W is None else _f64(W)
        if w is not None and w.shape != (M1, self.D):
            raise ValueError("W must be (M+1) x D")
        self._chk(self.L.dg_set_basis(self.h, FI, K, M1, M1, None if w is None else w.ctypes.data,
                                      self.D))
        self.K = K

    def tune(self, table_path=None):
        """Time pass-2 chunk counts for this problem (mixed precision), keep the
        fastest and optionally append it to a table file; returns the count."""
        nc = C.c_int()
        self._chk(self.L.dg_tune(self.h, None if table_path is None else
                                 os.fsencode(table_path), C.byref(nc)))
        return nc.value

    def estep(self, isDA=True, temp=1.0):
        o = DgEstepOut()
        self._chk(self.L.dg_estep(self.h, int(isDA), float(temp), C.byref(o)))
        return o.llh, o.nfree, o.nzero

    def mstep(self):
        b = C.c_double()
        self._chk(self.L.dg_mstep(self.h, C.byref(b)))
        return b.value

    def anneal(self, isDA=True, schedule=SCHED_EXP, maxloop=0, alpha=0.99, delta=0.1,
               eps=1e-6, start_temp=1.0):
        p = DgAnnealParams(isDA=int(isDA), schedule=schedule, maxloop=maxloop, alpha=alpha,
                           delta=delta, eps=eps, start_temp=start_temp)
        r = DgAnnealResult()
        rows = []

        def _cb(_u, nloop, lloop, temp, nfree, llh, diff, beta):
            rows.append((nloop, lloop, temp, nfree, llh, diff, beta))

        cb = ITER_CB(_cb)
        self._chk(self.L.dg_anneal(self.h, C.byref(p), C.byref(r), cb, None))
        return r, rows

    def next_ctemp(self, temp, eps, gg=None):
        """Critical temperatures at the current state (needs an estep there).
        gg=None uses the G of the last mstep, as the reference's loop does.
        Returns (max, per-centre array)."""
        g = None if gg is None else _f64(gg)
        if g is not None and g.shape != (self.K,):
            raise ValueError("gg must have K entries")
        ck = np.zeros(self.K); cmax = C.c_double()
        self._chk(self.L.dg_next_ctemp(self.h, float(temp), float(eps),
                                       None if g is None else g.ctypes.data, C.byref(cmax),
                                       ck.ctypes.data))
        return cmax.value, ck

    def map(self, X):
        X = _f64(X)
        Mo = np.zeros((self.N, X.shape[1]))
        self._chk(self.L.dg_map(self.h, X, X.shape[1], X.shape[1], Mo, X.shape[1]))
        return Mo

    def map_noalloc(self, X):
        """dg_map through the library's fixed device scratch (no allocation)."""
        X = _f64(X)
        Mo = np.zeros((self.N, X.shape[1]))
        self._chk(self.L.dg_map_noalloc(self.h, X, X.shape[1], X.shape[1], Mo, X.shape[1]))
        return Mo

    def resp_dense(self):
        R = np.zeros((self.K, self.N))
        self._chk(self.L.dg_resp_dense(self.h, R, self.N))
        return R

    def download(self):
        W = np.zeros((self.M1, self.D)); Y = np.zeros((self.K, self.D)); b = C.c_double()
        self._chk(self.L.dg_download(self.h, W.ctypes.data, self.D, Y.ctypes.data, self.D,
                                     C.byref(b)))
        return W, Y, b.value

    def stats(self):
        G = np.zeros(self.K); RT = np.zeros((self.K, self.D))
        self._chk(self.L.dg_get_stats(self.h, G.ctypes.data, RT.ctypes.data, self.D))
        return G, RT

    def device_peaks(self):
        """Measured (FP16 MFMA TFLOP/s, v_exp_f32 per s, FP64 FMA TFLOP/s)."""
        m, e, f = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.dg_device_peaks(self.h, C.byref(m), C.byref(e), C.byref(f)))
        return m.value, e.value, f.value

    def pass2_blocks(self):
        """(visited, total) (centre group, block) pairs of the last mixed E-pass 2."""
        v, t = C.c_long(0), C.c_long(0)
        self._chk(self.L.dg_pass2_blocks(self.h, C.byref(v), C.byref(t)))
        return v.value, t.value

    def set_timing(self, on=True):
        self._chk(self.L.dg_set_timing(self.h, int(on)))

    def timing(self):
        t = (C.c_double * 4)()
        self._chk(self.L.dg_get_timing(self.h, t))
        return list(t)


def batch_anneal(problems, precision=FP64, lam=0.0, workers=0, device=0, **anneal_kw):
    """Anneal independent instances concurrently (dg_batch_anneal).  workers:
    0 = batched launches replayed as hipGraphs (mixed precisions), -1 = the
    same without graphs, 1..4 = host worker threads.

    problems: list of (T, FI, Y0, beta0) or (T, FI, Y0, beta0, params_dict); the
    keyword arguments are the default annealing parameters (as Engine.anneal).
    Returns (list of (status, DgAnnealResult, W, Y), [upload ms, anneal ms])."""
    L = load()
    n = len(problems)
    items = (DgBatchItem * n)()
    res = (DgAnnealResult * n)()
    status = (C.c_int * n)()
    keep = []
    outs = []
    for i, p in enumerate(problems):
        T, FI, Y0, b0 = (_f64(p[0]), _f64(p[1]), _f64(p[2]), float(p[3]))
        kw = dict(isDA=True, schedule=SCHED_EXP, maxloop=0, alpha=0.99, delta=0.1, eps=1e-6,
                  start_temp=1.0)
        kw.update(anneal_kw)
        if len(p) > 4:
            kw.update(p[4])
        N, D = T.shape
        K, M1 = FI.shape
        W = np.zeros((M1, D))
        Y = np.zeros((K, D))
        keep += [T, FI, Y0]
        outs.append((W, Y))
        items[i] = DgBatchItem(T=T.ctypes.data, N=N, D=D, ldT=D, FI=FI.ctypes.data, K=K, M1=M1,
                               ldFI=M1, Y0=Y0.ctypes.data, ldY0=D, beta0=b0,
                               params=DgAnnealParams(isDA=int(kw["isDA"]),
                                                     schedule=kw["schedule"],
                                                     maxloop=kw["maxloop"], alpha=kw["alpha"],
                                                     delta=kw["delta"], eps=kw["eps"],
                                                     start_temp=kw["start_temp"]),
                               W=W.ctypes.data, ldW=D, Y=Y.ctypes.data, ldY=D)
    cfg = DgConfig(device=device, precision=precision, lambda_=lam)
    ph = (C.c_double * 2)()
    rc = L.dg_batch_anneal(C.byref(cfg), int(workers), items, n, res, status, ph)
    if rc:
        raise DgError(rc, "dg_batch_anneal could not start")
    return [(status[i], res[i], outs[i][0], outs[i][1]) for i in range(n)], list(ph)
