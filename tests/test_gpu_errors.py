"""Error behaviour of the C ABI on a GPU: every entry point rejects bad
arguments and out-of-order calls with its status code and a message, never
aborts, and leaves the context usable (the reference's convention of returning
a status, comm.h:122-131)."""
import ctypes as C

import numpy as np
import pytest

import dagtm_gpu as dg
import datagen

pytestmark = pytest.mark.gpu


def _ctx(precision=dg.FP64, lam=0.0):
    L = dg.load()
    h = C.c_void_p()
    cfg = dg.DgConfig(device=0, precision=precision, lambda_=lam)
    assert L.dg_create(C.byref(cfg), C.byref(h)) == dg.DG_OK
    return L, h


def _msg(L, h):
    return L.dg_last_error(h).decode()


def _problem(N=300, D=5, K=16, M=4, L=2):
    T = datagen.gmm(N, D, 1)
    X, FI = datagen.latent_basis(L, K, M, 1.0)
    W0, Y0, b0 = datagen.pca_init(T, FI, X)
    return T, FI, Y0, b0


def test_create_rejects_bad_config():
    L = dg.load()
    h = C.c_void_p()
    for prec, lam in ((7, 0.0), (dg.FP64, -1.0)):
        cfg = dg.DgConfig(device=0, precision=prec, lambda_=lam)
        assert L.dg_create(C.byref(cfg), C.byref(h)) == dg.DG_EINVAL and not h.value
    cfg = dg.DgConfig(device=999, precision=dg.FP64, lambda_=0.0)
    assert L.dg_create(C.byref(cfg), C.byref(h)) == dg.DG_EHIP and not h.value


def test_calls_out_of_order():
    L, h = _ctx()
    try:
        o = dg.DgEstepOut()
        assert L.dg_estep(h, 1, 1.0, C.byref(o)) == dg.DG_ESTATE and "upload" in _msg(L, h)
        assert L.dg_mstep(h, None) == dg.DG_ESTATE
        T, FI, Y0, b0 = _problem()
        N, D = T.shape; K, M1 = FI.shape
        assert L.dg_upload(h, T, N, D, D, FI, K, M1, M1, Y0, D, b0) == dg.DG_OK
        assert L.dg_mstep(h, None) == dg.DG_ESTATE and "estep" in _msg(L, h)
        assert L.dg_get_stats(h, None, None, 0) == dg.DG_ESTATE
        assert L.dg_resp_dense(h, np.zeros((K, N)), N) == dg.DG_ESTATE
        assert L.dg_estep(h, 1, 2.0, C.byref(o)) == dg.DG_OK
        assert L.dg_mstep(h, None) == dg.DG_OK
        assert L.dg_mstep(h, None) == dg.DG_ESTATE          # a second M-step needs a new E-step
    finally:
        L.dg_destroy(h)


def test_upload_rejects_bad_shapes_and_recovers():
    L, h = _ctx()
    T, FI, Y0, b0 = _problem()
    N, D = T.shape; K, M1 = FI.shape
    try:
        bad = [
            (T, 0, D, D, FI, K, M1, M1, Y0, D, b0),          # empty data
            (T, N, 0, D, FI, K, M1, M1, Y0, D, b0),          # D = 0
            (T, N, D, D - 1, FI, K, M1, M1, Y0, D, b0),      # ldT < D
            (T, N, D, D, FI, 0, M1, M1, Y0, D, b0),          # K = 0
            (T, N, D, D, FI, K, M1, M1 - 1, Y0, D, b0),      # ldFI < M1
            (T, N, D, D, FI, K, M1, M1, Y0, D, 0.0),         # beta <= 0
            (T, N, D, D, FI, K, M1, M1, Y0, D, float("nan")),
        ]
        for args in bad:
            assert L.dg_upload(h, *args) == dg.DG_EINVAL, args[1:4]
            assert _msg(L, h)
        Tw = np.zeros((N, 193)); Yw = np.zeros((K, 193))
        assert L.dg_upload(h, Tw, N, 193, 193, FI, K, M1, M1, Yw, 193, b0) == dg.DG_EINVAL
        assert "192" in _msg(L, h)
        FIw = np.zeros((K, 2049))
        assert L.dg_upload(h, T, N, D, D, FIw, K, 2049, 2049, Y0, D, b0) == dg.DG_EINVAL
        # the context still works after the rejections
        assert L.dg_upload(h, T, N, D, D, FI, K, M1, M1, Y0, D, b0) == dg.DG_OK
        o = dg.DgEstepOut()
        assert L.dg_estep(h, 1, 0.0, C.byref(o)) == dg.DG_EINVAL     # temperature must be > 0
        assert L.dg_estep(h, 1, 2.0, C.byref(o)) == dg.DG_OK and np.isfinite(o.llh)
    finally:
        L.dg_destroy(h)


def test_other_entry_points_validate_arguments():
    with dg.Engine(precision=dg.FP64) as e:
        T, FI, Y0, b0 = _problem()
        e.upload(T, FI, Y0, b0)
        e.estep(True, 2.0)
        N, D = T.shape; K = FI.shape[0]
        L, h = e.L, e.h
        X4 = np.zeros((K, 4))
        assert L.dg_map(h, X4, 4, 4, np.zeros((N, 4)), 4) == dg.DG_EINVAL   # L <= 3
        assert L.dg_download(h, None, 0, np.zeros((K, D)).ctypes.data, D - 1, None) == dg.DG_EINVAL
        assert L.dg_set_state(h, Y0, D, -1.0) == dg.DG_EINVAL
        p = dg.DgAnnealParams(isDA=1, schedule=9, maxloop=5, alpha=0.9, delta=0.1, eps=1e-6,
                              start_temp=2.0)
        r = dg.DgAnnealResult()
        assert L.dg_anneal(h, C.byref(p), C.byref(r), dg.ITER_CB(0), None) == dg.DG_EINVAL
        assert "schedule" in L.dg_last_error(h).decode()
        e.estep(True, 2.0)                            # still usable


def test_mixed_rejects_wide_float_data():
    T = datagen.gmm(500, 140, 3)                      # not exact in FP16, D > 128
    X, FI = datagen.latent_basis(2, 16, 4, 1.0)
    Y0 = np.tile(T.mean(0), (16, 1))
    with dg.Engine(precision=dg.MIXED) as e:
        with pytest.raises(dg.DgError) as ex:
            e.upload(T, FI, Y0, 1.0)
        assert ex.value.code == dg.DG_EINVAL and "FP16" in str(ex.value)
    with dg.Engine(precision=dg.FP64) as e:          # the FP64 path takes it
        e.upload(T, FI, Y0, 1.0)
        assert np.isfinite(e.estep(True, 2.0)[0])
