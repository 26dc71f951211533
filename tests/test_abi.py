"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/dagtm_gpu.h declares; no compute call is made."""
import ctypes as C
import os
import re

import pytest

import dagtm_gpu as dg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dagtm_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_[a-z_0-9]+)\s*\(", src)) - {"dg_iter_cb"})


def test_library_exports_every_declared_symbol():
    lib = dg.load()
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(dg.EXPORTS)


def test_version_string():
    assert dg.load().dg_version().decode().startswith("dagtm_gpu")


def test_create_fails_cleanly_without_device():
    """Without a usable device dg_create returns DG_EHIP instead of aborting."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    lib = dg.load()
    h = C.c_void_p()
    cfg = dg.DgConfig(device=0, precision=dg.FP64, lambda_=0.0)
    assert lib.dg_create(C.byref(cfg), C.byref(h)) == dg.DG_EHIP
    assert not h.value


def test_null_context_is_rejected():
    lib = dg.load()
    assert lib.dg_estep(None, 1, 1.0, None) == dg.DG_EINVAL
    assert lib.dg_mstep(None, None) == dg.DG_EINVAL
    assert lib.dg_next_ctemp(None, 1.0, 0.0, None, None, None) == dg.DG_EINVAL
    assert lib.dg_destroy(None) == dg.DG_OK
    assert lib.dg_last_error(None) == b"null context"
