"""lbm_step_hist_static, the recording fused step that allocates nothing,
through the C ABI: its history rows and end state equal lbm_step's bit for
bit on two calls in a row (the second shorter, so the static array holds the
first call's rows when the second starts), in FP64 and FP16 storage, with
the rows written to host memory and to device memory; a call larger than the
static array fails with LBM_ERR_CAPACITY before it runs anything."""
import ctypes
import importlib

import numpy as np
import pytest

from lbmtest import pkg

pytestmark = pytest.mark.gpu


def two_disk_small(storage):
    cfg, part = importlib.import_module("2d_nnf_lbm_amd.configs").two_disk_bingham(storage)
    cfg = dict(cfg, nx=80, ny=200)
    part = part.copy()
    part[0] = [38.0, 42.0]
    part[1] = [120.0, 165.0]
    return cfg, part


def fresh(cfg, part):
    s = pkg().D2Q9(cfg)
    s.initialize()
    s.set_particles(part)
    return s


def static_call(s, kk0, nsteps, out=None):
    """lbm_step_hist_static through the C ABI; rows into `out` (a device
    address) or a host array."""
    lib = s.lib
    row = s.np * 12 + 2
    if out is None:
        hist = np.full((nsteps, row), np.nan)
        rc = lib.lbm_step_hist_static(s._h, kk0, nsteps, hist.ctypes.data_as(ctypes.c_void_p))
    else:
        hist = None
        rc = lib.lbm_step_hist_static(s._h, kk0, nsteps, ctypes.c_void_p(out))
    return rc, hist


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_static_history_equals_lbm_step(storage):
    print("case: lbm_step_hist_static vs lbm_step,", storage, flush=True)
    cfg, part = two_disk_small(storage)
    a, b = fresh(cfg, part), fresh(cfg, part)
    kk = 0
    for n in (70, 23):
        ha = a.step(kk, n, record=True)
        rc, hb = static_call(b, kk, n)
        assert rc == 0, b.lib.lbm_last_error(b._h).decode()
        assert hb.shape == ha.shape
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.f, b.f)
        assert np.array_equal(a.get_particles(), b.get_particles())
        kk += n
    a.close()
    b.close()


def hip_runtime():
    """The HIP runtime the library itself is linked to (for a device buffer
    owned by the test)."""
    import os
    for path in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        if path.startswith("/") and not os.path.exists(path):
            continue
        rt = ctypes.CDLL(path)
        rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rt.hipFree.argtypes = [ctypes.c_void_p]
        return rt
    pytest.skip("HIP runtime library not found")


def test_static_history_into_device_memory():
    print("case: lbm_step_hist_static into device memory", flush=True)
    rt = hip_runtime()
    cfg, part = two_disk_small("fp16")
    a, b = fresh(cfg, part), fresh(cfg, part)
    kk = 0
    for n in (40, 9):
        ha = a.step(kk, n, record=True)
        nbytes = n * (b.np * 12 + 2) * 8
        dev = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(dev), nbytes) == 0
        try:
            rc, _ = static_call(b, kk, n, out=dev.value)
            assert rc == 0, b.lib.lbm_last_error(b._h).decode()
            back = np.full(ha.shape, np.nan)
            assert rt.hipMemcpy(back.ctypes.data_as(ctypes.c_void_p), dev, nbytes, 2) == 0
        finally:
            rt.hipFree(dev)
        assert np.array_equal(ha, back)
        kk += n
    a.close()
    b.close()


def test_static_history_rejects_oversize_call():
    print("case: lbm_step_hist_static capacity", flush=True)
    cfg, part = two_disk_small("fp16")
    s = fresh(cfg, part)
    f0 = s.f
    words = 1 << 19
    per_step = 2 * 12 * s.np + 66
    n = words // per_step + 1
    rc, _ = static_call(s, 0, n)
    assert rc == pkg()._abi.LBM_ERR_CAPACITY
    assert "static" in s.lib.lbm_last_error(s._h).decode()
    assert np.array_equal(s.f, f0)  # nothing ran
    s.close()
