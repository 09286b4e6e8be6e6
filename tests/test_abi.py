"""CPU checks of the C-ABI library: it loads, exports every entry point the
header declares, and rejects bad parameters before touching a device."""
import ctypes
import os
import re

import pytest

from lbmtest import pkg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(REPO, "include", "lbm_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*|lbm_handle \*)\s*(lbm_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    abi = pkg()._abi
    lib = abi.load()
    names = header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(abi.exported_names()) == names


def test_params_struct_matches_header_size():
    abi = pkg()._abi
    # 10 int32 + 14 doubles + wall[8] + corner[2] + 6 contact + 2 oscillation doubles
    assert ctypes.sizeof(abi.Params) == 10 * 4 + (1 + 6 + 1 + 4 + 2 + 8 + 2 + 6 + 2) * 8


@pytest.mark.parametrize("bad", [
    dict(omega=2.5),            # tau = 0.4 < 0.5 (F6:115-118)
    dict(nx=1),
    dict(rheology=7),
    dict(storage=9),
    dict(max_particles=-1),
])
def test_create_rejects_bad_parameters(bad):
    abi = pkg()._abi
    lib = abi.load()
    cfg = dict(nx=10, ny=10, omega=1.0, storage=0, max_particles=1)
    cfg.update(bad)
    p = pkg().make_params(cfg)
    h = ctypes.c_void_p()
    rc = lib.lbm_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == abi.LBM_ERR_PARAM
    assert not h.value
    # with no handle, the reason is kept per thread for lbm_last_error(NULL)
    msg = lib.lbm_last_error(None)
    assert msg and msg != b"null handle", msg


def test_solver_raises_on_bad_tau():
    with pytest.raises(pkg().LBMError):
        pkg().D2Q9(dict(nx=10, ny=10, omega=3.0))
