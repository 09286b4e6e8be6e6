"""CPU checks of the C-ABI library: it loads, exports every entry point the
header declares, and rejects bad parameters before touching a device."""
import ctypes
import os
import re
import shutil
import subprocess

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


def test_params_struct_matches_header_layout(tmp_path):
    """sizeof(lbm_params) and every field offset, as the C compiler lays out
    the header's struct, equal the ctypes mirror's."""
    abi = pkg()._abi
    fields = [n for n, _ in abi.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "lbm_amd.h"\n'
                   "int main(void) {\n"
                   '  printf("%zu\\n", sizeof(lbm_params));\n'
                   + "".join('  printf("%%zu\\n", offsetof(lbm_params, %s));\n' % n for n in fields)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = [int(v) for v in subprocess.run([str(exe)], stdout=subprocess.PIPE, check=True,
                                          text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(abi.Params)
    assert out[1:] == [getattr(abi.Params, n).offset for n in fields]


def test_fortran_params_type_matches_header(tmp_path):
    """The Fortran binding's TYPE, BIND(C) :: LBM_PARAMS has the C struct's
    size and the offsets of its last fields (flang's C_SIZEOF and LOC
    differences), so a Fortran caller passes the same block."""
    flang = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(flang):
        pytest.skip("flang not present")
    abi = pkg()._abi
    src = tmp_path / "layout.f90"
    src.write_text(
        "PROGRAM LAYOUT\n"
        "  USE LBM_AMD\n"
        "  USE ISO_C_BINDING\n"
        "  IMPLICIT NONE\n"
        "  TYPE(LBM_PARAMS), TARGET :: P\n"
        "  INTEGER(C_INTPTR_T) :: B\n"
        "  B = TRANSFER(C_LOC(P), B)\n"
        "  PRINT *, C_SIZEOF(P)\n"
        "  PRINT *, TRANSFER(C_LOC(P%OSC_PERIOD), B) - B\n"
        "  PRINT *, TRANSFER(C_LOC(P%CONTACT), B) - B\n"
        "  PRINT *, TRANSFER(C_LOC(P%WALL_EDGE), B) - B\n"
        "END PROGRAM LAYOUT\n")
    exe = tmp_path / "layout"
    pkgdir = os.path.join(REPO, "2d_nnf_lbm_amd")
    mod = os.path.join(pkgdir, "fortran", "lbm_amd.f90")
    abi.load()  # the library must be built (the module's helpers reference it)
    subprocess.run([flang, "-J", str(tmp_path), mod, str(src), "-o", str(exe), "-L" + pkgdir,
                    "-llbm_amd", "-Wl,-rpath," + pkgdir], check=True, cwd=str(tmp_path),
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    out = [int(v) for v in subprocess.run([str(exe)], stdout=subprocess.PIPE, check=True,
                                          text=True).stdout.split()]
    P = abi.Params
    assert out == [ctypes.sizeof(P), P.osc_period.offset, P.contact.offset, P.wall_edge.offset]


@pytest.mark.parametrize("bad", [
    dict(omega=2.5),            # tau = 0.4 < 0.5 (F6:115-118)
    dict(nx=1),
    dict(rheology=7),
    dict(storage=9),
    dict(max_particles=-1),
    dict(boundary=3),
    dict(contact=3),
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
