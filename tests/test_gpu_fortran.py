"""End-to-end drop-in check: the Fortran driver (the reference's MAIN loop on
the GPU through the iso_c_binding module) writes the reference's own output
files, which must match the reference program's files for the same case."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cases
from lbmtest import rel_err
from test_gpu_parity import check_field_errors, field_errors
from refio import parse_run

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "2d_nnf_lbm_amd", "fortran", "lbm_f6_driver")


@pytest.mark.parametrize("name,fused", [("f6_bing_local", True), ("f6_bing_local", False),
                                        ("f6_newt", True), ("f6_newt_lid", False),
                                        ("f6_cas_local", True), ("f6_zouhe_lid", True),
                                        ("f6_eq_bing_local", True), ("f6_eq_bing_local", False)])
def test_fortran_driver_matches_reference_files(name, fused):
    """The driver's output files against the reference program's for the same
    case, including the variants selected by the RHEOLOGY and BOUNDARY
    namelist switches (Casson; F1's Zou-He and F4's equilibrium edges,
    against builds of F6 with those routines)."""
    assert os.path.exists(DRIVER), "build the Fortran driver first"
    meta, ref = cases.load_golden(name)
    nx, ny, n = meta["nx"], meta["ny"], meta["time_max"]
    uxn = meta.get("uxn") or meta.get("zouhe_uxn") or "0D0"
    extra = ""
    if meta.get("rheology") == "casson":
        extra += ", RHEOLOGY=2"
    if meta.get("boundary") == "zouhe":
        extra += ", BOUNDARY=1"
    elif meta.get("boundary") == "equilibrium":
        extra += ", BOUNDARY=2"
    print("case: fortran driver", name, "fused" if fused else "subroutines", flush=True)
    with tempfile.TemporaryDirectory(prefix="fdrv_") as d:
        nml = os.path.join(d, "run.nml")
        with open(nml, "w") as fh:
            fh.write("&RUN NX=%d, NY=%d, TIME_MAX=%d, NON_NEWTONIAN=%s, WRITE_OUTPUT=.TRUE., "
                     "FUSED=%s, STORAGE=0, UXN=%s%s /\n"
                     % (nx, ny, n, ".TRUE." if meta["non_newtonian"] else ".FALSE.",
                        ".TRUE." if fused else ".FALSE.", uxn, extra))
        r = subprocess.run([DRIVER, nml], cwd=d, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        print(r.stdout[-2000:], flush=True)
        assert r.returncode == 0
        out, _ = parse_run(d, nx, ny)
    check_field_errors(field_errors(out["f"], out["rho"], out["ux"], out["uy"], out["g"],
                                    out["omega_bn"], ref), name)
    assert np.max(np.abs(out["stress"] - ref["stress"])) <= 1e-13
    h, hr = out["hist"], ref["hist"]
    assert h.shape == hr.shape
    for col in (1, 2, 5):     # Y, X, Uy/U_C (F6:315-316); Ux and OMG stay at round-off
        assert rel_err(h[:, col], hr[:, col]) <= 1e-9, col
    assert rel_err(out["avgden"], ref["avgden"]) <= 1e-13
    assert rel_err(out["particle"][:2], ref["particle"][:2]) <= 1e-9
    if not meta["non_newtonian"]:
        assert np.array_equal(out["omega_mrt"], ref["omega_mrt"])


SHIM_CASES = ["f6_bing_lagged", "f6_newt_lid", "f6_newt_mohammad"]


@pytest.mark.parametrize("name", SHIM_CASES)
def test_reference_main_with_gpu_subroutines(name):
    """The reference program itself (oracle/_ref/f6_shim_<case>: F6's MAIN and
    every routine but the nine time-step subroutines, which come from
    2d_nnf_lbm_amd/fortran/lbm_f6_shim.F90 and run on the GPU) writes the
    same files as the all-CPU reference.  f6_bing_lagged keeps the
    reference's lagged module-global rate."""
    exe = os.path.join(REPO, "oracle", "_ref", "f6_shim_" + name[3:])
    if not os.path.exists(exe):
        pytest.skip("shim programs are built by oracle/build_ref.py where the reference exists")
    meta, ref = cases.load_golden(name)
    nx, ny = meta["nx"], meta["ny"]
    print("case: reference MAIN with GPU subroutines", name, flush=True)
    with tempfile.TemporaryDirectory(prefix="fshim_") as d:
        r = subprocess.run([exe], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=600)
        print(r.stdout[-1500:], flush=True)
        assert r.returncode == 0
        out, _ = parse_run(d, nx, ny)
    check_field_errors(field_errors(out["f"], out["rho"], out["ux"], out["uy"], out["g"],
                                    out["omega_bn"], ref), name)
    assert np.max(np.abs(out["stress"] - ref["stress"])) <= 1e-13
    h, hr = out["hist"], ref["hist"]
    assert h.shape == hr.shape
    for col in (1, 2, 5):
        assert rel_err(h[:, col], hr[:, col]) <= 1e-9, col
    assert rel_err(out["avgden"], ref["avgden"]) <= 1e-13
    assert rel_err(out["omega_mrt"], ref["omega_mrt"]) <= 1e-8
    assert np.mean(out["image"] == ref["image"]) > 0.999
