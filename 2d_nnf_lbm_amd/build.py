"""Build the HIP library liblbm_amd.so in-tree for gfx950.

The library is compiled with contraction off (-ffp-contract=off) so the FP64
path keeps the reference's rounding, and with packed FP32 instructions
disabled.
"""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm_amd.so")
SOURCES = [os.path.join(CSRC, "lbm_amd.hip")]
# every source and header the library is built from (a change to any of
# them makes the built library stale)
DEPS = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.h")) +
              glob.glob(os.path.join(os.path.dirname(HERE), "include", "*.h")))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN_DIR, "lbm_f6_driver")
FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")


def build_fortran():
    """Compile the iso_c_binding module and the Fortran driver, linked to
    liblbm_amd.so (found next to the package through the run path)."""
    import tempfile
    srcs = [os.path.join(FORTRAN_DIR, "lbm_amd.f90"), os.path.join(FORTRAN_DIR, "lbm_f6_driver.f90")]
    with tempfile.TemporaryDirectory(prefix="lbmmod_") as moddir:
        cmd = [FLANG, "-O2", "-J", moddir] + srcs + [
            "-L" + HERE, "-llbm_amd", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(force=True, verbose=False):
    """Compile liblbm_amd.so (always when force, else only when stale)."""
    if not force and os.path.exists(LIB):
        t = os.path.getmtime(LIB)
        if all(os.path.getmtime(d) <= t for d in DEPS):
            return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + SOURCES
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout)
    if verbose:
        print(r.stdout)
    os.replace(tmp, LIB)
    if os.path.exists(FLANG):
        build_fortran()
    return LIB


if __name__ == "__main__":
    print(build(force=True))
