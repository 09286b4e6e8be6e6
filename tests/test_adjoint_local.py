"""CPU check of the hand-written per-node adjoint (collision, smoothed-profile
forcing, hydrodynamic-force contribution, Papanastasiou rate): its input
adjoints must equal fourth-order central differences of the same forward
pass, over 300 random near-equilibrium states that cover both rate sets,
Newtonian and Bingham nodes, strain rates from 1e-7 to 1e-3 and nodes
outside the particle (phi = 0)."""
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_local_adjoint_matches_finite_differences():
    with tempfile.TemporaryDirectory(prefix="adjlocal_") as d:
        exe = os.path.join(d, "check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe,
                        os.path.join(HERE, "adjoint_local_check.cpp"), "-lm"], check=True)
        out = subprocess.run([exe, "300", "1e-9"], check=True, stdout=subprocess.PIPE,
                             text=True).stdout
    worst = float(re.search(r"worst ([0-9.e+-]+)", out).group(1))
    assert worst <= 1e-4, out[-2000:]
    assert out.count("rheology 0") > 50 and out.count("rheology 1") > 150
