"""Write the golden fixtures under tests/golden/ from the compiled reference.

Runs each executable built by oracle/build_ref.py (oracle/_ref/<variant>) in a
scratch directory, parses what the reference program itself writes (all
list-directed text, which flang prints in shortest round-trip form, so the
parsed doubles are exact), and stores them as compressed .npz fixtures:

  final state dump (WRITE_OUTPUT block, F6:377-405): F, UX, UY, RHO, G,
      OMEGA_BN, DOMAIN_IMAGE, particle CENTER/VELOCITY/FORCE, OMEGA_MRT;
  final field file (WriteResults, F6:1416-1463): STRESS;
  per-step particle history (F6:315-318) and mean density (F6:348-349);
  run constants printed at start-up (F6:104-111): TAU_LB, U_C, TAU_Y, A_G.

and f3_contact.npz from oracle/_ref/f3_contact (the reference's N-body
ComputeCollisionForce, F3:632-806, driven by oracle/f3_contact_driver.f90):
disk positions and diameters in, F_COL of the second and third call out.

Usage: python tests/golden/make_golden.py [variant ...]   (needs oracle/_ref/ built here)
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))
import build_ref  # noqa: E402
from refio import parse_run  # noqa: E402


def contact_positions():
    """40 disks in the reference's 512 x 512 box: clusters with touching and
    overlapping pairs (distances inside r_i + r_j + 1), disks within contact
    range of each wall, mixed diameters.  Coordinates are multiples of 1/64
    so their decimal text is exact.  Disk 1 sits far from the left wall
    (its left-wall distance exceeds every r_i + r_j); the last disk touches
    the left wall (distance 7.5 + 8 <= 16)."""
    rng = np.random.default_rng(20191003)
    pts = [(200.0, 300.0, 16.0)]
    centres = [(120.0, 200.0), (300.0, 100.0), (420.0, 380.0), (60.0, 450.0)]
    for cx, cy in centres:
        for _ in range(8):
            d = float(rng.choice([12.0, 16.0, 16.0, 20.0]))
            x = cx + np.round(rng.uniform(-14, 14) * 64) / 64
            y = cy + np.round(rng.uniform(-14, 14) * 64) / 64
            pts.append((x, y, d))
    pts += [(256.0, 8.5, 16.0), (503.25, 256.0, 16.0), (300.0, 503.0, 16.0),
            (9.0, 9.0, 16.0), (502.0, 505.0, 12.0), (100.0, 60.0, 16.0), (7.5, 100.0, 16.0)]
    assert len(pts) == 40
    return np.array(pts)


def make_contact():
    exe = os.path.join(build_ref.OUT_DIR, "f3_contact")
    if not os.path.exists(exe):
        raise SystemExit("missing %s: run python oracle/build_ref.py" % exe)
    pos = contact_positions()
    text = "%d\n" % len(pos) + "".join("%r %r %r\n" % tuple(map(float, p)) for p in pos)
    r = subprocess.run([exe], input=text, stdout=subprocess.PIPE, check=True, text=True)
    rows = [ln.split() for ln in r.stdout.strip().split("\n")]
    a_g, rho_s, rho_f, pi = (float(v) for v in rows[0][:4])
    nx, ny = int(rows[0][4]), int(rows[0][5])
    fc = np.array([[float(v) for v in row] for row in rows[1:]])
    n = len(pos)
    path = os.path.join(HERE, "f3_contact.npz")
    meta = dict(a_g=a_g, rho_s=rho_s, rho_f=rho_f, pi=pi, nx=nx, ny=ny)
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), x=pos[:, 0], y=pos[:, 1],
                        diam=pos[:, 2], fcol_call2=fc[:n], fcol_call3=fc[n:])
    print("wrote", os.path.relpath(path, REPO), os.path.getsize(path), "bytes")


def main(names):
    """names: the variants to write (default: every variant and the contact
    fixture)."""
    os.makedirs(HERE, exist_ok=True)
    if not names:
        make_contact()
    for name, cfg in build_ref.VARIANTS.items():
        if names and name not in names:
            continue
        exe = os.path.join(build_ref.OUT_DIR, name)
        if not os.path.exists(exe):
            raise SystemExit("missing %s: run python oracle/build_ref.py" % exe)
        with tempfile.TemporaryDirectory(prefix="golden_", dir="/tmp") as d:
            subprocess.run([exe], cwd=d, check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL)
            out, consts = parse_run(d, cfg["nx"], cfg["ny"])
        meta = dict(cfg)
        meta["consts"] = consts
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **out)
        print("wrote", os.path.relpath(path, REPO), os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
