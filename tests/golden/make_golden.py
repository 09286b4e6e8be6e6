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

contact_models.npz from oracle/_ref/f4_contact_wt, f4_contact_lj and
f1_contact_wt (the reference's Wan-Turek and Lennard-Jones contact forces,
driven by oracle/contact_model_driver.f90): wall edges and disks in, F_COL out;
and f3_contact.npz from oracle/_ref/f3_contact (the reference's N-body
ComputeCollisionForce, F3:632-806, driven by oracle/f3_contact_driver.f90):
disk positions and diameters in, F_COL of the second and third call out;
f3_multi.npz and f2_dkt.npz from oracle/_ref/f3_multi and f2_dkt (F3's
whole N-particle program, 12 disks, 300 steps; F2's two-disk program, 400
steps): disk centres in, per-step disk states and the final F, UX, UY, RHO,
G out.

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


def contact_model_cases(name, nx, ny):
    """Inputs for the Wan-Turek / Lennard-Jones builds: per case the wall
    edges and two disks (Wan-Turek builds also get one-disk cases).  Pair
    distances sweep from inside r_i + r_j to beyond r_i + r_j + 2 (every
    branch for zeta 1 and 2, plus d = r_i + r_j exactly); wall cases put one
    disk near one or two walls of the model's wall geometry (Wan-Turek:
    mirror images about the edges; Lennard-Jones: the images of F4:600-621
    on the module's NX x NY box) with the other disk out of range.
    Coordinates are multiples of 1/64 so their decimal text is exact."""
    rng = np.random.default_rng(20190926 + len(name))
    q = lambda v: float(np.round(v * 64) / 64)
    lj = name.endswith("_lj")
    boxes = [(0.0, 100.0, 80.0, 0.0), (2.5, 97.25, 77.5, 3.75)]
    cases = []
    for k in range(40):  # pairs near (50, 40), out of every wall's range
        d1, d2 = (float(rng.choice([8.0, 10.0, 12.0, 16.0])) for _ in range(2))
        s = 0.5 * d1 + 0.5 * d2
        dist = rng.uniform(0.7 * s, s + 2.6)
        a = rng.uniform(0, 2 * np.pi)
        x1, y1 = q(50 + rng.uniform(-2, 2)), q(40 + rng.uniform(-2, 2))
        x2, y2 = q(x1 + dist * np.cos(a)), q(y1 + dist * np.sin(a))
        cases.append((boxes[k % 2], [(x1, y1, d1), (x2, y2, d2)]))
    cases.append((boxes[0], [(40.0, 40.0, 10.0), (50.0, 40.0, 10.0)]))  # d = r_i + r_j
    cases.append((boxes[0], [(40.0, 40.0, 10.0), (40.0, 51.0, 12.0)]))  # the same, along y
    far = (50.0, 40.0, 8.0)
    for k in range(48):  # one disk near a wall (two near the corners)
        dm = float(rng.choice([8.0, 10.0, 16.0]))
        box = boxes[k % 2]
        w = k % 4
        if lj:
            # wall distances of the L-J images: y, NX - x, 2 (NY - y), 2 x
            g = rng.uniform(0.5 * dm, dm + 1.5)
            x, y = q(0.5 * nx), q(0.5 * ny)
            if w == 0: y = q(g)
            elif w == 1: x = q(nx - g)
            elif w == 2: y = q(ny - 0.5 * g)
            else: x = q(0.5 * g)
            if k >= 40:
                x, y = q(0.5 * g), q(g)
            cases.append((box, [(x, y, dm), (q(0.5 * nx + 100.0), q(0.5 * ny), 8.0)]))
            continue
        gap = rng.uniform(0.25 * dm, 0.5 * dm + 1.3)  # mirror distance 2 gap
        x, y = q(0.5 * (box[1] + box[3])), q(0.5 * (box[0] + box[2]))
        if w == 0: y = q(box[0] + gap)
        elif w == 1: x = q(box[1] - gap)
        elif w == 2: y = q(box[2] - gap)
        else: x = q(box[3] + gap)
        if k >= 40:
            x, y = q(box[3] + gap), q(box[0] + gap)
        far_disk = far if abs(x - far[0]) + abs(y - far[1]) > 30 else (q(x + 40.0), q(y + 20.0), 8.0)
        cases.append((box, [(x, y, dm), far_disk]))
    if not lj:
        cases.append((boxes[0], [(5.0, 40.0, 10.0)]))   # mirror distance = D exactly
        cases.append((boxes[0], [(4.75, 3.0, 10.0)]))   # one disk, two walls
        cases.append((boxes[1], [(50.0, 74.0, 8.0)]))
    return cases


def make_contact_models():
    out = {}
    for name in build_ref.CONTACT_MODELS:
        exe = os.path.join(build_ref.OUT_DIR, name)
        if not os.path.exists(exe):
            raise SystemExit("missing %s: run python oracle/build_ref.py" % exe)
        nx, ny = (int(v) for v in subprocess.run([exe], input="0\n", stdout=subprocess.PIPE,
                                                  check=True, text=True).stdout.split())
        cases = contact_model_cases(name, nx, ny)
        text = "%d\n" % len(cases)
        for box, disks in cases:
            text += "%r %r %r %r\n%d\n" % (box + (len(disks),))
            text += "".join("%r %r %r\n" % d for d in disks)
        r = subprocess.run([exe], input=text, stdout=subprocess.PIPE, check=True, text=True)
        rows = [ln.split() for ln in r.stdout.strip().split("\n")][1:]
        out[name + "_nx"] = np.array([nx, ny])
        out[name + "_edges"] = np.array([box for box, _ in cases])
        out[name + "_n"] = np.array([len(d) for _, d in cases])
        out[name + "_disks"] = np.array([d for _, ds in cases for d in ds])
        out[name + "_fcol"] = np.array([[float(v) for v in row] for row in rows])
        assert len(out[name + "_fcol"]) == len(out[name + "_disks"])
        nz = np.count_nonzero(np.any(out[name + "_fcol"] != 0, axis=1))
        print(name, len(cases), "cases,", nz, "nonzero forces")
    path = os.path.join(HERE, "contact_models.npz")
    np.savez_compressed(path, **out)
    print("wrote", os.path.relpath(path, REPO), os.path.getsize(path), "bytes")


def f3_multi_positions():
    """Disk centres of the F3 multi-particle run (12 disks of F3's D = 16 in
    a 160 x 160 cavity), multiples of 1/64 so their decimal text is exact:
      * two pairs and a triple whose surfaces are 0.25-0.75 lu apart, so their
        smoothed profiles (phi > 0 within r + zeta/2) overlap and the
        highest-index-wins overwrite of G (F6:487-494) decides nodes, and the
        pair spring (d <= r_i + r_j + zeta, F3:746-770) acts;
      * one disk within wall range of the bottom wall (F3:773-804);
      * the last disk 7.75 lu from the left wall: its left-wall distance 15.5
        <= r_i + r_j is the stale D_PW the next call's pair branch reads
        (F3:760), so later steps take the E_P term."""
    q = lambda v: float(np.round(v * 64) / 64)
    pts = [(40.0, 120.0), (q(40.0 + 16.25 * np.cos(0.3)), q(120.0 + 16.25 * np.sin(0.3))),
           (110.0, 125.0), (q(110.0 + 16.75 * np.cos(2.0)), q(125.0 + 16.75 * np.sin(2.0))),
           (75.0, 80.0), (91.5, 80.0), (q(83.25 + 0.25), q(80.0 + 14.5)),
           (128.0, 8.375), (30.0, 60.0), (130.0, 60.0), (80.0, 140.0), (7.75, 100.0)]
    return np.array(pts)


def make_nparticle(name):
    """tests/golden/<name>.npz from oracle/_ref/<name> (a whole
    multi-particle reference program, oracle/build_ref.py NPARTICLE_BUILDS):
    the disk centres in, the run constants, the per-step CENTER, VELOCITY,
    FORCE and ANGLE of every disk and the final state dump out."""
    cfg = build_ref.NPARTICLE_BUILDS[name]
    exe = os.path.join(build_ref.OUT_DIR, name)
    if not os.path.exists(exe):
        raise SystemExit("missing %s: run python oracle/build_ref.py" % exe)
    if cfg["src"] == "F3":
        pos = f3_multi_positions()
        assert len(pos) == cfg["n_particle"]
    else:
        pos = np.array([[float(v.replace("D", "e")) for v in c] for c in cfg["centers"]])
    n = len(pos)
    with tempfile.TemporaryDirectory(prefix="golden_", dir="/tmp") as d:
        if cfg["src"] == "F3":
            with open(os.path.join(d, "0_Particle_Initial_Position.txt"), "w") as fh:
                fh.write("".join("%r %r\n" % tuple(map(float, p)) for p in pos))
        subprocess.run([exe], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        with open(os.path.join(d, "ref_constants.txt")) as fh:
            omega, nu, a_g, rho_s, rho_f, pi = (float(t) for t in fh.read().split())
        with open(os.path.join(d, "ref_history.txt")) as fh:
            toks = fh.read().split()
        rows = np.array([float(t) for t in toks]).reshape(cfg["time_max"], 1 + 9 * n)
        assert np.array_equal(rows[:, 0], np.arange(1, cfg["time_max"] + 1))
        hist = rows[:, 1:].reshape(cfg["time_max"], n, 9)  # x y vx vy w fx fy tq angle
        nx, ny = cfg["nx"], cfg["ny"]
        nn = (nx + 1) * (ny + 1)

        def grid(fname, nc):
            with open(os.path.join(d, fname)) as fh:
                a = np.array([float(t) for t in fh.read().split()])
            assert a.size == nn * nc, fname
            a = a.reshape(nx + 1, ny + 1, nc).transpose(1, 0, 2)
            return np.ascontiguousarray(a if nc > 1 else a[..., 0])
        out = dict(f=grid("0_F_output.txt", 9), ux=grid("0_UX_output.txt", 1),
                   uy=grid("0_UY_output.txt", 1), rho=grid("0_RHO_output.txt", 1),
                   g=grid("0_Gspm_output.txt", 2))
    meta = dict(nx=cfg["nx"], ny=cfg["ny"], time_max=cfg["time_max"], src=cfg["src"],
                omega=omega, nu=nu, a_g=a_g, rho_s=rho_s, rho_f=rho_f, pi=pi)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), pos=pos, hist=hist, **out)
    print("wrote", os.path.relpath(path, REPO), os.path.getsize(path), "bytes")


def main(names):
    """names: the variants to write (default: every variant and the contact
    fixture)."""
    os.makedirs(HERE, exist_ok=True)
    if not names or "contact_models" in names:
        make_contact_models()
    if not names:
        make_contact()
    for name in build_ref.NPARTICLE_BUILDS:
        if not names or name in names:
            make_nparticle(name)
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
