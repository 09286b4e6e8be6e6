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

Usage: python tests/golden/make_golden.py   (needs oracle/_ref/ built here)
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
import build_ref  # noqa: E402


def _tokens(path):
    with open(path) as fh:
        return fh.read().split()


def _floats(path):
    return np.array([float(t) for t in _tokens(path)], dtype=np.float64)


def parse_run(d, nx, ny):
    nn = (nx + 1) * (ny + 1)
    out = {}
    # I outer, J inner (F6:389-399) -> reorder to [j][i]
    def grid(name, ncomp):
        a = _floats(os.path.join(d, name))
        assert a.size == nn * ncomp, (name, a.size)
        a = a.reshape(nx + 1, ny + 1, ncomp).transpose(1, 0, 2)
        return np.ascontiguousarray(a if ncomp > 1 else a[..., 0])
    out["f"] = grid("0_F_output.txt", 9)
    out["ux"] = grid("0_UX_output.txt", 1)
    out["uy"] = grid("0_UY_output.txt", 1)
    out["rho"] = grid("0_RHO_output.txt", 1)
    out["g"] = grid("0_Gspm_output.txt", 2)
    out["omega_bn"] = grid("0_OmegaBN_output.txt", 1)
    out["image"] = grid("0_DomainImage_output.txt", 1).astype(np.int32)
    out["particle"] = _floats(os.path.join(d, "0_Particle_output.txt"))
    out["omega_mrt"] = _floats(os.path.join(d, "0_OmegaMRT_output.txt"))
    # velocity field file: 2 header lines then 9 blocks of (ny+1)*(nx+1)
    vf = [f for f in os.listdir(d) if f.startswith("02_VelocityField_")]
    assert len(vf) == 1, vf
    with open(os.path.join(d, vf[0])) as fh:
        lines = fh.read().split("\n")[2:]
    blk = np.array([float(t) for t in " ".join(lines).split()])
    blk = blk.reshape(9, ny + 1, nx + 1)
    out["stress"] = np.ascontiguousarray(np.stack([blk[6], blk[7], blk[8]], axis=-1))
    # per-step history (skip header)
    with open(os.path.join(d, "01_ParticleTimehistory.plt")) as fh:
        rows = fh.read().split("\n")[1:]
    toks = " ".join(rows).split()
    out["hist"] = np.array([float(t) for t in toks]).reshape(-1, 11)
    with open(os.path.join(d, "01_AveragedDensity.PLT")) as fh:
        rows = fh.read().split("\n")[1:]
    toks = " ".join(rows).split()
    out["avgden"] = np.array([float(t) for t in toks]).reshape(-1, 2)[:, 1]
    consts = {}
    with open(os.path.join(d, "00_Simulation_Parameters.txt")) as fh:
        for line in fh:
            if "=" in line:
                k, v = line.split("=", 1)
                try:
                    consts[k.strip()] = float(v.split()[0])
                except ValueError:
                    pass
    return out, consts


def main():
    os.makedirs(HERE, exist_ok=True)
    for name, cfg in build_ref.VARIANTS.items():
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
    main()
