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
sys.path.insert(0, os.path.join(REPO, "tests"))
import build_ref  # noqa: E402
from refio import parse_run  # noqa: E402


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
