"""Readers of the text files the reference program (and the Fortran driver,
which keeps its formats) writes: the state dump of F6:389-403, the field file
of WriteResults (F6:1416-1463), the particle history (F6:315-318), the mean
density (F6:348-349) and the start-up constants (F6:104-111).  flang prints
list-directed doubles in shortest round-trip form, so parsed values are exact.
"""
import os

import numpy as np


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
    pfile = os.path.join(d, "00_Simulation_Parameters.txt")
    if not os.path.exists(pfile):
        return out, consts
    with open(pfile) as fh:
        for line in fh:
            if "=" in line:
                k, v = line.split("=", 1)
                try:
                    consts[k.strip()] = float(v.split()[0])
                except ValueError:
                    pass
    return out, consts
