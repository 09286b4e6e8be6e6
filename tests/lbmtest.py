"""Shared helpers for the parity tests: set up the oracle and the HIP path
from the same configuration, and compare them."""
import importlib

import numpy as np

import cases


def pkg():
    return importlib.import_module("2d_nnf_lbm_amd")


def gpu_solver(cfg, storage="fp64", max_particles=8):
    c = dict(cfg)
    c["storage"] = storage
    c["max_particles"] = max_particles
    if c.get("rheology", 0) != 0:
        c["rate_mode"] = 1
    return pkg().D2Q9(c)


def oracle_state(orc, cfg, part):
    p = orc.make_params(cfg)
    st = orc.State(p, part.shape[1] if part is not None else 0)
    st.initialize()
    if part is not None and part.shape[1]:
        st.set_particles(part)
    return st


def f6_case(nx, ny, bingham, uxn=0.0):
    cfg = cases.f6_params(nx, ny, bingham, local_rate=bingham, uxn=uxn)
    part = cases.f6_disk(nx, ny)[:, None]
    return cfg, part


def rel_err(a, b):
    """max |a-b| / max |b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    return 0.0 if num == 0 else num / (den if den > 0 else 1.0)


def disk_record(x, y, diam, rho_p, zeta=1.0, vx=0.0, vy=0.0, w=0.0):
    rec = np.zeros(22)
    rec[0], rec[1] = x, y
    rec[3], rec[4], rec[5] = vx, vy, w
    rec[6] = diam
    rec[12] = rho_p
    rec[13] = rho_p * cases.PI * (0.5 * diam) * (0.5 * diam)
    rec[14] = 0.5 * rec[13] * (0.5 * diam) * (0.5 * diam)
    rec[15] = zeta
    return rec
