"""Run configurations of the benchmark workloads, as lbm_params dicts plus
initial particle records.

Every float literal that the reference writes as a default REAL is rounded to
float32 first (f32), as in the Fortran source.
"""
import math

import numpy as np

from . import _abi


def f32(x):
    return float(np.float32(x))


PI = f32(3.14159265359)  # F6:3


def f6_rates():
    """MRT rates of ComputeRelaxation_MRT (F6:1009-1011)."""
    return dict(q_mode=1, s0=0.0, s1=f32(1.1), s2=1.25, s3=0.0, s5=0.0, sq=0.0)


def mohammad_rates():
    """The alternative rate set of F6:1017-1018 / F1:1044-1045."""
    return dict(q_mode=0, s0=1.0, s1=f32(1.4), s2=f32(1.4), s3=1.0, s5=1.0, sq=f32(1.2))


def disk(x, y, diam, rho_p, zeta):
    """Particle record of a disk at rest (F6:161-169)."""
    rec = np.zeros(_abi.PF_COUNT)
    rec[_abi.PF["x"]], rec[_abi.PF["y"]] = x, y
    rec[_abi.PF["diam"]] = diam
    rec[_abi.PF["rhop"]] = rho_p
    m_p = rho_p * PI * (0.5 * diam) * (0.5 * diam)
    rec[_abi.PF["mp"]] = m_p
    rec[_abi.PF["ip"]] = 0.5 * m_p * (0.5 * diam) * (0.5 * diam)
    rec[_abi.PF["zeta"]] = zeta
    return rec


def two_disk_bingham(storage="fp64"):
    """Two disks settling in a Papanastasiou-Bingham fluid.

    Grid 400 x 1600 cells (401 x 1601 nodes), disks of diameter 40 at
    (198, 1300) and (200, 1380), density ratio 1.01, plastic viscosity from
    tau = 0.8 (mu0 = 0.1), regularisation m = 1e5, Bingham number 0.5.
    Gravity follows the reference's definition A_G = -Re*nu^2/(L^3 drho rho)
    (F6:28) with Re = 0.78 (F6:16) and L = D; the yield stress follows
    TAU_Y = Bn*drho*|g|*L (F6:30).  Interface width 2, bounce-back walls,
    F6's MRT rates and Glowinski contact constants (F6:691-693).
    """
    nx, ny = 400, 1600
    d = 40.0
    rho_s, rho_f = 1.01, 1.0
    tau = 0.8
    mu0 = (tau - 0.5) / 3.0
    re_s = f32(0.78)
    a_g = -re_s * mu0 * mu0 / ((d * d * d) * (rho_s - rho_f) * rho_f)
    bn = 0.5
    tau_y = bn * (rho_s - rho_f) * (-a_g) * d
    zeta = 2.0
    omega = 1.0 / (3.0 * mu0 + 0.5)
    cij = -((((((rho_s - rho_f) * a_g) * PI) * 0.5) * d) * 0.5) * d
    cfg = dict(nx=nx, ny=ny, rheology=_abi.LBM_BINGHAM, rate_mode=_abi.LBM_RATE_LOCAL,
               update_angle=1, storage=storage, max_particles=2, omega=omega, nu=mu0,
               mu0=mu0, tau_y=tau_y, m_pap=1e5, gamma_min=f32(1e-15), rho0=1.0, gy=a_g,
               wall=[0.0] * 8, corner=[0.0, 0.0], zeta_c=1.0, cij=cij,
               eps_p=0.5, eps_w=0.5, e_p=10.0, e_w=10.0)
    cfg.update(f6_rates())
    part = np.stack([disk(198.0, 1300.0, d, rho_s, zeta), disk(200.0, 1380.0, d, rho_s, zeta)],
                    axis=1)
    return cfg, part


def nodes(cfg):
    return (cfg["nx"] + 1) * (cfg["ny"] + 1)


# Algorithmic bytes per node update of the fused step: each stored
# population read once and written once, plus the per-node rate of the
# Bingham path (read and written).
def bytes_per_node(storage, bingham):
    pop = {"fp64": 8, "fp32": 4, "fp16": 2, "bf16": 2}[storage]
    rate = (8 if storage == "fp64" else 4) if bingham else 0
    return 2 * (9 * pop + rate)


def settling_metrics(hist, np_, cfg, p=0):
    """Terminal velocity, drag coefficient C_d = (F_y + F_in,y)/(nu*(-V_y))
    (F6:317) of particle p at the last step."""
    H = _abi.H
    row = hist[-1]
    base = p * _abi.H_PCOUNT
    vy = row[base + H["vy"]]
    fy_prev = hist[-2][base + H["fy"]] if len(hist) > 1 else 0.0
    cd = (fy_prev + row[base + H["fhy"]]) / (cfg["nu"] * -vy) if vy != 0 else math.nan
    return vy, cd
