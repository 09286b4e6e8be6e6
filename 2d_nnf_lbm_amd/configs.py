"""Run configurations of the benchmark workloads, as lbm_params dicts plus
initial particle records.

Every float literal that the reference writes as a default REAL is rounded to
float32 first (f32), as in the Fortran source.
"""
import math
import os

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


def read_initial_positions(path, n=None, diam=None, rho_p=None, zeta=1.0):
    """Particle records from the reference's initial-condition file
    0_Particle_Initial_Position.txt, read as its list-directed READs do.

    With diam and rho_p given, each record holds CENTER only (F3:145-149:
    READ(200,*) DISK(IP)%CENTER, with DIAM = L_C and RHO_P = RHO_S set from
    the parameters, F3:152-153); without them, CENTER, DIAM, RHO_P
    (F1:170-174).  Mass and inertia follow F1:177-178 / F3:154-155, with
    ZITA = zeta (F1:176).  Each READ starts at a new line and takes its
    values from as many lines as it needs (blank lines skipped); values left
    on its last line are ignored.  n: number of particles to read
    (N_PARTICLE); by default every record in the file.  Returns a
    (PF_COUNT, n) array."""
    per = 2 if diam is not None else 4
    if (diam is None) != (rho_p is None):
        raise ValueError("give both diam and rho_p (CENTER-only records) or neither")
    with open(path) as fh:
        lines = [ln.replace("D", "E").replace("d", "e").split() for ln in fh]
    recs, k = [], 0
    while (n is None or len(recs) < n) and k < len(lines):
        vals = []
        while len(vals) < per and k < len(lines):
            vals.extend(lines[k])
            k += 1
        if len(vals) < per:
            if vals:
                raise ValueError("%s: incomplete record %d" % (path, len(recs) + 1))
            break
        vals = [float(v) for v in vals[:per]]
        d, r = (diam, rho_p) if per == 2 else (vals[2], vals[3])
        recs.append(disk(vals[0], vals[1], float(d), float(r), zeta))
    if n is not None and len(recs) < n:
        raise ValueError("%s: %d records, expected %d" % (path, len(recs), n))
    return np.stack(recs, axis=1)


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


def one_disk_newtonian(storage="fp64"):
    """One disk falling from rest in a Newtonian fluid in a closed channel.

    Grid 200 x 600 cells (201 x 601 nodes), disk of diameter 25 at (100, 480),
    density ratio 1.5, rho0 = 1, tau = 0.6 (nu = 1/30), the MRT rate set
    {1, 1.4, 1.4, 1, 1.2, 1, 1.2, 1/tau, 1/tau} (F6:1017-1018, F1:1044-1045,
    float literals as in the source), interface width 2, gravity 1e-5 along
    -y, bounce-back walls, F6's Glowinski constants (F6:691-693).
    """
    nx, ny = 200, 600
    d = 25.0
    rho_s, rho_f = 1.5, 1.0
    tau = 0.6
    nu = (tau - 0.5) / 3.0
    a_g = -1e-5
    cij = -((((((rho_s - rho_f) * a_g) * PI) * 0.5) * d) * 0.5) * d
    cfg = dict(nx=nx, ny=ny, rheology=_abi.LBM_NEWTONIAN, rate_mode=_abi.LBM_RATE_LOCAL,
               update_angle=1, storage=storage, max_particles=1, omega=1.0 / tau, nu=nu,
               mu0=nu, tau_y=0.0, m_pap=0.0, gamma_min=f32(1e-15), rho0=1.0, gy=a_g,
               wall=[0.0] * 8, corner=[0.0, 0.0], zeta_c=1.0, cij=cij,
               eps_p=0.5, eps_w=0.5, e_p=10.0, e_w=10.0)
    cfg.update(mohammad_rates())
    part = disk(100.0, 480.0, d, rho_s, 2.0)[:, None]
    return cfg, part


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DENSE_FILE = os.path.join(DATA, "dense_504_seed0.txt")


def dense_positions(n=504, diam=16.0, gap=4.0, nx=2048, ny=2048, frac=0.6, seed=0):
    """Rejection sampling of n disk centres (numpy PCG64, default_rng(seed)):
    uniform in the upper `frac` of the cavity, at least `gap` lattice units
    between disks and between a disk and a wall.  Candidates are drawn as
    (x, y) pairs and accepted in order."""
    rng = np.random.default_rng(seed)
    r = 0.5 * diam
    xlo, xhi = r + gap, nx - r - gap
    ylo, yhi = (1.0 - frac) * ny + r + gap, ny - r - gap
    pts = np.zeros((0, 2))
    while len(pts) < n:
        x, y = rng.uniform(xlo, xhi), rng.uniform(ylo, yhi)
        if len(pts) == 0 or np.min(np.hypot(pts[:, 0] - x, pts[:, 1] - y)) >= diam + gap:
            pts = np.vstack([pts, [x, y]])
    return pts


def dense_sedimentation(bn=0.1, storage="bf16", positions=None, size=2048):
    """504 disks (D = 16, density ratio 1.1) settling from the upper 60 % of
    a closed 2048 x 2048 cavity (2049^2 nodes).  Positions: the committed
    file data/dense_504_seed0.txt (dense_positions() with seed 0).  Fluid:
    tau = 0.8 (mu0 = 0.1); Bn = 0 runs Newtonian, Bn > 0 Papanastasiou-Bingham
    with m = 1e5 and TAU_Y = Bn*drho*|g|*D (F6:30); gravity from
    A_G = -Re*mu0^2/(D^3 drho rho) with Re = 0.78 (F6:16, F6:28).  Contact: the
    N-body Glowinski model with F3's constants (F3:739-742).  `size` and
    `positions` give smaller cavities of the same fluid and disks for tests.
    """
    nx = ny = size
    d = 16.0
    rho_s, rho_f = 1.1, 1.0
    tau = 0.8
    mu0 = (tau - 0.5) / 3.0
    re_s = f32(0.78)
    a_g = -re_s * mu0 * mu0 / ((d * d * d) * (rho_s - rho_f) * rho_f)
    tau_y = bn * (rho_s - rho_f) * (-a_g) * d
    cij = -((((((rho_s - rho_f) * a_g) * PI) * 0.5) * d) * 0.5) * d
    if positions is None:
        positions = np.loadtxt(DENSE_FILE)
    cfg = dict(nx=nx, ny=ny, rheology=_abi.LBM_BINGHAM if bn > 0 else _abi.LBM_NEWTONIAN,
               rate_mode=_abi.LBM_RATE_LOCAL, update_angle=1, storage=storage,
               max_particles=len(positions), omega=1.0 / tau, nu=mu0, mu0=mu0, tau_y=tau_y,
               m_pap=1e5, gamma_min=f32(1e-15), rho0=1.0, gy=a_g, wall=[0.0] * 8,
               corner=[0.0, 0.0], zeta_c=1.0, cij=cij, eps_p=2.0, eps_w=1.0, e_p=20.0,
               e_w=10.0)
    cfg.update(f6_rates())
    part = np.stack([disk(float(x), float(y), d, rho_s, 2.0) for x, y in positions], axis=1)
    return cfg, part


def oscillating_cavity(side, bn, storage="fp16", wall_speed=0.05, tau=0.8, rho_p=1.5):
    """A square lid- and side-wall-driven cavity of `side` cells with one
    rigid cylinder (D = side/8) at the centre in prescribed x-oscillation:
    amplitude 0.1 D, period 2000 steps, V_x = -U cos(2 pi KK / T) with
    U = 2 pi A / T (F4:221-226).  The top wall moves along +x and the right
    wall along +y at `wall_speed`; Papanastasiou-Bingham fluid with
    mu0 = (tau - 1/2)/3, m = 1e5 and TAU_Y = Bn mu0 U_w / side; no gravity."""
    d = side / 8.0
    mu0 = (tau - 0.5) / 3.0
    period = 2000.0
    u_osc = 2.0 * PI * (0.1 * d) / period
    wall = [0.0] * 8
    wall[4] = wall_speed   # uxN, lid
    wall[3] = wall_speed   # uyR, right wall
    cfg = dict(nx=side, ny=side, rheology=_abi.LBM_BINGHAM, rate_mode=_abi.LBM_RATE_LOCAL,
               update_angle=1, storage=storage, max_particles=1,
               motion=_abi.LBM_MOTION_OSC_X, osc_u=u_osc, osc_period=period,
               omega=1.0 / tau, nu=mu0, mu0=mu0, tau_y=bn * mu0 * wall_speed / side,
               m_pap=1e5, gamma_min=f32(1e-15), rho0=1.0, gy=0.0, wall=wall,
               corner=[0.0, 0.0], zeta_c=1.0, cij=0.0, eps_p=0.5, eps_w=0.5, e_p=10.0,
               e_w=10.0)
    cfg.update(f6_rates())
    part = disk(side / 2.0, side / 2.0, d, rho_p, 2.0)[:, None]
    return cfg, part


def cavity_ensemble(n=256, seed=1, storage="fp16"):
    """n oscillating-cylinder cavities with sides drawn uniformly from
    {128..1024} and Bn uniform in [0, 10] (numpy PCG64, default_rng(seed):
    sides first, then Bn)."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(128, 1025, size=n)
    bns = rng.uniform(0.0, 10.0, size=n)
    return [oscillating_cavity(int(s), float(b), storage) for s, b in zip(sides, bns)]


def settling_adjoint(storage="fp32", tau_y=None, mu0=0.1, rho_p=1.2, nx=512, ny=1024,
                     diam=32.0):
    """One disk (D = 32, density ratio 1.2) settling from rest at
    (nx/2, ny - 7 D) in a closed Papanastasiou-Bingham channel of 512 x 1024
    cells; the parameters of the sensitivity study are the yield stress
    tau_y, the plastic viscosity mu0 and the particle density rho_p.
    Gravity is fixed at A_G = -Re mu^2/(D^3 drho rho) with Re = 0.78,
    mu = 0.1, drho = 0.2 (F6:16, F6:28) whatever mu0 and rho_p are; the
    default tau_y is Bn = 0.5 on that scale (F6:30).  OMEGA = 1/(3 mu0 + 1/2)
    (F6:101); m = 1e5; interface width 2; walls at rest."""
    rho_f = 1.0
    re_s = f32(0.78)
    a_g = -re_s * 0.1 * 0.1 / ((diam * diam * diam) * 0.2 * rho_f)
    if tau_y is None:
        tau_y = 0.5 * 0.2 * (-a_g) * diam
    cij = -((((((rho_p - rho_f) * a_g) * PI) * 0.5) * diam) * 0.5) * diam
    cfg = dict(nx=nx, ny=ny, rheology=_abi.LBM_BINGHAM, rate_mode=_abi.LBM_RATE_LOCAL,
               update_angle=1, storage=storage, max_particles=1, omega=1.0 / (3.0 * mu0 + 0.5),
               nu=mu0, mu0=mu0, tau_y=tau_y, m_pap=1e5, gamma_min=f32(1e-15), rho0=1.0,
               gy=a_g, wall=[0.0] * 8, corner=[0.0, 0.0], zeta_c=1.0, cij=cij,
               eps_p=0.5, eps_w=0.5, e_p=10.0, e_w=10.0)
    cfg.update(f6_rates())
    part = disk(nx / 2.0, ny - 7.0 * diam, diam, rho_p, 2.0)[:, None]
    return cfg, part


def kinetic_energy(hist, np_, cfg, part, p=0):
    """Translational kinetic energy 1/2 m V^2 of particle p per step."""
    H = _abi.H
    base = p * _abi.H_PCOUNT
    vx, vy = hist[:, base + H["vx"]], hist[:, base + H["vy"]]
    return 0.5 * part[_abi.PF["mp"], p] * (vx * vx + vy * vy)


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


def mean_drag_coefficient(hist, cfg, p=0, window=1000):
    """C_d of F6:317 averaged over the last `window` steps of the history
    (each step's value uses the force of the step before, as the reference's
    history line does)."""
    H = _abi.H
    base = p * _abi.H_PCOUNT
    w = min(window, len(hist) - 1)
    vals = []
    for k in range(len(hist) - w, len(hist)):
        vy = hist[k][base + H["vy"]]
        vals.append((hist[k - 1][base + H["fy"]] + hist[k][base + H["fhy"]]) / (cfg["nu"] * -vy))
    return float(sum(vals) / len(vals))
