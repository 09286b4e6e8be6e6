"""Run constants of the reference's single-disk program (F6 module PARAM,
F6:2-83), restated for a given grid so the oracle and the HIP path can be set
up exactly like a patched reference build.

Every default-REAL literal in F6 is a float32 value widened to double; f32()
reproduces that.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def f32(x):
    return float(np.float32(x))


PI = f32(3.14159265359)  # F6:3


def f6_constants(nx, ny):
    """Compile-time constants of F6's module PARAM for grid nx, ny."""
    re_s, bn_s, rho_s, rho_f, nu = f32(0.78), f32(0.033), 2.0, 1.0, f32(0.025)
    l_c = float(nx // 4)                                       # F6:26
    a_g = -(re_s * nu * nu) / ((l_c * l_c * l_c) * (rho_s - rho_f) * rho_f)  # F6:28
    u_c = (l_c * l_c) * (rho_s - rho_f) * (-a_g) / nu          # F6:29
    tau_y = bn_s * (rho_s - rho_f) * (-a_g) * l_c              # F6:30
    omega = 1.0 / (3.0 * nu + 0.5)                             # F6:101
    return dict(re_s=re_s, rho_s=rho_s, rho_f=rho_f, nu=nu, l_c=l_c, a_g=a_g,
                u_c=u_c, tau_y=tau_y, omega=omega)


def f6_params(nx, ny, non_newtonian, local_rate=False, uxn=0.0):
    c = f6_constants(nx, ny)
    wall = [0.0] * 8
    wall[4] = uxn
    d = c["l_c"]
    cij = -((((((c["rho_s"] - c["rho_f"]) * c["a_g"]) * PI) * 0.5) * d) * 0.5) * d  # F6:691
    return dict(
        nx=nx, ny=ny, rheology=1 if non_newtonian else 0,
        rate_mode=1 if local_rate else 0, q_mode=1, update_angle=1,
        omega=c["omega"], s0=0.0, s1=f32(1.1), s2=1.25, s3=0.0, s5=0.0, sq=0.0,  # F6:1009-1011
        nu=c["nu"], mu0=c["nu"], tau_y=c["tau_y"], m_pap=5.5e6, gamma_min=f32(1e-15),
        rho0=1.0, gy=c["a_g"], wall=wall, corner=[0.0, 0.0],
        zeta_c=1.0, cij=cij, eps_p=0.5, eps_w=0.5, e_p=10.0, e_w=10.0,  # F6:691-693
    )


def f6_disk(nx, ny):
    """Initial disk of F6:161-169 as a particle record (oracle field order)."""
    c = f6_constants(nx, ny)
    d = c["l_c"]
    rho_p = c["rho_s"]
    m_p = rho_p * PI * (0.5 * d) * (0.5 * d)
    i_p = 0.5 * m_p * (0.5 * d) * (0.5 * d)
    rec = np.zeros(22)
    rec[0] = 0.5 * float(nx)
    rec[1] = float(ny) - 7.0 * c["l_c"]
    rec[6] = d
    rec[12] = rho_p
    rec[13] = m_p
    rec[14] = i_p
    rec[15] = 1.0
    return rec


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return meta, {k: z[k] for k in z.files if k != "meta"}


MOHAMMAD_RATES = dict(q_mode=0, s0=1.0, s1=f32(1.4), s2=f32(1.4), s3=1.0, s5=1.0,
                      sq=f32(1.2))  # F6:1017-1018, F1:1044-1045


def golden_params(meta):
    cfg = f6_params(meta["nx"], meta["ny"], meta["non_newtonian"],
                    meta.get("local_rate", False),
                    float(meta["uxn"].replace("D", "e")) if meta.get("uxn") else 0.0)
    if meta.get("rates") == "mohammad":
        cfg.update(MOHAMMAD_RATES)
    if meta.get("rheology") == "casson":
        cfg["rheology"] = 2  # F6:941, F6:944
    if meta.get("boundary") == "zouhe":
        cfg["boundary"] = 1  # F1:1222-1417
        if meta.get("zouhe_uxn"):
            cfg["wall"][4] = float(meta["zouhe_uxn"].replace("D", "e"))
    if meta.get("boundary") == "equilibrium":
        cfg["boundary"] = 2  # F4:1205-1402
    return cfg


# Contact-model builds of the reference routine (oracle/build_ref.py):
# contact model and the constants each block sets (F4:625, F1:647, F4:588).
CONTACT_MODELS = {
    "f4_contact_wt": dict(contact=1, zeta_c=1.0, eps_p=1.0, eps_w=2.0),
    "f4_contact_lj": dict(contact=2, zeta_c=1.0),
    "f1_contact_wt": dict(contact=1, zeta_c=2.0, eps_p=1.0, eps_w=2.0),
}


def contact_model_cases(name):
    """(nx, ny, [(wall_edge, disks (n, 3: x y diam), expected F_COL (n, 2))])
    of one build in the contact_models fixture."""
    z = np.load(os.path.join(GOLDEN, "contact_models.npz"))
    nx, ny = (int(v) for v in z[name + "_nx"])
    out, off = [], 0
    for edges, n in zip(z[name + "_edges"], z[name + "_n"]):
        out.append((list(edges), z[name + "_disks"][off:off + n], z[name + "_fcol"][off:off + n]))
        off += n
    return nx, ny, out


# Contact constants and disk size of the whole-program builds (F3:739-742,
# F2:710-713; D = L_C, F3:22, F2:21) and whether the program integrates the
# angle (F2:331; F3:327 has it commented out).
NPARTICLE_PROGRAMS = {
    "F3": dict(diam=16.0, eps_p=2.0, eps_w=1.0, e_p=20.0, e_w=10.0, update_angle=0),
    "F2": dict(diam=20.0, eps_p=1.0, eps_w=1.0, e_p=20.0, e_w=20.0, update_angle=1),
}


def nparticle_setup(meta, pos):
    """Run parameters and disk records of a whole multi-particle program
    build (F3's or F2's module PARAM with the grid the build sets): the
    constants printed by the program itself (OMEGA, NU, A_G, RHO_S, RHO_F,
    PI), the rate set (the same ComputeRelaxation_MRT as F6), the program's
    Glowinski contact constants, and its disks (D = L_C, RHO_P = RHO_S,
    ZITA = 1, mass and inertia in the reference's operation order, F3:151-155,
    F2:166-172)."""
    prog = NPARTICLE_PROGRAMS[meta["src"]]
    d = prog["diam"]
    cfg = f6_params(meta["nx"], meta["ny"], False)
    # CIJ = -(RHO_S-RHO_F)*A_G*PI*0.5*D*0.5*D, left to right (F3:739, F2:710)
    cij = -((((((meta["rho_s"] - meta["rho_f"]) * meta["a_g"]) * meta["pi"]) * 0.5) * d) * 0.5) * d
    cfg.update(omega=meta["omega"], nu=meta["nu"], mu0=meta["nu"], tau_y=0.0,
               gy=meta["a_g"], update_angle=prog["update_angle"], zeta_c=1.0, cij=cij,
               eps_p=prog["eps_p"], eps_w=prog["eps_w"], e_p=prog["e_p"], e_w=prog["e_w"])
    n = len(pos)
    part = np.zeros((22, n))
    rho_p = meta["rho_s"]
    m_p = ((rho_p * meta["pi"]) * (0.5 * d)) * (0.5 * d)
    part[0], part[1] = pos[:, 0], pos[:, 1]
    part[6] = d
    part[12] = rho_p
    part[13] = m_p
    part[14] = ((0.5 * m_p) * (0.5 * d)) * (0.5 * d)
    part[15] = 1.0
    return cfg, part
