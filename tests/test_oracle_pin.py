"""Pins the CPU oracle to the reference program itself.

The fixtures under tests/golden/ were written by the compiled reference
(oracle/build_ref.py + tests/golden/make_golden.py): F6 run for 1000 steps
(300 for the moving-lid case) on small grids.  The oracle is run from the same
initial state for the same number of steps and every output must agree
bit-for-bit.
"""
import numpy as np
import pytest

import cases

VARIANTS = ["f6_newt", "f6_bing_lagged", "f6_bing_local", "f6_newt_lid", "f6_newt_mohammad",
            "f6_cas_local", "f6_zouhe_lid", "f6_eq_bing_local"]


def _run_oracle(orc, meta):
    cfg = cases.golden_params(meta)
    p = orc.make_params(cfg)
    st = orc.State(p, 1)
    st.initialize()
    st.set_particles(cases.f6_disk(meta["nx"], meta["ny"])[:, None])
    hist = st.run(0, meta["time_max"])
    return cfg, st, hist


@pytest.mark.parametrize("name", VARIANTS)
def test_constants_match_reference_printout(name):
    meta, _ = cases.load_golden(name)
    c = cases.f6_constants(meta["nx"], meta["ny"])
    k = meta["consts"]
    assert c["a_g"] == k["A_G"]
    assert c["tau_y"] == k["Tau_Y_LB"]
    assert c["u_c"] == k["U_LB_max"]
    assert 1.0 / c["omega"] == k["TAU_LB"]


@pytest.mark.parametrize("name", VARIANTS)
def test_oracle_bitwise_vs_reference(orc, name):
    meta, g = cases.load_golden(name)
    cfg, st, hist = _run_oracle(orc, meta)
    assert np.array_equal(st.f, g["f"])
    assert np.array_equal(st.rho, g["rho"])
    assert np.array_equal(st.ux, g["ux"])
    assert np.array_equal(st.uy, g["uy"])
    assert np.array_equal(st.g, g["g"])
    assert np.array_equal(st.omega_bn, g["omega_bn"])
    assert np.array_equal(st.image, g["image"])
    assert np.array_equal(st.stress, g["stress"])
    assert np.array_equal(st.s_mrt, g["omega_mrt"])
    part = st.part[:, 0]
    ref = g["particle"]  # CENTER(2), VELOCITY(3), FORCE(3)
    mine = np.array([part[orc.PF_X], part[orc.PF_Y], part[orc.PF_VX], part[orc.PF_VY],
                     part[orc.PF_W], part[orc.PF_FX], part[orc.PF_FY], part[orc.PF_TQ]])
    assert np.array_equal(mine, ref)


@pytest.mark.parametrize("name", VARIANTS)
def test_oracle_history_vs_reference(orc, name):
    """Per-step particle history line (F6:315-318), every step."""
    meta, g = cases.load_golden(name)
    cfg, st, hist = _run_oracle(orc, meta)
    c = cases.f6_constants(meta["nx"], meta["ny"])
    H = orc.H_PCOUNT
    ref = g["hist"]
    n = meta["time_max"]
    assert ref.shape[0] == n
    y, x, ang = hist[:, orc.H_Y], hist[:, orc.H_X], hist[:, orc.H_ANG]
    vx, vy, w = hist[:, orc.H_VX], hist[:, orc.H_VY], hist[:, orc.H_W]
    assert np.array_equal(ref[:, 1], y)
    assert np.array_equal(ref[:, 2], x)
    assert np.array_equal(ref[:, 3], ang)
    assert np.array_equal(ref[:, 4], vx)
    assert np.array_equal(ref[:, 5], vy / c["u_c"])
    assert np.array_equal(ref[:, 6], w)
    # Fx + F_H_IN_x: the line holds the previous step's hydrodynamic force
    fx_prev = np.concatenate([[0.0], hist[:-1, orc.H_FX]])
    assert np.array_equal(ref[:, 7], fx_prev + hist[:, orc.H_FHX])
    # mean density per step, F6:348
    nodes = (meta["nx"] + 1) * (meta["ny"] + 1)
    assert np.array_equal(g["avgden"], hist[:, H] / float(nodes))


def _contact_setup(orc, z, meta, n):
    """Oracle state holding the first n disks of the f3_contact fixture, with
    F3's contact constants (F3:739-742)."""
    d1 = float(z["diam"][0])
    # CIJ = -(RHO_S-RHO_F)*A_G*PI*0.5*D1*0.5*D1, evaluated left to right (F3:739)
    cij = -((((((meta["rho_s"] - meta["rho_f"]) * meta["a_g"]) * meta["pi"]) * 0.5) * d1) * 0.5) * d1
    cfg = cases.f6_params(meta["nx"], meta["ny"], False)
    cfg.update(zeta_c=1.0, cij=cij, eps_p=2.0, eps_w=1.0, e_p=20.0, e_w=10.0)
    st = orc.State(orc.make_params(cfg), n)
    part = np.zeros((orc.PF_COUNT, n))
    part[orc.PF_X] = z["x"][:n]
    part[orc.PF_Y] = z["y"][:n]
    part[orc.PF_DIAM] = z["diam"][:n]
    st.set_particles(part)
    return st


def test_oracle_contact_vs_reference_routine(orc):
    """N-body contact force (F3:632-806) against the reference routine itself,
    compiled from F3 and driven by oracle/f3_contact_driver.f90: three calls,
    the stale D_PW of the pair branch (F3:760) carried from call to call.  The
    second call sees particle 1's left-wall distance (> r_i + r_j, spring term
    only); the third sees the last particle's (<= r_i + r_j, with the E_P
    term).  Every F_COL must agree bit for bit."""
    meta, z = cases.load_golden("f3_contact")
    n = len(z["x"])
    first = _contact_setup(orc, z, meta, 1)
    first.contact()
    st = _contact_setup(orc, z, meta, n)
    st.c.dpw_stale = first.c.dpw_stale
    pairs, npairs = st.contact(max_pairs=4096)
    got2 = np.stack([st.part[orc.PF_FCX, :n], st.part[orc.PF_FCY, :n]], axis=1)
    st.contact()
    got3 = np.stack([st.part[orc.PF_FCX, :n], st.part[orc.PF_FCY, :n]], axis=1)
    assert np.array_equal(got2, z["fcol_call2"])
    assert np.array_equal(got3, z["fcol_call3"])
    # the two calls take different pair branches, and both have pairs
    assert npairs > 10
    assert not np.array_equal(got2, got3)
    # pair list: brute force over all i < j, d <= r_i + r_j + zeta
    x, y, d = z["x"], z["y"], z["diam"]
    brute = [(i, j) for i in range(n) for j in range(i + 1, n)
             if np.sqrt((x[i] - x[j]) ** 2 + (y[i] - y[j]) ** 2) <= 0.5 * d[i] + 0.5 * d[j] + 1.0]
    assert [tuple(p) for p in pairs] == brute


def _branches(name, nx, ny, cases_):
    """Which branches of the model the fixture's cases reach, computed from
    the inputs alone: pair overlap / gap, wall overlap / gap."""
    zeta = cases.CONTACT_MODELS[name]["zeta_c"]
    seen = set()
    for edges, disks, _ in cases_:
        if len(disks) == 2:
            (x1, y1, d1), (x2, y2, d2) = disks
            d, s = np.hypot(x1 - x2, y1 - y2), 0.5 * d1 + 0.5 * d2
            if d <= s + zeta:
                seen.add("pair overlap" if d < s else "pair gap")
        for x, y, dm in disks:
            if name.endswith("_lj"):
                dw = [y, nx - x, 2 * (ny - y), 2 * x]
            else:
                dw = [2 * abs(y - edges[0]), 2 * abs(edges[1] - x), 2 * abs(edges[2] - y),
                      2 * abs(x - edges[3])]
            for v in dw:
                if v <= dm + zeta:
                    seen.add("wall overlap" if v < dm else "wall gap")
    return seen


@pytest.mark.parametrize("name", sorted(cases.CONTACT_MODELS))
def test_oracle_contact_models_vs_reference_routine(orc, name):
    """Wan-Turek (F4:624-677 with zeta 1; F1:646-699 with zeta 2) and
    Lennard-Jones (F4:587-622) contact forces against the reference routine
    itself, compiled from F4 / F1 with that model switched on and driven by
    oracle/contact_model_driver.f90: every F_COL bit for bit, over cases that
    reach every branch of the model."""
    nx, ny, cs = cases.contact_model_cases(name)
    cfg = dict(cases.f6_params(nx, ny, False), **cases.CONTACT_MODELS[name])
    states = {}
    for c, (edges, disks, want) in enumerate(cs):
        n = len(disks)
        if n not in states:
            states[n] = orc.State(orc.make_params(cfg), n)
        st = states[n]
        for k in range(4):
            st.p.wall_edge[k] = edges[k]
        part = np.zeros((orc.PF_COUNT, n))
        part[orc.PF_X], part[orc.PF_Y], part[orc.PF_DIAM] = disks.T
        st.set_particles(part)
        st.contact()
        got = np.stack([st.part[orc.PF_FCX, :n], st.part[orc.PF_FCY, :n]], axis=1)
        assert np.array_equal(got, want), (c, got, want)
    seen = _branches(name, nx, ny, cs)
    want_b = {"pair overlap", "pair gap", "wall overlap", "wall gap"}
    if name.endswith("_lj"):  # one formula inside the cut-off
        want_b = {"pair overlap", "wall overlap"}
    assert want_b <= seen, seen


def test_oracle_contact_model_bites(orc):
    """The fixture discriminates between models: the Glowinski terms
    (contact 0) on the Wan-Turek cases' inputs miss most of them."""
    nx, ny, cs = cases.contact_model_cases("f4_contact_wt")
    cfg = dict(cases.f6_params(nx, ny, False), **cases.CONTACT_MODELS["f4_contact_wt"])
    cfg["contact"] = 0  # Glowinski terms on the same inputs
    differ = 0
    for edges, disks, want in cs:
        n = len(disks)
        st = orc.State(orc.make_params(dict(cfg, wall_edge=edges)), n)
        part = np.zeros((orc.PF_COUNT, n))
        part[orc.PF_X], part[orc.PF_Y], part[orc.PF_DIAM] = disks.T
        st.set_particles(part)
        st.contact()
        got = np.stack([st.part[orc.PF_FCX, :n], st.part[orc.PF_FCY, :n]], axis=1)
        differ += not np.array_equal(got, want)
    assert differ > len(cs) // 2


def test_oracle_prescribed_oscillation_closed_form(orc):
    """Prescribed motion of F4:221-226: V_x(KK) = -U cos(2 pi KK / T), centre
    advanced by the mean of the old and new velocity, and the added-mass
    term (rho0/rho_p) m (V - V_old) of F4:256-258.  Checked against the
    integral of the velocity law: after n updates from rest,
    X_n - X_0 = -U T/(2 pi) sin(2 pi n/T) + U/2 (the first half-step starts
    from V = 0 instead of -U) up to the trapezoid error, which is below
    n U (2 pi/T)^2 / 12.  pi is F4's default-REAL literal (float32)."""
    cfg = cases.f6_params(40, 40, False)
    u, period = 0.003, 200.0
    cfg.update(motion=1, osc_u=u, osc_period=period)
    st = orc.State(orc.make_params(cfg), 1)
    rec = np.zeros(orc.PF_COUNT)
    rec[orc.PF_X], rec[orc.PF_Y], rec[orc.PF_DIAM] = 20.0, 20.0, 5.0
    rec[orc.PF_RHOP], rec[orc.PF_MP], rec[orc.PF_IP], rec[orc.PF_ZETA] = 1.5, 29.45, 92.0, 1.0
    st.set_particles(rec[:, None])
    n = 150
    xs, fh = [], []
    for kk in range(1, n + 1):
        st.call("orc_update", kk)
        xs.append(st.part[orc.PF_X, 0])
        fh.append(st.part[orc.PF_FHX, 0])
    k = np.arange(1, n + 1)
    pi = cases.PI
    x_exact = 20.0 - u * period / (2 * pi) * np.sin(2 * pi * k / period) + 0.5 * u
    assert np.max(np.abs(np.array(xs) - x_exact)) <= n * u * (2 * pi / period) ** 2 / 12
    v = -u * np.cos(2 * pi * k / period)
    dv = np.diff(np.concatenate([[0.0], v]))
    assert np.allclose(fh, (1.0 / 1.5) * 29.45 * dv, rtol=1e-12, atol=1e-18)
    assert st.part[orc.PF_Y, 0] == 20.0 and st.part[orc.PF_VY, 0] == 0.0


def nparticle_oracle(orc, name):
    """The oracle set up as a whole multi-particle program build: its
    parameters and disks, and the stale wall distance of the contact force at
    0, the value the compiled program's saved D_PW holds before its first
    call (static storage starts at zero; F3:760 reads it before F3:773-804
    assign it)."""
    meta, g = cases.load_golden(name)
    cfg, part = cases.nparticle_setup(meta, g["pos"])
    st = orc.State(orc.make_params(cfg), part.shape[1])
    st.initialize()
    st.set_particles(part)
    st.c.dpw_stale = 0.0
    return meta, g, cfg, part, st


HCOLS = ("H_X", "H_Y", "H_VX", "H_VY", "H_W", "H_FX", "H_FY", "H_TQ", "H_ANG")


def per_disk_history(orc, hist, n):
    """Oracle history rows as [step, disk, (x y vx vy w fx fy tq angle)], the
    layout of the fixtures' ref_history records."""
    H = orc.H_PCOUNT
    cols = [getattr(orc, c) for c in HCOLS]
    return np.stack([hist[:, [p * H + c for c in cols]] for p in range(n)], axis=1)


@pytest.mark.parametrize("name", ["f3_multi", "f2_dkt"])
def test_oracle_multi_particle_step_vs_reference_program(orc, name):
    """Whole multi-particle reference programs against the oracle, bit for
    bit (oracle/build_ref.py patch_nparticle): F3, 12 disks in a 160 x 160
    cavity, 300 steps; F2, two disks starting within contact range in a
    120 x 240 channel, 400 steps.  Every disk's CENTER, VELOCITY, FORCE and
    ANGLE at every step (the update loop over all disks, F3:312-328 without
    and F2:311-332 with the angle; the multi-disk hydrodynamic sums,
    F6:534-558; F2's assigned pair force, F2:715-744, which for two disks is
    the N-body sum's result), and the final F, RHO, UX, UY and G (with
    overlapping smoothed profiles, where the highest disk index sets G,
    F6:487-494)."""
    meta, g, cfg, part, st = nparticle_oracle(orc, name)
    hist = st.run(0, meta["time_max"])
    mine = per_disk_history(orc, hist, part.shape[1])
    assert mine.shape == g["hist"].shape
    assert np.array_equal(mine, g["hist"])
    assert np.array_equal(st.f, g["f"])
    assert np.array_equal(st.rho, g["rho"])
    assert np.array_equal(st.ux, g["ux"])
    assert np.array_equal(st.uy, g["uy"])
    assert np.array_equal(st.g, g["g"])


@pytest.mark.parametrize("name", ["f3_multi", "f2_dkt"])
def test_multi_particle_fixtures_reach_the_rules(orc, name):
    """The fixtures are ones where the rules in question decide results:
    nodes in the smoothed profiles (phi > 0) of two disks at once, pairs
    within contact range, every disk moving; F3's also has a disk within
    wall range."""
    meta, g, cfg, part, st = nparticle_oracle(orc, name)
    x, y, d = part[orc.PF_X], part[orc.PF_Y], part[orc.PF_DIAM]
    yy, xx = np.mgrid[0:meta["ny"] + 1, 0:meta["nx"] + 1]
    reach = [np.hypot(xx - x[p], yy - y[p]) < 0.5 * d[p] + 0.5 for p in range(len(x))]
    assert np.any(np.sum(reach, axis=0) >= 2)
    pairs, npairs = st.contact(max_pairs=64)
    assert npairs >= (4 if name == "f3_multi" else 1)
    if name == "f3_multi":
        assert np.any(part[orc.PF_Y] - 0.5 * part[orc.PF_DIAM] < 1.0)
    h = g["hist"]
    assert np.all(h[-1, :, 1] < h[0, :, 1])  # every disk falls
