"""Many-particle runs (more particles than the per-node loop handles, so the
fused step uses per-tile candidate lists): FP64 fused path against the
oracle, including contact forces and the pair list of the last contact
evaluation."""
import numpy as np
import pytest

import cases
from lbmtest import disk_record, gpu_solver, oracle_state, rel_err
from test_gpu_parity import check_field_errors, field_errors

pytestmark = pytest.mark.gpu


def _packed_disks(rng, nx, ny, n, d, gap_min, close_pairs):
    recs = []
    while len(recs) < n:
        if close_pairs and len(recs) % 2 == 1 and len(recs) // 2 < close_pairs:
            # partner placed 0.5 lu from contact with the previous disk
            x0, y0 = recs[-1][0], recs[-1][1]
            ang = rng.uniform(0, 2 * np.pi)
            x, y = x0 + (d + 0.5) * np.cos(ang), y0 + (d + 0.5) * np.sin(ang)
        else:
            x, y = rng.uniform(d, nx - d), rng.uniform(d, ny - d)
        if not (d <= x <= nx - d and d <= y <= ny - d):
            continue
        ok = all((x - r[0]) ** 2 + (y - r[1]) ** 2 >= (d + gap_min) ** 2 for r in recs)
        if ok:
            recs.append(disk_record(x, y, d, 1.5, zeta=1.5))
    return np.stack(recs, axis=1)


@pytest.mark.parametrize("bingham", [False, True])
def test_many_particles_fused_matches_oracle(orc, bingham):
    rng = np.random.default_rng(11)
    nx, ny, n, steps = 200, 200, 40, 300
    cfg = cases.f6_params(nx, ny, bingham, local_rate=bingham)
    cfg["gy"] = -2e-5
    d = 10.0
    cfg["cij"] = -((((((1.5 - 1.0) * cfg["gy"]) * cases.PI) * 0.5) * d) * 0.5) * d
    part = _packed_disks(rng, nx, ny, n, d, 0.3, close_pairs=8)
    print("case: many particles n=%d bingham=%s" % (n, bingham), flush=True)
    st = oracle_state(orc, cfg, part)
    ho = st.run(0, steps)
    pairs_o, npo = st.contact(max_pairs=4096)   # the contact the next step would do
    g = gpu_solver(cfg, max_particles=64)
    g.initialize()
    g.set_particles(part)
    hg = g.step(0, steps)
    rho, ux, uy, gg, om = g.get_fields()
    ref = dict(f=st.f, rho=st.rho, ux=st.ux, uy=st.uy, g=st.g, omega_bn=st.omega_bn)
    check_field_errors(field_errors(g.f, rho, ux, uy, gg, om, ref), "many")
    H = orc.H_PCOUNT
    for fld in (orc.H_X, orc.H_Y, orc.H_VX, orc.H_VY):
        cols = [p * H + fld for p in range(n)]
        e = rel_err(hg[:, cols], ho[:, cols])
        assert e <= 1e-9, (fld, e)
    # pair list found by the fused step's last contact evaluation (state
    # steps-1) equals the oracle's brute-force loop on the same state
    st2 = oracle_state(orc, cfg, part)
    st2.run(0, steps - 1, record=False)
    pairs_ref, nref = st2.contact(max_pairs=4096)
    pairs_g = g.last_pairs()
    print("pairs: ref %d gpu %d" % (nref, len(pairs_g)), flush=True)
    assert np.array_equal(pairs_g, pairs_ref)


# Reduced-precision storage on the many-particle path (FP32 candidate lists
# and profile records, contact forces on staged FP64 inputs) against the FP64
# oracle: the same 40 disks, Bingham fluid.  Bounds on every disk's settling
# velocity and position after the run, relative to the largest of them.
MANY_TOL = {"fp16": 1e-3, "bf16": 2e-3}


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_many_particles_reduced_storage_near_oracle(orc, storage):
    rng = np.random.default_rng(11)
    nx, ny, n, steps = 200, 200, 40, 1500
    cfg = cases.f6_params(nx, ny, True, local_rate=True)
    cfg["gy"] = -2e-5
    d = 10.0
    cfg["cij"] = -((((((1.5 - 1.0) * cfg["gy"]) * cases.PI) * 0.5) * d) * 0.5) * d
    part = _packed_disks(rng, nx, ny, n, d, 0.3, close_pairs=8)
    print("case: many particles n=%d %s vs FP64 oracle, %d steps" % (n, storage, steps),
          flush=True)
    st = oracle_state(orc, cfg, part)
    ho = st.run(0, steps)
    g = gpu_solver(cfg, storage=storage, max_particles=64)
    g.initialize()
    g.set_particles(part)
    hg = g.step(0, steps)
    assert np.all(np.isfinite(hg))
    H = orc.H_PCOUNT
    vy = [p * H + orc.H_VY for p in range(n)]
    ys = [p * H + orc.H_Y for p in range(n)]
    e_vy = rel_err(hg[-1, vy], ho[-1, vy])
    e_y = np.max(np.abs(hg[-1, ys] - ho[-1, ys]) / np.abs(ho[-1, ys]))
    e_mass = rel_err(hg[:, -2], ho[:, -2])
    print("%s: vy %.3e  y %.3e  mass %.3e  (max |vy| %.3e)"
          % (storage, e_vy, e_y, e_mass, np.max(np.abs(ho[-1, vy]))), flush=True)
    assert e_vy <= MANY_TOL[storage]
    assert e_y <= MANY_TOL[storage]
    assert e_mass <= 1e-4


def test_hundred_particles_fp64_matches_oracle(orc):
    """100 disks: the contact force of each disk by one block over inputs
    staged in LDS (65 to 128 disks), FP64 fused path against the oracle."""
    rng = np.random.default_rng(5)
    nx, ny, n, steps = 240, 240, 100, 200
    cfg = cases.f6_params(nx, ny, False)
    cfg["gy"] = -2e-5
    d = 10.0
    cfg["cij"] = -((((((1.5 - 1.0) * cfg["gy"]) * cases.PI) * 0.5) * d) * 0.5) * d
    part = _packed_disks(rng, nx, ny, n, d, 0.3, close_pairs=20)
    print("case: many particles n=%d FP64, %d steps" % (n, steps), flush=True)
    st = oracle_state(orc, cfg, part)
    ho = st.run(0, steps)
    g = gpu_solver(cfg, max_particles=128)
    g.initialize()
    g.set_particles(part)
    hg = g.step(0, steps)
    assert rel_err(g.f, st.f) <= 1e-12
    H = orc.H_PCOUNT
    for fld in (orc.H_X, orc.H_Y, orc.H_VX, orc.H_VY):
        cols = [p * H + fld for p in range(n)]
        e = rel_err(hg[:, cols], ho[:, cols])
        assert e <= 1e-9, (fld, e)
    st2 = oracle_state(orc, cfg, part)
    st2.run(0, steps - 1, record=False)
    pairs_ref, nref = st2.contact(max_pairs=4096)
    pairs_g = g.last_pairs()
    print("pairs: ref %d gpu %d" % (nref, len(pairs_g)), flush=True)
    assert nref > 0
    assert np.array_equal(pairs_g, pairs_ref)


def set_stale_wall_distance(g, value):
    """Set the contact force's stale wall distance (the D_PW the pair branch
    reads, F3:760) through a snapshot: its last two doubles are that distance
    and the module-global rate."""
    snap = g.snapshot()
    snap[-16:-8] = np.frombuffer(np.float64(value).tobytes(), dtype=np.uint8)
    g.load_snapshot(snap)


@pytest.mark.parametrize("name", ["f3_multi", "f2_dkt"])
def test_fused_fp64_matches_multi_particle_reference_program(name):
    """The fused FP64 path against whole multi-particle reference programs
    (tests/golden/f3_multi.npz: F3, 12 disks, 300 steps; f2_dkt.npz: F2, two
    disks in contact range, 400 steps), started like the compiled program
    (stale wall distance 0): every disk's position, velocity and angle at
    every step within 1e-9 relative, its force and torque within 1e-8, the
    final populations and fields within the parity bounds."""
    meta, gold = cases.load_golden(name)
    cfg, part = cases.nparticle_setup(meta, gold["pos"])
    n, steps = part.shape[1], meta["time_max"]
    print("case: %s, %d disks, %d steps, fused FP64" % (name, n, steps), flush=True)
    g = gpu_solver(cfg, max_particles=16)
    g.initialize()
    g.set_particles(part)
    set_stale_wall_distance(g, 0.0)
    hg = g.step(0, steps)
    H = 12
    # history columns in the fixture's order: x y vx vy w fx fy tq angle
    cols = [0, 1, 3, 4, 5, 6, 7, 8, 2]
    mine = np.stack([hg[:, [p * H + c for c in cols]] for p in range(n)], axis=1)
    ref = gold["hist"]
    for k, label in enumerate(("x", "y", "vx", "vy", "w", "fx", "fy", "tq", "angle")):
        e = rel_err(mine[:, :, k], ref[:, :, k])
        tol = 1e-8 if label in ("fx", "fy", "tq") else 1e-9
        assert e <= tol, (label, e)
    rho, ux, uy, gg, om = g.get_fields()
    ref_f = dict(f=gold["f"], rho=gold["rho"], ux=gold["ux"], uy=gold["uy"], g=gold["g"],
                 omega_bn=om)
    check_field_errors(field_errors(g.f, rho, ux, uy, gg, om, ref_f), name)
    g.close()
