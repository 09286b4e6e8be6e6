"""CPU checks of the workload builders: the committed dense-sedimentation
positions are what the pinned generator draws, and they satisfy the
placement rules (gap 4 between disks and to the walls, upper 60 % of the
cavity); the reader of the reference's initial-condition files."""
import os

import numpy as np
import pytest

from lbmtest import pkg


def configs():
    import importlib
    return importlib.import_module("2d_nnf_lbm_amd.configs")


def test_dense_positions_file_is_the_seeded_draw():
    c = configs()
    ref = np.loadtxt(c.DENSE_FILE)
    assert ref.shape == (504, 2)
    assert np.array_equal(c.dense_positions(), ref)


def test_dense_positions_placement_rules():
    pts = np.loadtxt(configs().DENSE_FILE)
    d, gap, n = 16.0, 4.0, 2048
    diff = pts[:, None, :] - pts[None, :, :]
    dist = np.hypot(diff[..., 0], diff[..., 1]) + np.eye(len(pts)) * 1e9
    assert dist.min() >= d + gap
    assert pts[:, 0].min() >= 0.5 * d + gap and pts[:, 0].max() <= n - 0.5 * d - gap
    assert pts[:, 1].min() >= 0.4 * n + 0.5 * d + gap and pts[:, 1].max() <= n - 0.5 * d - gap


def test_workload_parameters():
    c = configs()
    cfg, part = c.one_disk_newtonian()
    assert (cfg["nx"], cfg["ny"]) == (200, 600)
    assert abs(1.0 / cfg["omega"] - 0.6) < 1e-15
    assert cfg["q_mode"] == 0 and cfg["sq"] == float(np.float32(1.2))
    assert part[pkg()._abi.PF["x"], 0] == 100.0 and part[pkg()._abi.PF["y"], 0] == 480.0
    cfg, part = c.dense_sedimentation(1.0)
    assert part.shape[1] == 504 and cfg["max_particles"] == 504
    # Bn = tau_y / (drho |g| D)
    assert abs(cfg["tau_y"] / (0.1 * -cfg["gy"] * 16.0) - 1.0) < 1e-12
    cfg0, _ = c.dense_sedimentation(0.0)
    assert cfg0["rheology"] == 0


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reads_reference_f3_initial_positions():
    """The reference's own 504-disk initial-condition file (F3's
    0_Particle_Initial_Position.txt, CRLF lines after a blank first line)
    read as READ(200,*) DISK(IP)%CENTER does (F3:145-149)."""
    c = configs()
    pf = pkg()._abi.PF
    part = c.read_initial_positions(GOLDEN + "/f3_initial_positions.txt", n=504, diam=16.0,
                                    rho_p=1.01)
    assert part.shape == (pkg()._abi.PF_COUNT, 504)
    assert part[pf["x"], :3].tolist() == [14.0, 32.0, 50.0]
    assert part[pf["y"], :3].tolist() == [498.0, 498.0, 498.0]
    assert part[pf["x"], -1] == 498.0 and part[pf["y"], -1] == 192.0
    # a square lattice of spacing 18 inside F3's 512 x 512 cavity
    x, y = part[pf["x"]], part[pf["y"]]
    dist = np.hypot(x[:, None] - x[None], y[:, None] - y[None]) + np.eye(504) * 1e9
    assert dist.min() == 18.0
    assert x.min() - 8.0 > 0 and x.max() + 8.0 < 512 and y.max() + 8.0 < 512
    # F3:152-155: DIAM = L_C, RHO_P = RHO_S, mass and inertia from them
    pi32 = float(np.float32(3.14159265359))
    m = 1.01 * pi32 * 8.0 * 8.0
    assert np.all(part[pf["mp"]] == m) and np.all(part[pf["ip"]] == 0.5 * m * 8.0 * 8.0)
    assert np.all(part[pf["zeta"]] == 1.0) and np.all(part[pf["vy"]] == 0.0)
    # reading more records than the file holds is an error
    with pytest.raises(ValueError):
        c.read_initial_positions(GOLDEN + "/f3_initial_positions.txt", n=505, diam=16.0,
                                 rho_p=1.01)


def test_reads_f1_layout_records(tmp_path):
    """F1:170-174: CENTER, DIAM, RHO_P per record; a record may continue on
    the next line and values after it on its last line are skipped, as in
    Fortran list-directed input."""
    c = configs()
    pf = pkg()._abi.PF
    f = tmp_path / "0_Particle_Initial_Position.txt"
    f.write_text("120\t480\t30\t1.25\t\n\n 60 100\n 20 2.0D0 99\n")
    part = c.read_initial_positions(str(f))
    assert part.shape[1] == 2
    assert part[[pf["x"], pf["y"], pf["diam"], pf["rhop"]], 0].tolist() == [120, 480, 30, 1.25]
    assert part[[pf["x"], pf["y"], pf["diam"], pf["rhop"]], 1].tolist() == [60, 100, 20, 2.0]
    f.write_text("1 2 3\n")
    with pytest.raises(ValueError):
        c.read_initial_positions(str(f))
