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
