"""Batched ensemble: many grids of different sizes and parameters in one
launch per step.

* Each simulation of a batch gives bit-for-bit the result of the same
  simulation run alone through lbm_step (FP64 and FP16 storage).
* Steps replayed from the captured hipGraph equal eager launches, checked
  after every call (each call replays the graph twice), with host work and
  synchronisation between the calls.
* The tiling chosen per simulation (tiles per block) and the autotuner's
  choice leave the results unchanged.
* The prescribed oscillation (F4:221-226) in FP64 matches the oracle.
* The 256-cavity ensemble (about 1.04e8 nodes per step) runs in FP16.
"""
import importlib

import numpy as np
import pytest

from lbmtest import oracle_state, rel_err

pytestmark = pytest.mark.gpu


def pkg():
    return importlib.import_module("2d_nnf_lbm_amd")


def configs():
    return importlib.import_module("2d_nnf_lbm_amd.configs")


CASES = [(40, 2.0), (56, 7.5), (33, 0.5), (64, 4.0)]


def small_batch(storage):
    return [configs().oscillating_cavity(side, bn, storage) for side, bn in CASES]


def make_batch(cases):
    b = pkg().Batch([c for c, _ in cases])
    b.initialize()
    for i, (_, part) in enumerate(cases):
        b.set_particles(i, part)
    return b


def state(sim):
    return sim.f, sim.get_particles()


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_batch_equals_single_simulations(storage):
    cases = small_batch(storage)
    n = 30
    print("case: batch of %d, %s, %d steps" % (len(cases), storage, n), flush=True)
    b = make_batch(cases)
    hb = b.step(0, n)
    hb2 = b.step(n, 7)
    for i, (cfg, part) in enumerate(cases):
        s = pkg().D2Q9(cfg)
        s.initialize()
        s.set_particles(part)
        hs = s.step(0, n)
        hs2 = s.step(n, 7)
        assert np.array_equal(hb[i], hs), i
        assert np.array_equal(hb2[i], hs2), i
        fb, pb = state(b.sims[i])
        fs, ps = state(s)
        assert np.array_equal(fb, fs), i
        assert np.array_equal(pb, ps), i
        s.close()
    b.close()


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_batch_mixed_boundaries_equal_single_simulations(storage):
    """One batch whose simulations use the three boundary rules (bounce-back,
    Zou-He, equilibrium): each equals its run alone, bit for bit."""
    cases = small_batch(storage)
    for (cfg, _), bnd in zip(cases, (0, 1, 2, 2)):
        cfg["boundary"] = bnd
    n = 30
    print("case: mixed-boundary batch, %s, %d steps" % (storage, n), flush=True)
    b = make_batch(cases)
    hb = b.step(0, n)
    for i, (cfg, part) in enumerate(cases):
        s = pkg().D2Q9(cfg)
        s.initialize()
        s.set_particles(part)
        hs = s.step(0, n)
        assert np.array_equal(hb[i], hs), i
        fb, pb = state(b.sims[i])
        fs, ps = state(s)
        assert np.array_equal(fb, fs), i
        assert np.array_equal(pb, ps), i
        s.close()
    b.close()


def test_batch_mixed_contact_models_equal_single_simulations():
    """One FP64 batch whose simulations use the three contact models
    (Glowinski, Wan-Turek with the cavity as wall box, Lennard-Jones), with
    a second disk next to the first so the pair term acts: each equals its
    run alone, bit for bit."""
    cases = small_batch("fp64")
    models = (0, 1, 2, 1)
    for (cfg, part), model in zip(cases, models):
        side = cfg["nx"]
        cfg["contact"] = model
        cfg["max_particles"] = 2
        if model == 1:
            cfg.update(zeta_c=1.0, eps_p=1.0, eps_w=2.0, wall_edge=[0.0, side, side, 0.0])
        elif model == 2:
            cfg["zeta_c"] = 1.0
    two = []
    for cfg, part in cases:
        b = part[:, 0].copy()
        b[0] = part[0, 0] + part[6, 0] + 0.5  # just outside contact, inside range
        two.append((cfg, np.stack([part[:, 0], b], axis=1)))
    n = 25
    print("case: mixed-contact batch, %d steps" % n, flush=True)
    bt = make_batch(two)
    hb = bt.step(0, n)
    for i, (cfg, part) in enumerate(two):
        s = pkg().D2Q9(cfg)
        s.initialize()
        s.set_particles(part)
        hs = s.step(0, n)
        assert np.array_equal(hb[i], hs), i
        fb, pb = state(bt.sims[i])
        fs, ps = state(s)
        assert np.array_equal(fb, fs), i
        assert np.array_equal(pb, ps), i
        s.close()
    bt.close()


def test_batch_graph_replays_equal_eager_launches():
    cases = small_batch("fp16")
    eager = make_batch(cases)
    graph = make_batch(cases)
    graph.set_graph(4)
    kk = 0
    for call in range(3):
        n = 10  # steps 1..8 from two replays of the 4-step graph
        print("case: graph call %d" % call, flush=True)
        he = eager.step(kk, n)
        hg = graph.step(kk, n)
        kk += n
        for i in range(len(cases)):
            assert np.array_equal(he[i], hg[i]), (call, i)
            fe, pe = state(eager.sims[i])
            fg, pg = state(graph.sims[i])
            assert np.array_equal(fe, fg), (call, i)
            assert np.array_equal(pe, pg), (call, i)
    eager.close()
    graph.close()


def test_batch_tiling_and_tuner_leave_results_unchanged():
    cases = small_batch("fp16")
    ref = make_batch(cases)
    ref.step(0, 12, record=False)
    tiled = make_batch(cases)
    tiled.tiles = [1, 2, 4, 8]
    assert list(tiled.tiles) == [1, 2, 4, 8]
    tiled.step(0, 12, record=False)
    for i in range(len(cases)):
        assert np.array_equal(state(ref.sims[i])[0], state(tiled.sims[i])[0]), i
    tuned = pkg().Batch([c for c, _ in cases])
    choice = tuned.tune(steps=4)
    print("tuner choice", list(choice), flush=True)
    assert set(choice) <= {1, 2, 4, 8}
    assert list(tuned.tiles) == list(choice)
    tuned.initialize()
    for i, (_, part) in enumerate(cases):
        tuned.set_particles(i, part)
    tuned.step(0, 12, record=False)
    for i in range(len(cases)):
        assert np.array_equal(state(ref.sims[i])[0], state(tuned.sims[i])[0]), i
        assert np.array_equal(state(ref.sims[i])[1], state(tuned.sims[i])[1]), i
    for b in (ref, tiled, tuned):
        b.close()


def test_prescribed_oscillation_matches_oracle(orc):
    cfg, part = configs().oscillating_cavity(48, 3.0, "fp64")
    n = 300
    st = oracle_state(orc, cfg, part)
    ho = st.run(0, n)
    g = pkg().D2Q9(dict(cfg, storage="fp64"))
    g.initialize()
    g.set_particles(part)
    hg = g.step(0, n)
    assert rel_err(g.f, st.f) <= 1e-12
    for col in (orc.H_X, orc.H_VX, orc.H_FHX):
        assert rel_err(hg[:, col], ho[:, col]) <= 1e-9, col
    assert rel_err(hg[:, orc.H_FX], ho[:, orc.H_FX]) <= 1e-9
    g.close()


def test_cavity_ensemble_256():
    ens = configs().cavity_ensemble()
    nodes = sum(configs().nodes(c) for c, _ in ens)
    print("case: 256 cavities, %.3g nodes, fp16" % nodes, flush=True)
    b = make_batch(ens)
    b.set_graph(2)
    n = 6
    hist = b.step(0, n)
    u = np.array([c["osc_u"] for c, _ in ens])
    t = ens[0][0]["osc_period"]
    k = np.arange(1, n + 1)
    H = pkg()._abi.H
    for i, (cfg, part) in enumerate(ens):
        vx = hist[i][:, H["vx"]]
        # the record of step k holds the particle state of loop iteration k
        expect = -u[i] * np.cos(2 * configs().PI * k / t)
        assert np.allclose(vx, expect, rtol=1e-12, atol=1e-15), i
        assert np.all(np.isfinite(hist[i]))
        m = hist[i][:, -2]
        assert abs(m[-1] / configs().nodes(cfg) - 1.0) < 1e-4, i
    b.close()
