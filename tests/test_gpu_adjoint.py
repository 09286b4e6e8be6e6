"""Adjoint of the settling velocity (lbm_adjoint_velocity).

* FP64 adjoint gradient d(mean V_y)/d(tau_y, mu0, rho_p) against central
  finite differences (h = 1e-4 relative) of the FP64 forward run, on a
  reduced channel and on the full 513 x 1025 node channel over 4096 steps:
  within 1e-4 relative.
* The FP32-stored run's gradient against the FP64 one.
* The checkpoint interval changes memory, not the result.
* torch.autograd.Function returns the same gradient.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def pkg():
    return importlib.import_module("2d_nnf_lbm_amd")


def configs():
    return importlib.import_module("2d_nnf_lbm_amd.configs")


SMALL = dict(nx=96, ny=192, diam=16.0)


def forward_loss(theta, nsteps, window, storage="fp64", **shape):
    cfg, part = configs().settling_adjoint(storage, *theta, **shape)
    s = pkg().D2Q9(cfg)
    s.initialize()
    s.set_particles(part)
    hist = s.step(0, nsteps)
    s.close()
    return float(np.mean(hist[-window:, pkg()._abi.H["vy"]]))


def adjoint(theta, nsteps, window, chk, storage="fp64", **shape):
    cfg, part = configs().settling_adjoint(storage, *theta, **shape)
    s = pkg().D2Q9(cfg)
    s.initialize()
    s.set_particles(part)
    loss, grad, mem = s.adjoint_velocity(nsteps, window, chk)
    s.close()
    return loss, grad, mem


def theta0(**shape):
    cfg, part = configs().settling_adjoint("fp64", **shape)
    return [cfg["tau_y"], cfg["mu0"], float(part[pkg()._abi.PF["rhop"], 0])]


def central_differences(theta, nsteps, window, **shape):
    g = np.zeros(3)
    for i in range(3):
        h = 1e-4 * theta[i]
        tp, tm = list(theta), list(theta)
        tp[i] += h
        tm[i] -= h
        g[i] = (forward_loss(tp, nsteps, window, **shape) -
                forward_loss(tm, nsteps, window, **shape)) / (2 * h)
    return g


def check_against_fd(theta, nsteps, window, chk, **shape):
    loss, grad, mem = adjoint(theta, nsteps, window, chk, **shape)
    fd = central_differences(theta, nsteps, window, **shape)
    print("loss", loss, "adjoint", grad, "fd", fd, "mem %.3g" % mem, flush=True)
    assert loss == pytest.approx(forward_loss(theta, nsteps, window, **shape), rel=1e-12)
    for i in range(3):
        assert abs(grad[i] - fd[i]) <= 1e-4 * abs(fd[i]), (i, grad[i], fd[i])
    return grad


def test_adjoint_fp64_matches_finite_differences_small():
    theta = theta0(**SMALL)
    check_against_fd(theta, 400, 100, 32, **SMALL)


def test_checkpoint_interval_changes_memory_not_gradient():
    theta = theta0(**SMALL)
    l1, g1, m1 = adjoint(theta, 300, 50, 7, **SMALL)
    l2, g2, m2 = adjoint(theta, 300, 50, 50, **SMALL)
    l3, g3, m3 = adjoint(theta, 300, 50, 300, **SMALL)
    assert l1 == l2 == l3
    assert np.array_equal(g1, g2) and np.array_equal(g1, g3)
    print("memory", m1, m2, m3, flush=True)
    assert m2 < m3


def test_adjoint_fp32_storage_close_to_fp64():
    theta = theta0(**SMALL)
    _, g64, _ = adjoint(theta, 400, 100, 32, **SMALL)
    _, g32, _ = adjoint(theta, 400, 100, 32, storage="fp32", **SMALL)
    rel = np.abs(g32 - g64) / np.abs(g64)
    print("fp32 vs fp64", g32, g64, rel, flush=True)
    assert np.all(rel <= 1e-4)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    ag = importlib.import_module("2d_nnf_lbm_amd.autograd")
    theta = theta0(**SMALL)

    def build(tau_y, mu0, rho_p):
        return configs().settling_adjoint("fp64", tau_y, mu0, rho_p, **SMALL)

    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    loss = ag.settling_velocity(t, build, 200, 50, 25)
    (2.0 * loss).backward()
    _, g, _ = adjoint(theta, 200, 50, 25, **SMALL)
    assert np.allclose(t.grad.numpy(), 2.0 * g, rtol=1e-14, atol=0)


def test_adjoint_full_channel_4096_steps():
    """The sensitivity study at its real shape: 513 x 1025 nodes, 4096
    steps, checkpoints every 64 steps, loss over the last 512 steps."""
    theta = theta0()
    n, w, k = 4096, 512, 64
    g64 = check_against_fd(theta, n, w, k)
    _, g32, mem32 = adjoint(theta, n, w, k, storage="fp32")
    rel = np.abs(g32 - g64) / np.abs(g64)
    print("full channel fp32", g32, "rel to fp64", rel, "mem %.3g" % mem32, flush=True)
    assert np.all(rel <= 1e-4)
