"""Failure handling of the fused step, forced from the host: a non-finite
value put into the state through the C ABI makes the next lbm_step return
LBM_ERR_NAN (the kernels run to their end normally and report through the
handle's error bits), and the handle keeps working after the error.  The
reference's only check is tau < 0.5 (F6:115-118), which lbm_create
reports as LBM_ERR_PARAM."""
import numpy as np
import pytest

from lbmtest import f6_case, gpu_solver, pkg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_nan_population_reported(storage):
    print("case: NaN population", storage, flush=True)
    lbm = pkg()
    cfg, part = f6_case(40, 120, True)
    g = gpu_solver(cfg, storage=storage)
    g.initialize()
    g.set_particles(part)
    g.step(0, 5, record=False)
    f = g.f
    f[60, 20, 3] = np.nan
    g.f = f
    with pytest.raises(lbm.LBMError) as e:
        g.step(5, 4, record=True)
    assert e.value.code == lbm._abi.LBM_ERR_NAN
    # the error bits were cleared: the handle runs on from a fresh state
    g.initialize()
    g.set_particles(part)
    h = g.step(0, 5)
    assert np.all(np.isfinite(h))
    g.close()


@pytest.mark.parametrize("storage", ["fp64", "bf16"])
def test_nan_particle_state_reported(storage):
    print("case: NaN particle velocity", storage, flush=True)
    lbm = pkg()
    cfg, part = f6_case(40, 120, True)
    g = gpu_solver(cfg, storage=storage)
    g.initialize()
    bad = part.copy()
    bad[lbm._abi.PF["vy"], 0] = np.nan
    g.set_particles(bad)
    with pytest.raises(lbm.LBMError) as e:
        g.step(0, 3, record=False)
    assert e.value.code == lbm._abi.LBM_ERR_NAN
    g.close()

