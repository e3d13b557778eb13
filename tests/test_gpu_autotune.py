"""The launch-shape autotuner: every candidate shape gives bit-identical
results (checked inside tune_solver), the solver's arrays are restored, and
a run with the tuned shapes equals a run with the built-in ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def solver(n=64, precision="mixed"):
    from warpflow import FlowSolver, build_topology
    from warpflow.config import CaseConfig, validate
    c = CaseConfig()
    c.grid.nx = c.grid.ny = c.grid.nz = n
    c.numerics.mg_levels = 4
    c.numerics.mg_precision = precision
    c.time.variable_dt = False
    c.time.dt_init = 2e-4
    validate(c)
    return FlowSolver(c, build_topology(c), "hip:0")


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_tuned_shapes_change_no_result(gpu, precision):
    from warpflow import autotune
    autotune.apply({k: 0 for k in autotune.KNOBS})
    try:
        a = solver(precision=precision)
        a.step()
        keep = {nm: a.f[nm].numpy() for nm in ("u", "ustar", "p")}
        phi = a.mg.levels[0].phi.numpy()
        res = autotune.tune_solver(a, reps=2)
        assert set(res["settings"]) == {"cd_ich", "sweep_tile", "sweep_ich"}
        assert autotune.settings() == res["settings"]
        for nm, v in keep.items():
            assert np.array_equal(a.f[nm].numpy(), v), nm
        assert np.array_equal(a.mg.levels[0].phi.numpy(), phi)
        a.step()
        autotune.apply({k: 0 for k in autotune.KNOBS})
        b = solver(precision=precision)
        b.step()
        b.step()
        for nm in ("u", "v", "w", "p"):
            assert np.array_equal(a.f[nm].numpy(), b.f[nm].numpy()), nm
    finally:
        autotune.apply({k: 0 for k in autotune.KNOBS})


def test_knob_validation(gpu):
    from warpflow import _hip, autotune
    with pytest.raises(_hip.HipError):
        autotune.set("sweep_tile", 3)
    autotune.set("cd_ich", 24)
    assert autotune.get("cd_ich") == 24
    autotune.set("cd_ich", 0)
