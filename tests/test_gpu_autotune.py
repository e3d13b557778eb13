"""The launch-shape autotuner: every candidate shape gives bit-identical
results (checked inside tune_solver), entries apply only to the shape they
were tuned on, the solver's arrays are restored, and a run with the tuned
shapes equals a run with the built-in ones."""
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
    autotune.reset()
    try:
        a = solver(precision=precision)
        a.step()
        keep = {nm: a.f[nm].numpy() for nm in ("u", "ustar", "p")}
        phi = a.mg.levels[0].phi.numpy()
        res = autotune.tune_solver(a, reps=2)
        knobs = {k for k, _, _ in res["entries"]}
        assert autotune.CD_ICH in knobs and autotune.PROJ_ICH in knobs
        for k, n3, v in res["entries"]:
            assert autotune.get_shape(k, n3) == v
            assert autotune.get_shape(k, (n3[0] + 1, n3[1], n3[2])) == 0  # other shapes untouched
        for nm, v in keep.items():
            assert np.array_equal(a.f[nm].numpy(), v), nm
        assert np.array_equal(a.mg.levels[0].phi.numpy(), phi)
        a.step()
        autotune.reset()
        b = solver(precision=precision)
        b.step()
        b.step()
        for nm in ("u", "v", "w", "p"):
            assert np.array_equal(a.f[nm].numpy(), b.f[nm].numpy()), nm
    finally:
        autotune.reset()


def test_tuning_table(gpu):
    from warpflow import _hip, autotune
    autotune.reset()
    with pytest.raises(_hip.HipError):
        autotune.set_shape(autotune.SWEEP_TILE_F32, (8, 8, 8), 5)
    autotune.set_shape(autotune.CD_ICH, (64, 64, 64), 24)
    assert autotune.get_shape(autotune.CD_ICH, (64, 64, 64)) == 24
    assert autotune.get_shape(autotune.CD_ICH, (64, 64, 32)) == 0
    autotune.set_shape(autotune.CD_ICH, (64, 64, 64), 0)
    assert autotune.get_shape(autotune.CD_ICH, (64, 64, 64)) == 0
    autotune.reset()
