"""Failure handling on the HIP path, as the reference raises and warns:
"time step collapsed" (solver.py:142-144), "projection diverged"
(solver.py:190-193) and the non-convergence warning (solver.py:205-209), for
the single-cavity solver and per member of a batched ensemble.

The conditions are planted in the fields from the host (a NaN, a velocity
spike) or by the outer-iteration limit.  Every kernel still runs to its end on
the data it is given; the errors are raised by the host code that reads the
reductions back."""
import numpy as np
import pytest

from test_gpu_ensemble import DT, specs
from test_gpu_step import gpu_solver, make_cfg

pytestmark = pytest.mark.gpu

CELL = (8, 9, 10)  # an interior cell of a 16^3 field (3 ghost layers)


def plant(arr, value, at=CELL):
    from warpflow import _hip
    a = arr.numpy()
    a[at] = value
    arr.assign(a)
    _hip.call("wf_synchronize")


@pytest.mark.parametrize("fused,variable", [(True, True), (True, False), (False, True),
                                            (False, False)])
def test_nan_velocity_collapses_the_time_step(gpu, fused, variable):
    """A NaN in u makes max|u| and dtvisc NaN (the reductions propagate NaN),
    so dtmax is not below 1e10 and compute_dt raises -- at once with a
    variable dt, with the first calmas read-back of the step with a fixed dt."""
    print("case: nan u, fused=%s variable=%s" % (fused, variable), flush=True)
    sol = gpu_solver(make_cfg(n3=(16, 16, 16), variable=variable, dt=1e-3), fused=fused)
    sol.step()
    plant(sol.f["u"], np.nan)
    with pytest.raises(RuntimeError, match="time step collapsed"):
        sol.step()


@pytest.mark.parametrize("fused", [True, False])
def test_velocity_spike_diverges_the_projection(gpu, fused):
    """u = 1e4 in one cell: dt stays finite, but the first divergence of the
    projection is far above 10 (rmax = max|div dV|)."""
    print("case: spike u, fused=%s" % fused, flush=True)
    sol = gpu_solver(make_cfg(n3=(16, 16, 16), variable=False, dt=1e-3), fused=fused)
    sol.step()
    plant(sol.f["u"], 1.0e4)
    with pytest.raises(RuntimeError, match="projection diverged"):
        sol.step()


@pytest.mark.parametrize("fused", [True, False])
def test_outer_iteration_limit_warns(gpu, capsys, fused):
    """max_outer_iters = 1: the break needs it > 1, so the loop always runs
    out and the warning is printed; the step still completes."""
    cfg = make_cfg(n3=(16, 16, 16), variable=False, dt=1e-3)
    cfg.numerics.max_outer_iters = 1
    sol = gpu_solver(cfg, fused=fused)
    info = sol.step()
    assert info["iters"] == 1 and np.isfinite(info["rmax"])
    assert "warning: projection did not converge" in capsys.readouterr().out


def test_ensemble_nan_member_collapses_the_time_step(gpu):
    from warpflow.ensemble import Ensemble
    ens = Ensemble(specs(), DT, mg_levels=4, graph=True)
    try:
        ens.step()
        plant(ens.field("u", 2), np.nan)
        with pytest.raises(RuntimeError, match=r"time step collapsed in members \[2\]"):
            ens.step()
    finally:
        ens.close()


def test_ensemble_spike_member_diverges(gpu):
    from warpflow.ensemble import Ensemble
    ens = Ensemble(specs(), DT, mg_levels=4, graph=True)
    try:
        ens.step()
        plant(ens.field("u", 3), 1.0e4)
        with pytest.raises(RuntimeError, match=r"projection diverged .* members \[3\]"):
            ens.step()
    finally:
        ens.close()


def test_ensemble_nan_rmax_counts_as_diverged(gpu):
    """A NaN that appears after checkdt (here: planted in p, which the
    projection's calvel reads) gives a NaN rmax; the member is stopped as
    diverged instead of iterating to the limit."""
    from warpflow.ensemble import Ensemble
    ens = Ensemble(specs(), DT, mg_levels=4, graph=True)
    try:
        ens.step()
        plant(ens.field("p", 1), np.nan)
        with pytest.raises(RuntimeError, match=r"projection diverged .* members \[1\]"):
            ens.step()
    finally:
        ens.close()


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_outer_iteration_limit_warns(gpu, capsys, graph):
    from warpflow.ensemble import Ensemble
    ens = Ensemble(specs(), DT, mg_levels=4, graph=graph, maxcy=1)
    try:
        out = ens.step()
        assert list(out["iters"]) == [1] * ens.nmem
        msg = capsys.readouterr().out
        assert "warning: projection did not converge in members [0, 1, 2, 3, 4]" in msg
    finally:
        ens.close()
