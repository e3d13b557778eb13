"""Mixed-precision pressure solve on the GPU: against the oracle's FP32
restatement, against the plain FP64 solve, to a tight tolerance, and inside
the flow solver against the plain FP64 form."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import np_oracle as O
from test_gpu_mg import _Cfg
from test_gpu_step import gpu_solver, make_cfg

pytestmark = pytest.mark.gpu


def topo_for(n3):
    from warpflow.topology import SerialComm, Topology
    return Topology(0, 1, SerialComm(), (1, 1, 1), (0, 0, 0),
                    {"x": (-1, -1), "y": (-1, -1), "z": (-1, -1)}, n3, (0, 0, 0))


def demean(a):
    s = (slice(3, -3),) * 3
    return a[s] - a[s].mean()


@pytest.mark.parametrize("n3,levels", [((32, 32, 32), 3), ((16, 32, 24), 3), ((64, 64, 64), 5)])
def test_mixed_solve_matches_oracle_and_fp64(gpu, n3, levels):
    from warpflow.device import array, zeros
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(2).standard_normal(shape)
    mgm = MixedMultigrid(_Cfg(levels), topo_for(n3), None, *h, max_cycles=4,
                         result_mean=False)
    pp = zeros(shape)
    mgm.solve(array(su), pp, 3)
    got = pp.numpy()
    assert mgm.last_cycles == 4
    ref32 = O.Multigrid(n3, levels, *h, dtype=np.float32)
    pp_ref = np.zeros(shape)
    O.solve_mixed(ref32, su, pp_ref, 3, 4, defect_mean=False, result_mean=False)
    assert rel_err(got, pp_ref) < 1e-5
    mg64 = Multigrid(_Cfg(levels), topo_for(n3), None, *h)
    pp64 = zeros(shape)
    mg64.solve(array(su), pp64, 3)
    assert rel_err(got, pp64.numpy()) < 1e-5
    # gauge: with result_mean the interior mean is removed, gradients unchanged
    mgm.result_mean = True
    mgm.solve(array(su), pp, 3)
    g2 = pp.numpy()
    assert abs(g2[3:-3, 3:-3, 3:-3].mean()) < 1e-12 * np.abs(g2).max()
    assert rel_err(demean(g2), demean(got)) < 1e-12


def test_tolerance_mode_reaches_1e10(gpu):
    """Standalone Poisson (tests/test_mg.py:35-45 rhs): the mixed solve
    converges to max|r| <= 1e-10 max|b| in FP64, like plain FP64 V-cycles."""
    from warpflow.device import array, zeros
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    n3, levels = (64, 64, 64), 5
    h = tuple(1.0 / n for n in n3)
    b = np.random.default_rng(0).standard_normal(tuple(n + 2 for n in n3))
    b[0], b[-1], b[:, 0], b[:, -1], b[:, :, 0], b[:, :, -1] = 0, 0, 0, 0, 0, 0
    b[1:-1, 1:-1, 1:-1] -= b[1:-1, 1:-1, 1:-1].mean()
    su = np.zeros(tuple(n + 6 for n in n3))
    su[2:-2, 2:-2, 2:-2] = b
    mgm = MixedMultigrid(_Cfg(levels), topo_for(n3), None, *h, rtol=1e-10, max_cycles=40,
                         result_mean=True)
    pp = zeros(su.shape)
    mgm.solve(array(su), pp, 3)
    assert mgm.last_residual <= 1e-10 and mgm.last_cycles < 20, (mgm.last_cycles, mgm.last_residual)
    mg64 = Multigrid(_Cfg(levels), topo_for(n3), None, *h)
    lev0 = mg64.levels[0]
    lev0.rhs.assign(b)
    lev0.phi.zero_()
    bmax = np.abs(b).max()
    for cyc in range(40):
        mg64._vcycle()
        if mg64.residual_max(0) <= 1e-10 * bmax:
            break
    p64 = np.zeros(su.shape)
    p64[2:-2, 2:-2, 2:-2] = lev0.phi.numpy()
    assert rel_err(demean(pp.numpy()), demean(p64)) < 1e-8
    # the independent residual of the mixed result, in FP64 on the host
    ref = O.Level(n3)
    O.mgcoeff(ref, {s: True for s in O.SIDES}, *h)
    ref.rhs[...] = b
    ref.phi[1:-1, 1:-1, 1:-1] = pp.numpy()[3:-3, 3:-3, 3:-3]
    assert np.abs(O.residual(ref)).max() <= 1.01e-10 * bmax


@pytest.mark.parametrize("cycles,bound", [(4, 1e-10), (2, 1e-7)])
def test_flow_solver_mixed_tracks_fp64(gpu, cycles, bound):
    """Velocity rms deviation of the mixed form from the plain FP64 form over
    20 steps (32^3, Re 400, WALE): at the plain form's 4 cycles only FP32
    rounding separates them; at the default 2 cycles the deviation stays far
    below the 1e-6 gate, with the same outer-iteration and rmax history."""
    cfg = make_cfg(n3=(32, 32, 32), precision="mixed")
    cfg.numerics.mg_mixed_cycles = cycles
    a = gpu_solver(cfg)
    b = gpu_solver(make_cfg(n3=(32, 32, 32), precision="fp64"))
    for _ in range(20):
        ia, ib = a.step(), b.step()
        assert ia["iters"] == ib["iters"]
        assert ia["rmax"] == pytest.approx(ib["rmax"], rel=1e-2)
    for nm in ("u", "v", "w"):
        x = a.f[nm].numpy()[3:-3, 3:-3, 3:-3]
        y = b.f[nm].numpy()[3:-3, 3:-3, 3:-3]
        assert np.sqrt(np.mean((x - y) ** 2)) < bound, nm


@pytest.mark.parametrize("mixed", [False, True])
def test_solve_takes_the_zero_start_unread(gpu, mixed):
    """Both solves start their first V-cycle from phi = 0 without zeroing or
    reading level 0's phi (the first sweep takes the zero iterate through an
    empty descriptor): NaN in the real cells of that phi beforehand gives
    the same pp bit for bit, on a 64^3 level (full-sweep smoother) with
    garbage also in pp's real cells and ring."""
    from warpflow.device import array
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    n3 = (64, 64, 64)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(4).standard_normal(shape)
    if mixed:
        mg = MixedMultigrid(_Cfg(4), topo_for(n3), None, *h, max_cycles=2)
    else:
        mg = Multigrid(_Cfg(4), topo_for(n3), None, *h)
    pp = array(np.zeros(shape))
    mg.solve(array(su), pp, 3)
    want = pp.numpy()
    lev = mg.levels[0]
    phi = lev.phi.numpy()
    phi[1:-1, 1:-1, 1:-1] = np.nan
    lev.phi.assign(phi)
    dirty = np.zeros(shape)
    dirty[2:-2, 2:-2, 2:-2] = 7.0  # real cells and the first ghost layer
    pp2 = array(dirty)
    mg.solve(array(su), pp2, 3)
    assert np.array_equal(pp2.numpy(), want)


@pytest.mark.parametrize("nb", [2, 4])
@pytest.mark.parametrize("mixed", [False, True])
def test_one_pass_sweep_chunks_change_no_result(gpu, mixed, nb):
    """Solves whose every level relaxes with the one-sweep kernel's one-pass
    chunks (nb planes per workgroup, all loads at once) give the built-in
    shapes' pp bit for bit -- including each V-cycle's first sweep from the
    unread zero iterate."""
    from warpflow import autotune
    from warpflow.device import array
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    n3 = (48, 40, 64)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(6).standard_normal(shape)
    if mixed:
        mg = MixedMultigrid(_Cfg(3), topo_for(n3), None, *h, max_cycles=2)
    else:
        mg = Multigrid(_Cfg(3), topo_for(n3), None, *h)
    kp = autotune.SWEEP_PAIRS_F32 if mixed else autotune.SWEEP_PAIRS_F64
    ki = autotune.SWEEP_ICH_F32 if mixed else autotune.SWEEP_ICH_F64
    autotune.reset()
    try:
        pp = array(np.zeros(shape))
        mg.solve(array(su), pp, 3)
        want = pp.numpy()
        for lev in mg.levels:
            autotune.set_shape(kp, lev.n3, 1)
            autotune.set_shape(ki, lev.n3, -nb)
        pp2 = array(np.zeros(shape))
        mg.solve(array(su), pp2, 3)
    finally:
        autotune.reset()
    assert np.array_equal(pp2.numpy(), want)


def test_deferred_gauge_mean_gives_the_same_p(gpu):
    """solve(defer_mean=True) + add_result (the shift of pp applied inside
    p += pp) leaves p bit for bit as solve + add_inplace of the shifted pp."""
    from warpflow import kernels as K
    from warpflow.device import array
    from warpflow.mixed import MixedMultigrid
    n3 = (32, 24, 40)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    rng = np.random.default_rng(8)
    su = rng.standard_normal(shape)
    p0 = rng.standard_normal(shape)
    mg = MixedMultigrid(_Cfg(3), topo_for(n3), None, *h, max_cycles=2)
    assert mg.result_mean
    pp, p = array(np.zeros(shape)), array(p0)
    mg.solve(array(su), pp, 3)
    K.add_inplace(tuple(n + 2 for n in n3), p, pp, 3)
    want = p.numpy()
    pp2, p2 = array(np.zeros(shape)), array(p0)
    mg.solve(array(su), pp2, 3, defer_mean=True)
    mg.add_result(p2, pp2, 3)
    assert np.array_equal(p2.numpy(), want)
    assert not np.array_equal(pp2.numpy(), pp.numpy())  # pp itself unshifted


def test_solve_with_the_first_defect_supplied(gpu):
    """rhs_ready: with (float)su already in level 0's rhs the solve skips its
    first defect and gives the same pp bit for bit."""
    from warpflow.device import array
    from warpflow.mixed import MixedMultigrid
    n3 = (32, 24, 40)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(12).standard_normal(shape)
    mg = MixedMultigrid(_Cfg(3), topo_for(n3), None, *h, max_cycles=2)
    pp = array(np.zeros(shape))
    mg.solve(array(su), pp, 3)
    want = pp.numpy()
    lev = mg.levels[0]
    r = np.full(lev.shape, 5.0, np.float32)  # the ring is not read by the solve's defect path
    r[1:-1, 1:-1, 1:-1] = su[3:-3, 3:-3, 3:-3].astype(np.float32)
    lev.rhs.assign(r)
    pp2 = array(np.zeros(shape))
    mg.solve(array(su), pp2, 3, rhs_ready=True)
    assert np.array_equal(pp2.numpy(), want)


@pytest.mark.parametrize("defect_mean", [False, True])
def test_unrolled_defect_matches_the_plane_walk(gpu, defect_mean):
    """The fixed-cycle solve forms its later defects four planes per load
    group; the tolerance mode forms every defect plane by plane (with the
    residual maxima).  Run to the same cycle count they give pp bit for bit.
    172 x 252 x 232 cells: 5-plane chunks (a full and a one-plane load group)
    and a 2-plane last chunk, the last k tile ragged."""
    from warpflow.device import array
    from warpflow.mixed import MixedMultigrid
    n3 = (172, 252, 232)
    h = tuple(1.0 / n for n in n3)
    shape = tuple(n + 6 for n in n3)
    su = np.random.default_rng(21).standard_normal(shape)
    kw = dict(defect_mean=defect_mean, result_mean=False)
    fixed = MixedMultigrid(_Cfg(3), topo_for(n3), None, *h, max_cycles=3, **kw)
    pp = array(np.zeros(shape))
    fixed.solve(array(su), pp, 3)
    assert fixed.last_cycles == 3
    tol = MixedMultigrid(_Cfg(3), topo_for(n3), None, *h, rtol=1e-300, max_cycles=3, **kw)
    pp2 = array(np.zeros(shape))
    tol.solve(array(su), pp2, 3)
    assert tol.last_cycles == 3
    assert np.array_equal(pp2.numpy(), pp.numpy())
