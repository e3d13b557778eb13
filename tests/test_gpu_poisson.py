"""Standalone Poisson solve at 256^3 (cell-centred FP64, homogeneous Neumann
walls, 4 levels): the plain FP64 RBGS V-cycle repeated to 1e-10 relative
residual against FP32 V-cycles inside FP64 defect correction to the same
tolerance.

The right-hand side is the reference's standalone multigrid test input
(tests/test_mg.py:35-45): a seeded standard-normal field on the level-0
array, ring zeroed, interior demeaned.  The residual norm is the
reference's max|rhs - A phi| over real cells (tests/test_mg.py:16-32),
re-evaluated on the CPU by the oracle for both results, and the two
pressures are compared after removing their means (the Neumann problem
fixes phi up to a constant)."""
import time

import numpy as np
import pytest

from oracle import np_oracle as O

pytestmark = pytest.mark.gpu
N = 256
LEVELS = 4
RTOL = 1e-10


def cfg_for(n, levels):
    from warpflow.config import CaseConfig, validate
    c = CaseConfig()
    c.grid.nx = c.grid.ny = c.grid.nz = n
    c.numerics.mg_levels = levels
    validate(c)
    return c


def make_rhs(n):
    rng = np.random.default_rng(0)
    h = rng.standard_normal((n + 2,) * 3)
    h[0], h[-1] = 0.0, 0.0
    h[:, 0], h[:, -1] = 0.0, 0.0
    h[:, :, 0], h[:, :, -1] = 0.0, 0.0
    h[1:-1, 1:-1, 1:-1] -= h[1:-1, 1:-1, 1:-1].mean()
    return h


def cpu_relres(phi, rhs, n):
    lev = O.Level((n,) * 3)
    O.mgcoeff(lev, {s: True for s in O.SIDES}, 1.0 / n, 1.0 / n, 1.0 / n)
    lev.rhs[...] = rhs
    lev.phi[...] = phi
    return float(np.abs(O.residual(lev)).max() / np.abs(rhs[1:-1, 1:-1, 1:-1]).max())


def test_poisson256_fp64_and_mixed_to_1e10(gpu):
    from warpflow import build_topology
    from warpflow.device import array, synchronize, zeros
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    cfg = cfg_for(N, LEVELS)
    topo = build_topology(cfg)
    h = 1.0 / N
    rhs = make_rhs(N)
    bmax = np.abs(rhs[1:-1, 1:-1, 1:-1]).max()

    # plain FP64: repeated V(6,6) cycles (300 coarsest sweeps) from phi = 0
    mg = Multigrid(cfg, topo, "hip:0", h, h, h)
    lev0 = mg.levels[0]
    lev0.rhs.assign(rhs)
    lev0.phi.zero_()
    mg._refill(lev0)
    hist = []
    synchronize()
    t0 = time.perf_counter()
    for _ in range(40):
        mg._vcycle()
        hist.append(mg.residual_max() / bmax)
        if hist[-1] <= RTOL:
            break
    t64 = time.perf_counter() - t0
    phi64 = lev0.phi.numpy()
    assert hist[-1] <= RTOL, hist
    # every cycle reduces the residual (geometric convergence, no stagnation)
    assert all(b < a for a, b in zip(hist, hist[1:])), hist

    # FP32 V-cycles inside FP64 defect correction, same tolerance
    mm = MixedMultigrid(cfg, topo, "hip:0", h, h, h, rtol=RTOL, max_cycles=40)
    su = array(rhs, np.float64)
    pp = zeros(rhs.shape, np.float64)
    synchronize()
    t0 = time.perf_counter()
    mm.solve(su, pp, 1)
    synchronize()
    tmx = time.perf_counter() - t0
    pmx = pp.numpy()
    assert mm.last_residual <= RTOL
    assert mm.last_cycles <= len(hist) + 2, (mm.last_cycles, len(hist))

    # the oracle's residual of both results
    r64 = cpu_relres(phi64, rhs, N)
    rmx = cpu_relres(pmx, rhs, N)
    assert r64 <= 1.01 * RTOL and rmx <= 1.01 * RTOL, (r64, rmx)
    a = phi64[1:-1, 1:-1, 1:-1]
    b = pmx[1:-1, 1:-1, 1:-1]
    a = a - a.mean()
    b = b - b.mean()
    err = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(a ** 2)))
    assert err < 1e-8, err
    print("\npoisson 256^3 L%d: fp64 %d cycles %.2f ms (res %.2e)  mixed %d cycles %.2f ms "
          "(res %.2e)  rms|p_mix - p_64|/rms|p_64| = %.2e  factors %s"
          % (LEVELS, len(hist), 1e3 * t64, r64, mm.last_cycles, 1e3 * tmx, rmx, err,
             " ".join("%.3f" % (b / a) for a, b in zip(hist, hist[1:]))))
