"""The two largest stated workloads, run as specified.

* 256^3 cavity (Re 1000, WALE cw 0.46, 1% noise, fixed dt 1e-4, 500 steps,
  divergence tolerance 1e-6, 6 multigrid levels, 8^3 coarsest): the mixed
  form (FP32 V-cycles inside FP64 defect correction, 2 cycles per solve)
  beside the plain FP64 form, step for step -- identical outer-iteration
  counts, both divergence trajectories below the tolerance and within a few
  per cent of each other, velocity rms deviation <= 1e-6 throughout; and the
  FP64 form's first two steps against the C oracle at 256^3 (<= 1e-10).
* 64 uneven cavities (n from {32, 48, 64, 96} per axis, Re in [100, 1000],
  seeded; WALE on, 1% noise with seed member + 1, fixed dt 5e-4, 200 steps,
  4 levels) batched with a graph-captured step: every member bit for bit
  equal to a single-cavity solver of the same configuration, and the cubic
  members within 1e-6 rms of the C oracle's plain FP64 run of that member.
"""
import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import np_oracle as O

pytestmark = pytest.mark.gpu

INNER = (slice(3, -3),) * 3


def rms(a, b):
    return float(np.sqrt(np.mean((a[INNER] - b[INNER]) ** 2)))


def cavity256(precision):
    from warpflow.config import CaseConfig, validate
    c = CaseConfig(name="cavity256_re1000")
    c.grid.nx = c.grid.ny = c.grid.nz = 256
    c.physics.reynolds = 1000.0
    c.physics.sgs_enabled = True
    c.physics.sgs_cw = 0.46
    c.initial.noise = 0.01
    c.numerics.mg_levels = 6
    c.numerics.div_tolerance = 1e-6
    c.numerics.max_outer_iters = 30
    c.numerics.mg_precision = precision
    c.time.variable_dt = False
    c.time.dt_init = 1e-4
    validate(c)
    return c


def solver(cfg):
    from warpflow import FlowSolver, build_topology
    return FlowSolver(cfg, build_topology(cfg), "hip:0")


@pytest.mark.timeout(600)
def test_cavity256_500_steps_mixed_beside_fp64(gpu):
    print("cavity256: 500 steps, mixed and fp64", flush=True)
    mx, f64 = solver(cavity256("mixed")), solver(cavity256("fp64"))
    dev, rmx, r64 = [], [], []
    for it in range(1, 501):
        a, b = mx.step(), f64.step()
        assert a["iters"] == b["iters"], (it, a, b)
        rmx.append(a["rmax"])
        r64.append(b["rmax"])
        if it in (1, 10, 100, 250, 500):
            d = max(rms(mx.f[c].numpy(), f64.f[c].numpy()) for c in "uvw")
            dev.append((it, d))
            assert d <= 1e-6, (it, d)
    rmx, r64 = np.array(rmx), np.array(r64)
    print("velocity rms deviation", dev, "rmax last", rmx[-1], r64[-1], flush=True)
    assert np.all(rmx < 1e-6) and np.all(r64 < 1e-6)
    assert np.max(np.abs(np.log10(rmx / r64))) < 0.05  # within 12% step by step
    assert mx.ctime == pytest.approx(0.05, rel=1e-12)


@pytest.mark.timeout(600)
def test_cavity256_fp64_first_steps_vs_c_oracle(gpu):
    print("cavity256: fp64 vs C oracle, 2 steps", flush=True)
    cfg = cavity256("fp64")
    sol = solver(cfg)
    ref = CO.CSolver(O.Case.from_config(cfg))
    for it in (1, 2):
        a, b = sol.step(), ref.step()
        assert a["iters"] == b["iters"], (it, a, b)
        assert a["rmax"] == pytest.approx(b["rmax"], rel=1e-6), it
    for c in ("u", "v", "w", "p"):
        got, exp = sol.f[c].numpy(), ref.field(c)
        if c == "p":  # the gauge is free: compare demeaned interiors
            got, exp = got[INNER] - got[INNER].mean(), exp[INNER] - exp[INNER].mean()
        err = np.abs(got - exp).max() / np.abs(exp).max()
        assert err <= 1e-10, (c, err)


@pytest.mark.timeout(900)
def test_ensemble64_200_steps(gpu):
    from warpflow.ensemble import Ensemble, draw_members, member_config
    specs = draw_members(64)
    dt, nsteps = 5e-4, 200
    print("ensemble64: %d members, %d steps" % (len(specs), nsteps), flush=True)
    ens = Ensemble(specs, dt, mg_levels=4, graph=True)
    outs = [ens.step() for _ in range(nsteps)]
    iters = np.array([o["iters"] for o in outs])
    rmax_last = outs[-1]["rmax"]
    fields = {m: {nm: ens.field(nm, m).numpy() for nm in ("u", "v", "w", "p", "vis")}
              for m in range(len(specs))}
    ens.close()
    cubic = []
    for m, sp in enumerate(specs):
        cfg = member_config(sp, dt, mg_levels=4)
        sol = solver(cfg)
        for k in range(nsteps):
            info = sol.step()
            assert info["iters"] == iters[k, m], (m, k)
        assert info["rmax"] == rmax_last[m], m
        for nm, got in fields[m].items():
            assert np.array_equal(got, sol.f[nm].numpy()), (m, nm)
        if len(set(sp.n3)) == 1:
            cubic.append((m, sp, cfg))
        del sol
    assert len(cubic) == 4  # 96^3, 48^3, 32^3 and 64^3 members of the seeded draw
    for m, sp, cfg in cubic:
        print("ensemble64: member %d %s vs C oracle" % (m, sp.n3), flush=True)
        ref = CO.CSolver(O.Case.from_config(cfg, seed=sp.seed))
        for _ in range(nsteps):
            ref.step()
        for c in "uvw":
            d = rms(fields[m][c], ref.field(c))
            assert d <= 1e-6, (m, c, d)
