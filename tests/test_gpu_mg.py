"""GPU parity of the multigrid pieces (matrix-free HIP) against the oracle's
stored-coefficient restatement of warpflow/multigrid.py, including the ring
state the reference leaves behind (faces mirrored, edges/corners zero)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import np_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


class _Cfg:
    """Minimal CaseConfig stand-in for Multigrid(cfg, ...)."""

    def __init__(self, levels):
        from warpflow.config import CaseConfig
        c = CaseConfig()
        c.numerics.mg_levels = levels
        self.__dict__.update(c.__dict__)


def build(n3, levels):
    from warpflow.multigrid import Multigrid
    from warpflow.topology import Topology, SerialComm
    topo = Topology(0, 1, SerialComm(), (1, 1, 1), (0, 0, 0),
                    {"x": (-1, -1), "y": (-1, -1), "z": (-1, -1)}, n3, (0, 0, 0))
    h = tuple(1.0 / n for n in n3)
    mg = Multigrid(_Cfg(levels), topo, None, *h)
    ref = O.Multigrid(n3, levels, *h)
    return mg, ref


def load(mg, ref, seed=0):
    rng = np.random.default_rng(seed)
    for lev, rlev in zip(mg.levels, ref.levels):
        phi = rng.standard_normal(rlev.shape)
        rhs = rng.standard_normal(rlev.shape)
        rlev.phi[...] = phi
        rlev.rhs[...] = rhs
        lev.phi.assign(phi)
        lev.rhs.assign(rhs)


@pytest.mark.parametrize("n3", [(16, 16, 16), (8, 12, 20)])
def test_coefficients_bitwise(gpu, n3):
    mg, ref = build(n3, 2)
    for lev, rlev in zip(mg.levels, ref.levels):
        for name in ("aw", "ae", "as_", "an", "ab", "at", "ap"):
            assert np.array_equal(getattr(lev, name).numpy(), getattr(rlev, name)), name


@pytest.mark.parametrize("n3", [(16, 16, 16), (8, 12, 20)])
def test_rbgs_refill_restrict_prolong(gpu, n3):
    mg, ref = build(n3, 2)
    load(mg, ref)
    f, c = mg.levels
    rf, rc = ref.levels
    from warpflow import _hip
    import ctypes as C
    for color in (0, 1):
        O.rbgs(rf, color)
        d = f.desc()
        _hip.call("wf_mg_rbgs", C.byref(d), color, None)
        assert rel_err(f.phi.numpy(), rf.phi) < TOL
    ref._refill(rf)
    mg._refill(f)
    assert rel_err(f.phi.numpy(), rf.phi) < TOL
    # restriction (+ coarse zeroing)
    O.restrict(rf, rc)
    rc.phi[...] = 0.0
    df, dc = f.desc(), c.desc()
    _hip.call("wf_mg_restrict", C.byref(df), C.byref(dc), None)
    assert rel_err(c.rhs.numpy()[1:-1, 1:-1, 1:-1], rc.rhs[1:-1, 1:-1, 1:-1]) < TOL
    cphi = c.phi.numpy()
    assert not cphi[1:-1, 1:-1, 1:-1].any()
    # prolongation reads the coarse ring: faces mirrored, edges zero
    rng = np.random.default_rng(3)
    cint = rng.standard_normal(rc.n3)
    rc.phi[...] = 0.0
    rc.phi[1:-1, 1:-1, 1:-1] = cint
    ref._refill(rc)
    c.phi.assign(rc.phi)
    O.prolong(rc, rf)
    ref._refill(rf)
    _hip.call("wf_mg_prolong", C.byref(dc), C.byref(df), None)
    assert rel_err(f.phi.numpy(), rf.phi) < TOL


@pytest.mark.parametrize("n3,nsweeps", [((8, 8, 8), 300), ((20, 20, 20), 300),
                                        ((20, 20, 20), 6), ((40, 40, 40), 3),
                                        ((16, 8, 24), 5), ((32, 32, 32), 4),
                                        ((64, 48, 40), 6), ((24, 72, 100), 2),
                                        ((9, 7, 11), 6), ((5, 13, 3), 7)])
def test_relax_paths(gpu, n3, nsweeps):
    """_relax: small levels (and long coarsest solves) run in one LDS-resident
    workgroup (colour-packed, zero-ringed; odd extents leave spare words),
    larger ones as out-of-place sweep launches (ping-pong, tile shapes,
    partial tiles, two-sweep launches), an odd leftover sweep as two
    half-sweep launches; all must match the oracle incl. the ring."""
    mg, ref = build(n3, 1)
    load(mg, ref, seed=4)
    ref._relax(0, nsweeps)
    mg._relax(0, nsweeps)
    got = mg.levels[0].phi.numpy()
    assert rel_err(got, ref.levels[0].phi) < 1e-11


@pytest.mark.parametrize("n3,levels", [((32, 32, 32), 3), ((16, 24, 32), 3), ((64, 64, 64), 4)])
def test_vcycle_and_solve(gpu, n3, levels):
    from warpflow.device import array, zeros
    mg, ref = build(n3, levels)
    rng = np.random.default_rng(5)
    shape = tuple(n + 6 for n in n3)
    su = rng.standard_normal(shape)
    pp_ref = np.zeros(shape)
    ref.solve(su, pp_ref, 3)
    pp = zeros(shape)
    mg.solve(array(su), pp, 3)
    got = pp.numpy()
    assert rel_err(got, pp_ref) < 1e-11
    ring = np.zeros(shape, bool)
    ring[2:-2, 2:-2, 2:-2] = True
    assert not got[~ring].any()  # outside interior+ring untouched (zeroed)
    # one more V-cycle from the current state through the API
    ref._vcycle()
    mg._vcycle()
    assert rel_err(mg.levels[0].phi.numpy(), ref.levels[0].phi) < 1e-11


def test_reference_mg_script_flow(gpu):
    """tests/test_mg.py:48-66 driven through the public API: random Neumann
    rhs, phi = 0, refill, V-cycles; the residual (computed from the
    materialised coefficient arrays) drops >= 1 order per cycle."""
    mg, _ = build((32, 32, 32), 3)
    lev0 = mg.levels[0]
    rng = np.random.default_rng(0)
    h = rng.standard_normal(lev0.shape)
    h[0], h[-1], h[:, 0], h[:, -1], h[:, :, 0], h[:, :, -1] = 0, 0, 0, 0, 0, 0
    h[1:-1, 1:-1, 1:-1] -= h[1:-1, 1:-1, 1:-1].mean()
    lev0.rhs.assign(h)
    lev0.phi.zero_()
    mg._refill(lev0)

    def residual(lev):
        phi, rhs = lev.phi.numpy(), lev.rhs.numpy()
        A = [getattr(lev, k).numpy() for k in ("aw", "ae", "as_", "an", "ab", "at", "ap")]
        n = lev.n3[0]
        s = slice(1, n + 1)
        r = rhs[s, s, s] - (A[6][s, s, s] * phi[s, s, s] + A[0][s, s, s] * phi[0:n, s, s]
                            + A[1][s, s, s] * phi[2:n + 2, s, s] + A[2][s, s, s] * phi[s, 0:n, s]
                            + A[3][s, s, s] * phi[s, 2:n + 2, s] + A[4][s, s, s] * phi[s, s, 0:n]
                            + A[5][s, s, s] * phi[s, s, 2:n + 2])
        return np.abs(r).max()
    r = [residual(lev0)]
    for _ in range(4):
        mg._vcycle()
        r.append(residual(lev0))
    assert all(b < 0.1 * a for a, b in zip(r, r[1:])), r


@pytest.mark.parametrize("n3,prec", [((8, 8, 8), np.float64), ((8, 8, 8), np.float32),
                                     ((4, 6, 8), np.float64), ((12, 12, 12), np.float64),
                                     ((16, 16, 16), np.float32), ((20, 20, 20), np.float64),
                                     ((20, 20, 20), np.float32), ((18, 24, 20), np.float64)])
def test_coarse_operator_equals_sweeps(gpu, n3, prec):
    """phi = M rhs (M formed by running the sweeps on unit vectors) equals
    300 RBGS sweeps from phi = 0, ring included (multigrid.py:313-316).
    Levels above 4096 cells take the row x column-chunk apply (20^3 is the
    coarsest level of the 80^3 case on 3 levels)."""
    import ctypes as C
    from warpflow import _hip
    from warpflow.multigrid import _Level
    h = tuple(1.0 / (4 * n) for n in n3)
    a, b = _Level(n3, h, None, prec), _Level(n3, h, None, prec)
    assert a.build_coarse_operator(300)
    rhs = np.random.default_rng(9).standard_normal(a.shape)
    for lev in (a, b):
        lev.rhs.assign(rhs)
        lev.phi.zero_()
    da, db = a.desc(), b.desc()
    _hip.call("wf_mg_coarse_op_apply", C.byref(da), None)
    _hip.call("wf_mg_relax", C.byref(db), 300, None)
    tol = 1e-11 if prec == np.float64 else 2e-5
    assert rel_err(a.phi.numpy(), b.phi.numpy()) < tol


@pytest.mark.parametrize("prec", [np.float64, np.float32])
@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_sweep_tiles_agree_bitwise(gpu, prec, tile):
    """Every fused-sweep tile shape (8x64, 16x32, 16x64, 32x32; one or two
    (even, odd) pairs per thread) and i-chunk -- also the one-pass chunks of
    2 and 4 planes (-2, -4), whose last chunk here holds one plane -- gives
    the built-in result bit for bit, on a level whose extents are multiples
    of none of them."""
    import ctypes as C
    from warpflow import _hip, autotune
    from warpflow.multigrid import _Level
    n3 = (37, 42, 70)
    h = tuple(1.0 / n for n in n3)
    lev = _Level(n3, h, None, prec)
    rng = np.random.default_rng(11)
    rhs = rng.standard_normal(lev.shape)
    phi0 = rng.standard_normal(lev.shape)
    d = lev.desc()
    knob_t = autotune.SWEEP_TILE_F64 if prec == np.float64 else autotune.SWEEP_TILE_F32
    knob_i = autotune.SWEEP_ICH_F64 if prec == np.float64 else autotune.SWEEP_ICH_F32
    knob_p = autotune.SWEEP_PAIRS_F64 if prec == np.float64 else autotune.SWEEP_PAIRS_F32
    lev.rhs.assign(rhs)
    autotune.reset()
    try:
        autotune.set_shape(knob_p, n3, 1)  # one-sweep launches only
        lev.phi.assign(phi0)
        _hip.call("wf_mg_relax", C.byref(d), 4, None)
        want = lev.phi.numpy()
        for ich in (1, 3, 16, -2, -4):
            autotune.set_shape(knob_t, n3, tile)
            autotune.set_shape(knob_i, n3, ich)
            lev.phi.assign(phi0)
            _hip.call("wf_mg_relax", C.byref(d), 4, None)
            assert np.array_equal(lev.phi.numpy(), want), (tile, ich)
    finally:
        autotune.reset()


@pytest.mark.parametrize("prec,tile", [(np.float64, 1), (np.float64, 2), (np.float32, 1),
                                       (np.float32, 2), (np.float32, 3), (np.float32, 4)])
def test_two_sweep_launches_agree_bitwise(gpu, prec, tile):
    """Pairs of sweeps as one two-sweep launch (4 lagged stages, halo
    recomputation) give the one-sweep launches' result bit for bit, for every
    tile (FP32 also with two pairs per thread: 16 x 64, 32 x 32) and several
    i-chunks (the first and last chunks touch the i faces), on a level whose
    extents are multiples of none of them."""
    import ctypes as C
    from warpflow import _hip, autotune
    from warpflow.multigrid import _Level
    n3 = (37, 42, 70)
    h = tuple(1.0 / n for n in n3)
    lev = _Level(n3, h, None, prec)
    rng = np.random.default_rng(12)
    rhs = rng.standard_normal(lev.shape)
    phi0 = rng.standard_normal(lev.shape)
    d = lev.desc()
    knob_t = autotune.SWEEP2_TILE_F64 if prec == np.float64 else autotune.SWEEP2_TILE_F32
    knob_i = autotune.SWEEP2_ICH_F64 if prec == np.float64 else autotune.SWEEP2_ICH_F32
    knob_p = autotune.SWEEP_PAIRS_F64 if prec == np.float64 else autotune.SWEEP_PAIRS_F32
    lev.rhs.assign(rhs)
    autotune.reset()
    try:
        autotune.set_shape(knob_p, n3, 1)
        lev.phi.assign(phi0)
        _hip.call("wf_mg_relax", C.byref(d), 4, None)
        want = lev.phi.numpy()
        autotune.set_shape(knob_p, n3, 2)
        for ich in (1, 3, 16, 40):
            autotune.set_shape(knob_t, n3, tile)
            autotune.set_shape(knob_i, n3, ich)
            lev.phi.assign(phi0)
            _hip.call("wf_mg_relax", C.byref(d), 4, None)
            assert np.array_equal(lev.phi.numpy(), want), (tile, ich)
    finally:
        autotune.reset()


@pytest.mark.parametrize("npre,npost", [(6, 6), (2, 4), (4, 2)])
def test_vcycle_two_sweep_parity(gpu, npre, npost):
    """A V-cycle relaxing with two-sweep launches equals one with one-sweep
    launches bit for bit on every level's phi -- including an odd number of
    two-sweep launches in one smoothing phase, which leaves the iterate in
    tmp and swaps the level's roles until the cycle ends.  The ring's edge
    and corner cells are zero, as the reference leaves them (prolongation
    reads them; no kernel writes them, in phi or tmp)."""
    from warpflow import _hip, autotune
    from warpflow.device import current_stream
    n3 = (48, 40, 64)
    mg, _ = build(n3, 3)
    load(mg, _, seed=9)
    phis = []
    for lev in mg.levels:
        p = lev.phi.numpy()
        ring = np.zeros(p.shape, int)
        for ax in range(3):
            sl = [slice(None)] * 3
            sl[ax] = [0, -1]
            ring[tuple(sl)] += 1
        p[ring >= 2] = 0.0
        phis.append(p)
    descs = mg._descs()
    res = []
    autotune.reset()
    try:
        for pairs in (1, 2):
            for lev in mg.levels:
                autotune.set_shape(autotune.SWEEP_PAIRS_F64, lev.n3, pairs)
            for lev, p in zip(mg.levels, phis):
                lev.phi.assign(p)
            _hip.call("wf_mg_vcycle", descs, len(mg.levels), npre, npost, 8, current_stream())
            res.append([lev.phi.numpy() for lev in mg.levels])
    finally:
        autotune.reset()
    for l, (a, b) in enumerate(zip(*res)):
        assert np.array_equal(a, b), (npre, npost, l, float(np.abs(a - b).max()))
