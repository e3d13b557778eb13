"""Pin the CPU oracle (oracle/np_oracle.py) before trusting it.

The reference ships no golden vectors (its only collected tests are the
invariants of tests/test_cavity.py) and cannot be imported here (warp-lang
and mpi4py are absent).  Each check below is chosen so that a plausible
transcription mistake -- a dropped term, a wrong stencil coefficient, a
half-cell index shift, a transposed operand -- fails it:

* closed forms: the 4-point interpolant/derivative are exact for cubics and
  the 4th-order Laplacian for quintics, so on polynomial fields conv4th,
  diff4th, calvel and calmas must reproduce analytic values at the staggered
  positions (a half-cell shift of the REPORT.md kind is an O(dx) error);
* WALE from matrix algebra (G @ G) with the reference's two documented
  substitutions, on linear velocity fields whose discrete gradient is exact;
* init_noise recomputed with Python integers;
* the multigrid operator assembled independently with scipy, transfer
  operators on linear fields, V-cycle convergence to the scipy solution;
* face ghosts against the independent rule set of the reference's
  tests/h3d_bridge.py:27-56;
* the reference's published numbers: step-1 rmax = 1.69e-7 at 80^3
  (REPORT.md:51,62) and the tests/test_cavity.py invariants at 24^3.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import np_oracle as O

N = O.NPL


# ---------------------------------------------------------------------------
# staggered-grid coordinates of every full-array index (kernels.py:9-12):
# u(i,j,k) on the +x face of cell (i,j,k), v on the +y face, w on the +z face
# ---------------------------------------------------------------------------
def coords(n3, h3, comp):
    axes = []
    for a in range(3):
        idx = np.arange(n3[a] + 2 * N) - N
        face = (a == {"u": 0, "v": 1, "w": 2}.get(comp, -1))
        axes.append((idx + (1.0 if face else 0.5)) * h3[a])
    return np.meshgrid(*axes, indexing="ij")


N3 = (12, 10, 14)
H3 = (0.1, 0.13, 0.07)
RNG = np.random.default_rng(7)


def lin_field(coef, X, Y, Z):
    return coef[0] + coef[1] * X + coef[2] * Y + coef[3] * Z


def interior(n3, dims=None):
    dims = dims or n3
    return tuple(slice(N, N + d) for d in dims)


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_conv4th_exact_on_linear_fields(comp):
    """With u, v, w linear in x, y, z every product is quadratic, so the
    stencils are exact and conv = d(c u)/dx + d(c v)/dy + d(c w)/dz at c's
    face position."""
    A = {c: RNG.normal(size=4) for c in "uvw"}
    f = {}
    for c in "uvw":
        X, Y, Z = coords(N3, H3, c)
        f[c] = lin_field(A[c], X, Y, Z)
    Xc, Yc, Zc = coords(N3, H3, comp)
    coo = RNG.normal(size=f["u"].shape)
    star = np.zeros_like(coo)
    dt, alpha = 0.01, 0.5
    kern = {"u": O.conv4th_u, "v": O.conv4th_v, "w": O.conv4th_w}[comp]
    dims = (N3[0] - 1, N3[1] - 1, N3[2] - 1)
    kern(f["u"], f["v"], f["w"], coo, star, *H3, dt, alpha, N, dims)
    # analytic: c = A0 + A1 x + A2 y + A3 z; d(c q)/dx_b = c_b q + c q_b
    val = {c: lin_field(A[c], Xc, Yc, Zc) for c in "uvw"}
    grad = {c: A[c][1:] for c in "uvw"}
    cc = val[comp]
    conv = sum(grad[comp][b] * val[q] + cc * grad[q][b] for b, q in enumerate("uvw"))
    s = interior(N3, dims)
    got = (coo[s] - star[s]) / (dt * alpha)
    np.testing.assert_allclose(got, conv[s], rtol=1e-10, atol=1e-10)


def test_conv4th_detects_half_cell_shift():
    """The check above is sensitive to the REPORT.md half-cell bug: shifting
    the cross-term interpolation point by one index changes the result by an
    O(dx) amount on the same fields."""
    A = {c: RNG.normal(size=4) for c in "uvw"}
    f = {c: lin_field(A[c], *coords(N3, H3, c)) for c in "uvw"}
    dims = (N3[0] - 1, N3[1] - 1, N3[2] - 1)
    B = O._box(N, dims)
    good = O._iy(B, f["u"], 0, 1, 0) * O._ix(B, f["v"], 1, 0, 0)
    shifted = O._iy(B, f["u"], 0, 0, 0) * O._ix(B, f["v"], 1, 0, 0)
    assert np.abs(good - shifted).max() > 1e-3


@pytest.mark.parametrize("comp,voff", [("u", (0, 0, 0)), ("v", (0, 1, 0)), ("w", (0, 0, 1))])
def test_diff4th_exact_on_quintics(comp, voff):
    """L4 is exact for quintics; viscosity is sampled at vis[i,j,k] (u),
    vis[i,j+1,k] (v), vis[i,j,k+1] (w) (kernels.py:172-174)."""
    X, Y, Z = coords(N3, H3, comp)
    c = X ** 5 - 2 * Y ** 4 * X + 0.5 * Z ** 5 + X * Y * Z ** 3
    lap = 20 * X ** 3 - 24 * Y ** 2 * X + 10 * Z ** 3 + 6 * X * Y * Z
    vis = 1.0 + np.arange(c.shape[1])[None, :, None] * 0.01 + \
        np.arange(c.shape[2])[None, None, :] * 0.003 + np.zeros_like(c)
    star = RNG.normal(size=c.shape)
    ref = star.copy()
    dt, alpha = 0.02, 1.0
    kern = {"u": O.diff4th_u, "v": O.diff4th_v, "w": O.diff4th_w}[comp]
    kern(c, vis, star, *H3, dt, alpha, N, N3)
    s = interior(N3)
    vs = tuple(slice(N + o, N + o + n) for o, n in zip(voff, N3))
    np.testing.assert_allclose((star[s] - ref[s]) / (dt * alpha), vis[vs] * lap[s],
                               rtol=1e-9, atol=1e-6)


def test_calvel_and_calmas_closed_form():
    """calvel on a linear p is the exact gradient; calmas of cubic
    velocities is the exact divergence at cell centres."""
    X, Y, Z = coords(N3, H3, "p")
    gp = RNG.normal(size=3)
    p = 0.3 + gp[0] * X + gp[1] * Y + gp[2] * Z
    dt, alpha = 0.05, 0.5
    for comp, kern, a in (("u", O.calvel_u, 0), ("v", O.calvel_v, 1), ("w", O.calvel_w, 2)):
        star = RNG.normal(size=p.shape)
        out = np.zeros_like(p)
        kern(star, p, out, H3[a], dt, alpha, N, N3)
        s = interior(N3)
        np.testing.assert_allclose(out[s], star[s] - dt * alpha * gp[a], rtol=1e-12, atol=1e-12)
    Xu, Yu, Zu = coords(N3, H3, "u")
    Xv, Yv, Zv = coords(N3, H3, "v")
    Xw, Yw, Zw = coords(N3, H3, "w")
    u = Xu ** 3 + Yu * Zu
    v = 2 * Yv ** 3 - Xv * Yv
    w = -Zw ** 3 + Xw * Zw ** 2
    div = 3 * X ** 2 + 6 * Y ** 2 - X - 3 * Z ** 2 + 2 * X * Z
    su = np.zeros_like(u)
    rmax = O.calmas(u, v, w, su, *H3, dt, N, N3)
    s = interior(N3)
    np.testing.assert_allclose(su[s] * dt, div[s], rtol=1e-10, atol=1e-10)
    assert rmax == pytest.approx(np.abs(div[s]).max() * H3[0] * H3[1] * H3[2], rel=1e-10)


def _wale_from_matrix(G, delta, cw, textbook=False):
    """WALE via matrix algebra: g2 = G @ G with G[a][b] = d u_a / d x_b; the
    reference replaces the last term of g2[z][x] by dwdz*dudz and of g2[z][y]
    by dwdz*dudy (kernels.py:336, :338)."""
    g2 = G @ G
    if not textbook:
        g2[2, 0] = G[2, 0] * G[0, 0] + G[2, 1] * G[1, 0] + G[2, 2] * G[0, 2]
        g2[2, 1] = G[2, 0] * G[0, 1] + G[2, 1] * G[1, 1] + G[2, 2] * G[0, 1]
    Sd = 0.5 * (g2 + g2.T) - np.trace(g2) / 3.0 * np.eye(3)
    S = 0.5 * (G + G.T)
    sdsd = np.sum(Sd * Sd)
    ss = np.sum(S * S)
    if sdsd <= 1e-8:
        return 0.0
    return (cw * delta) ** 2 * sdsd ** 1.5 / (ss ** 2.5 + sdsd ** 1.25)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_wale_linear_fields_match_matrix_form(seed):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(3, 3))
    f = {}
    for a, c in enumerate("uvw"):
        X, Y, Z = coords(N3, H3, c)
        f[c] = 0.2 * a + G[a, 0] * X + G[a, 1] * Y + G[a, 2] * Z
    vis = np.zeros_like(f["u"])
    rrey, cw = 1.0 / 400.0, 0.46
    O.wale_viscosity(f["u"], f["v"], f["w"], vis, *H3, rrey, cw, N, N3)
    delta = O.wale_delta(*H3)
    nut = _wale_from_matrix(G, delta, cw)
    np.testing.assert_allclose(vis[interior(N3)], rrey + nut, rtol=1e-12)
    # the substitution matters unless dwdz vanishes
    assert abs(_wale_from_matrix(G, delta, cw, textbook=True) - nut) > 1e-9 * nut


def test_wale_textbook_when_dwdz_zero_and_threshold():
    G = np.array([[0.3, -1.1, 0.4], [0.7, 0.2, -0.5], [0.9, 1.3, 0.0]])
    assert _wale_from_matrix(G, 0.1, 0.46) == pytest.approx(
        _wale_from_matrix(G, 0.1, 0.46, textbook=True), rel=1e-14)
    # uniform flow: vis == 1/Re exactly; tiny gradients below the 1e-8
    # switch-on threshold also give exactly 1/Re
    for scale in (0.0, 1e-6):
        f = {}
        for a, c in enumerate("uvw"):
            X, Y, Z = coords(N3, H3, c)
            f[c] = 1.0 + scale * (G[a, 0] * X + G[a, 1] * Y + G[a, 2] * Z)
        vis = np.zeros_like(f["u"])
        O.wale_viscosity(f["u"], f["v"], f["w"], vis, *H3, 0.0025, 0.46, N, N3)
        assert np.all(vis[interior(N3)] == 0.0025)


def test_dtvisc_uses_cell_centred_averages():
    """u alternating in sign along x averages to zero at cell centres, so the
    advective x-term drops out; v, w uniform; vis uniform."""
    shape = tuple(n + 2 * N for n in N3)
    U = 2.0
    u = U * (-1.0) ** np.arange(shape[0])[:, None, None] * np.ones(shape)
    v = np.full(shape, -0.5)
    w = np.full(shape, 0.25)
    vis = np.full(shape, 0.01)
    dx, dy, dz = H3
    got = O.dtvisc_reduce(u, v, w, vis, dx, dy, dz, N, N3)
    expect = 1.0 / (0.5 / dy + 0.25 / dz + 2 * 0.01 * (1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))
    assert got == pytest.approx(expect, rel=1e-14)
    assert O.maxabs_reduce(u, N, N3) == U


def _hash_py(i, j, k, seed):
    M = 0xFFFFFFFF
    h = ((i * 73856093) & M) ^ ((j * 19349663) & M) ^ ((k * 83492791) & M) ^ (seed & M)
    h ^= h >> 13
    h = (h * 1274126177) & M
    h ^= h >> 16
    return h


def test_init_noise_bitwise_and_statistics():
    n3 = (16, 16, 16)
    shape = tuple(n + 2 * N for n in n3)
    u, v, w = (np.zeros(shape) for _ in range(3))
    amp = 0.01
    O.init_noise(u, v, w, amp, 1, N, n3)
    for (i, j, k) in [(3, 3, 3), (3, 4, 5), (18, 3, 11), (10, 18, 7), (18, 18, 18)]:
        h = _hash_py(i, j, k, 1)
        r = (float(h) / 4294967295.0) * 2.0 - 1.0
        assert u[i, j, k] == amp * r and v[i, j, k] == amp * r and w[i, j, k] == -amp * r
    s = interior(n3)
    assert np.all(u[s] == v[s]) and np.all(w[s] == -u[s])
    assert np.abs(u[s]).max() <= amp
    hist, _ = np.histogram(u[s], bins=8, range=(-amp, amp))
    assert np.all(np.abs(hist / u[s].size - 1 / 8) < 0.02)
    ghosts = u.copy()
    ghosts[s] = 0.0
    assert not ghosts.any()


# ---------------------------------------------------------------------------
# multigrid
# ---------------------------------------------------------------------------
def neumann_matrix(n3, h3):
    """7-point cell-centred Laplacian with zero-flux faces, assembled from
    1-D second-difference matrices (scipy), no ghost cells."""
    ops = []
    for n, h in zip(n3, h3):
        main = -2.0 * np.ones(n)
        main[0] = main[-1] = -1.0
        ops.append(sp.diags([np.ones(n - 1), main, np.ones(n - 1)], [-1, 0, 1]) / h ** 2)
    I = [sp.identity(n) for n in n3]
    return (sp.kron(sp.kron(ops[0], I[1]), I[2]) + sp.kron(sp.kron(I[0], ops[1]), I[2])
            + sp.kron(sp.kron(I[0], I[1]), ops[2])).tocsr()


def test_mg_operator_matches_scipy_assembly():
    n3, h3 = (6, 5, 7), (0.2, 0.25, 1.0 / 7)
    lev = O.Level(n3)
    O.mgcoeff(lev, {s: True for s in O.SIDES}, *h3)
    lev.phi[...] = RNG.normal(size=lev.shape)  # ring values must not matter
    lev.rhs[...] = 0.0
    A = neumann_matrix(n3, h3)
    x = lev.phi[1:-1, 1:-1, 1:-1].ravel()
    np.testing.assert_allclose(-O.residual(lev).ravel(), A @ x, rtol=1e-12, atol=1e-10)


def test_rbgs_half_sweep_matches_pointwise_gauss_seidel():
    n3, h3 = (4, 5, 6), (0.25, 0.2, 1.0 / 6)
    lev = O.Level(n3)
    O.mgcoeff(lev, {s: True for s in O.SIDES}, *h3)
    lev.phi[...] = RNG.normal(size=lev.shape)
    lev.rhs[...] = RNG.normal(size=lev.shape)
    A = neumann_matrix(n3, h3).toarray()
    x = lev.phi[1:-1, 1:-1, 1:-1].ravel().copy()
    b = lev.rhs[1:-1, 1:-1, 1:-1].ravel()
    ii, jj, kk = np.unravel_index(np.arange(x.size), n3)
    for color in (0, 1):
        O.rbgs(lev, color)
        for q in range(x.size):  # plain Gauss-Seidel over the colour set
            if (ii[q] + jj[q] + kk[q] + color) % 2 == 1:
                x[q] = (b[q] - (A[q] @ x - A[q, q] * x[q])) / A[q, q]
        np.testing.assert_allclose(lev.phi[1:-1, 1:-1, 1:-1].ravel(), x, rtol=1e-12)


def test_transfer_operators_on_linear_fields():
    nf = (8, 8, 8)
    f, c = O.Level(nf), O.Level(tuple(n // 2 for n in nf))
    hf = (0.1, 0.1, 0.1)
    for lev, h in ((f, hf), (c, tuple(2 * x for x in hf))):
        O.mgcoeff(lev, {s: True for s in O.SIDES}, *h)
    a = RNG.normal(size=4)

    def centres(n3, h):
        ax = [(np.arange(n) + 0.5) * hh for n, hh in zip(n3, h)]
        return np.meshgrid(*ax, indexing="ij")
    # restriction of a residual equals the average of the 8 children: with
    # phi = 0 a linear rhs restricts to its value at the coarse centre
    Xf = centres(nf, hf)
    f.phi[...] = 0.0
    f.rhs[1:-1, 1:-1, 1:-1] = lin_field(a, *Xf)
    O.restrict(f, c)
    Xc = centres(c.n3, (0.2, 0.2, 0.2))
    np.testing.assert_allclose(c.rhs[1:-1, 1:-1, 1:-1], lin_field(a, *Xc), rtol=1e-12)
    # prolongation of a linear coarse function is exact away from the ring
    c.phi[1:-1, 1:-1, 1:-1] = lin_field(a, *Xc)
    f.phi[...] = 0.0
    O.prolong(c, f)
    inner = (slice(2, -2),) * 3
    np.testing.assert_allclose(f.phi[inner], lin_field(a, *Xf)[(slice(1, -1),) * 3],
                               rtol=1e-12)


def make_rhs(shape, seed=0):
    """tests/test_mg.py:35-45: seeded normal rhs, zero ring, demeaned."""
    h = np.random.default_rng(seed).standard_normal(shape)
    h[0], h[-1], h[:, 0], h[:, -1], h[:, :, 0], h[:, :, -1] = 0, 0, 0, 0, 0, 0
    h[1:-1, 1:-1, 1:-1] -= h[1:-1, 1:-1, 1:-1].mean()
    return h


def test_vcycle_converges_to_scipy_solution():
    n3 = (16, 16, 16)
    h = 1.0 / 16
    mg = O.Multigrid(n3, 3, h, h, h)
    lev0 = mg.levels[0]
    lev0.rhs[...] = make_rhs(lev0.shape)
    lev0.phi[...] = 0.0
    res = [np.abs(O.residual(lev0)).max()]
    for _ in range(6):
        mg._vcycle()
        res.append(np.abs(O.residual(lev0)).max())
    # test_mg.py:1-3: 1-2 orders of magnitude per cycle
    factors = np.array(res[1:]) / np.array(res[:-1])
    assert np.all(factors[:4] < 0.1), factors
    A = neumann_matrix(n3, (h, h, h))
    b = lev0.rhs[1:-1, 1:-1, 1:-1].ravel()
    x = spla.lsqr(A, b, atol=1e-14, btol=1e-14, iter_lim=20000)[0]
    got = lev0.phi[1:-1, 1:-1, 1:-1].ravel()
    np.testing.assert_allclose(got - got.mean(), x - x.mean(), atol=1e-6 * np.abs(x).max())


# ---------------------------------------------------------------------------
# ghost fills
# ---------------------------------------------------------------------------
_NORMAL = {"west": "u", "east": "u", "south": "v", "north": "v", "bottom": "w", "top": "w"}
_BND = {"west": "wall", "east": "wall", "south": "slip", "north": "slip",
        "bottom": "wall", "top": "moving_lid"}


def h3d_face_rule(g, name, n, ulid):
    """The face-ghost rules of tests/h3d_bridge.py:27-56 (wall: odd mirror of
    tangential components, slip/lid: even mirror, normal component 0, lid u
    = ulid), evaluated for face ghosts whose two other indices are real."""
    out = {}
    axes = {"west": 0, "east": 0, "south": 1, "north": 1, "bottom": 2, "top": 2}
    for side, kind in _BND.items():
        ax = axes[side]
        for m in range(N):
            if side in ("west", "south", "bottom"):
                gi, ri = N - 1 - m, N + m
            else:
                gi, ri = N + n + m, N + n - 1 - m
            if kind == "moving_lid" and name == "u":
                val = ulid
            elif name == _NORMAL[side]:
                val = 0.0
            else:
                val = (1.0 if kind in ("slip", "moving_lid") else -1.0) * np.take(g, ri, axis=ax)
            out[(side, m)] = (ax, gi, val)
    return out


@pytest.mark.parametrize("name", ["u", "v", "w"])
def test_face_ghosts_follow_h3d_rules(name):
    n = 8
    shape = (n + 2 * N,) * 3
    g = RNG.normal(size=shape)
    fill = {"u": lambda a: O.apply_u(a, n, 1.0), "v": lambda a: O.apply_v(a, n),
            "w": lambda a: O.apply_w(a, n)}[name]
    fill(g)
    real = slice(N, N + n - (1 if name == "u" else 0))
    real_all = slice(N, N + n)
    for (side, m), (ax, gi, val) in h3d_face_rule(g, name, n, 1.0).items():
        sl = [real_all, real_all, real_all]
        # the staggered component's own axis: the wall face N+n-1 is set by
        # the fill, the h3d rule treats it as real -- skip that position
        if name == "u":
            sl[0] = real
        if name == "v":
            sl[1] = slice(N, N + n - 1)
        if name == "w":
            sl[2] = slice(N, N + n - 1)
        if _NORMAL[side] == name:
            # normal component: the reference zeroes the ghosts and, on the
            # high side, starts at the wall face itself, so its three layers
            # are N+n-1..N+n+1 (bcs.py:208, :237, :264)
            sl[ax] = gi - 1 if side in ("east", "north", "top") else gi
            assert np.all(g[tuple(sl)] == 0.0), (side, m)
            continue
        sl[ax] = gi
        got = g[tuple(sl)]
        exp = val if np.isscalar(val) else val[tuple(s for a, s in enumerate(sl) if a != ax)]
        np.testing.assert_array_equal(got, np.broadcast_to(exp, got.shape), err_msg=str((side, m)))


def test_pbound_and_vis_faces():
    n = 6
    shape = (n + 2 * N,) * 3
    p = RNG.normal(size=shape)
    O.apply_pbound(p, n)
    s = slice(N, N + n)
    for ax in range(3):
        lo = [s, s, s]
        lo[ax] = N - 1
        src = [s, s, s]
        src[ax] = N
        np.testing.assert_array_equal(p[tuple(lo)], p[tuple(src)])
    vis = RNG.normal(size=shape)
    O.apply_vis(vis, n, 0.0025)
    assert np.all(vis[N, s, s] == 0.0025) and np.all(vis[N - 1, s, s] == 0.0025)
    assert np.all(vis[N + n - 1, s, s] == 0.0025) and np.all(vis[s, s, N] == 0.0025)


# ---------------------------------------------------------------------------
# whole step: the reference's published numbers and regression invariants
# ---------------------------------------------------------------------------
def test_step1_rmax_matches_report():
    """REPORT.md:51,62: 80^3, Re 400, WALE, 1% noise (seed 1), fixed dt 3e-4:
    step-1 divergence residual 1.69e-7."""
    case = O.Case(n3=(80, 80, 80), reynolds=400, mg_levels=3, variable_dt=False, dt_init=3e-4)
    sol = O.Solver(case)
    info = sol.step()
    assert info["rmax"] == pytest.approx(1.69e-7, abs=0.006e-7)


@pytest.fixture(scope="module")
def cavity24():
    """tests/test_cavity.py:24-65: 24^3, 30 steps, variable dt, WALE."""
    sol = O.Solver(O.Case(n3=(24, 24, 24), reynolds=400, mg_levels=3))
    infos = [sol.step() for _ in range(30)]
    return sol, infos


def test_cavity_projection_converges(cavity24):
    _, infos = cavity24
    for info in infos:
        assert info["rmax"] < 1.0e-6
        assert info["iters"] <= 6


def test_cavity_fields_finite_and_ghosts(cavity24):
    sol, _ = cavity24
    n = sol.ncells[0]
    for name in ("u", "v", "w", "p", "vis"):
        assert np.isfinite(sol.f[name]).all()
    s = slice(N, N + n)
    u, v, w, p = (sol.f[c] for c in "uvwp")
    assert np.allclose(u[N:N + n, N:N + n, N + n], sol.ulid)
    assert np.allclose(u[N - 1, s, s], 0) and np.allclose(u[N + n, s, s], 0)
    assert np.allclose(v[s, N - 1, s], 0) and np.allclose(v[s, N + n, s], 0)
    assert np.allclose(w[s, s, N - 1], 0) and np.allclose(w[s, s, N + n], 0)
    assert np.allclose(p[N - 1, s, s], p[N, s, s])
    assert np.allclose(p[N + n, s, s], p[N + n - 1, s, s])
    assert np.allclose(p[s, N - 1, s], p[s, N, s])
    assert np.allclose(p[s, N + n, s], p[s, N + n - 1, s])
    assert np.allclose(p[s, s, N - 1], p[s, s, N])
    assert np.allclose(p[s, s, N + n], p[s, s, N + n - 1])


def test_c_oracle_matches_numpy_oracle_bitwise():
    """Two independent transcriptions of the step schedule (NumPy and C)."""
    from oracle import c_oracle as CO
    for case in (O.Case(n3=(24, 24, 24)),
                 O.Case(n3=(16, 24, 32), sgs_enabled=False, variable_dt=False,
                        dt_init=1e-3, reynolds=100)):
        a, b = O.Solver(case), CO.CSolver(case)
        for _ in range(3):
            ia, ib = a.step(), b.step()
            assert ia["iters"] == ib["iters"] and ia["dt"] == ib["dt"]
            assert a.rmax_history == b.rmax_history()
        for nm in ("u", "v", "w", "p", "ustar", "su", "pp"):
            assert np.array_equal(a.f[nm], b.field(nm)), nm
        np.testing.assert_allclose(b.field("vis"), a.f["vis"], rtol=1e-15)


def test_mixed_form_reproduces_fp64_vcycles():
    """FP32 V-cycles inside FP64 defect correction, with the defect mean
    kept, equal the plain FP64 V-cycles up to FP32 rounding (a V-cycle from
    zero on the defect is the same stationary iteration)."""
    n3, h = (32, 32, 32), 1.0 / 32
    su = np.random.default_rng(1).standard_normal(tuple(n + 6 for n in n3))
    mg64 = O.Multigrid(n3, 3, h, h, h)
    pp64 = np.zeros_like(su)
    mg64.solve(su, pp64, 3)
    mg32 = O.Multigrid(n3, 3, h, h, h, dtype=np.float32)
    pp = np.zeros_like(su)
    O.solve_mixed(mg32, su, pp, 3, 4, defect_mean=False, result_mean=False)
    s = (slice(3, -3),) * 3
    assert np.abs(pp[s] - pp64[s]).max() < 1e-5 * np.abs(pp64[s]).max()


# --------------------------------------------------------------------------
# Edge and corner ghosts: every ghost value as a composition of face rules.
#
# The Hydro3D rules behind bcs.py:201-293 (its comments state them in
# Fortran form, e.g. "u(is-1-ly) = 0", "v(is-1-ly) = -v(is+ly)") are written
# below as a table of face rules per field, in the side order W, E, S, N, B, T.
# A rule either sets a set of planes to a constant, or mirrors layer m of a
# side onto the m-th plane inside it with a sign; it covers only its
# perpendicular index ranges.  A cell's final value is found by pulling, not
# by replaying the fills on an array: take the LAST rule covering the cell;
# a constant ends the chain, a mirror multiplies by its sign and moves to the
# reflected cell, where only rules EARLIER than this one may have written.
# The chain ends at a constant or at a cell no earlier rule covers, whose
# value is the field's value before the fill -- edges and corners outside
# some rule's range keep such earlier (history-dependent) values.  Each
# ghost value is then +-F[source] or a constant, and must equal the
# oracle's array-replay fill bit for bit on a field whose ghosts start out
# random.
# --------------------------------------------------------------------------
def _face_rules(name, n3, ulid=0.75, rrey=1.0 / 400.0):
    ni, nj, nk = n3
    nn = (ni, nj, nk)
    full = [(N - 1, N + x) for x in nn]           # first ghost layer .. first ghost layer
    short = [(N - 1, N + x - 1) for x in nn]      # staggered: up to the wall face

    def planes_lo(first, nl):                     # first, first-1, ..
        return [first - m for m in range(nl)]

    def planes_hi(first, nl):
        return [first + m for m in range(nl)]

    def mirror(ax, high, sign, perp, nl=3):
        # layer m: ghost N-1-m <- N+m (low); N+n+m <- N+n-1-m (high)
        n = nn[ax]
        pairs = ([(N + n + m, N + n - 1 - m) for m in range(nl)] if high
                 else [(N - 1 - m, N + m) for m in range(nl)])
        return ("mirror", ax, dict(pairs), sign, perp)

    def const(ax, planes, val, perp):
        return ("const", ax, set(planes), val, perp)

    def perp(ax, r):  # perpendicular ranges, keyed by axis
        return {a: r[a] for a in range(3) if a != ax}

    if name == "u":
        # u(is-1-ly) = 0 / u(ie+1+ly) = 0 from the wall face / u(js-1-ly) =
        # u(js+ly) / u(ks-1-ly) = -u(ks+ly) / u(ke+1+ly) = ulid
        ry = [short[0], full[1], full[2]]
        return [const(0, planes_lo(N - 1, 3), 0.0, perp(0, full)),
                const(0, planes_hi(N + ni - 1, 3), 0.0, perp(0, full)),
                mirror(1, False, 1.0, perp(1, ry)), mirror(1, True, 1.0, perp(1, ry)),
                mirror(2, False, -1.0, perp(2, ry)),
                const(2, planes_hi(N + nk, 3), ulid, perp(2, ry))]
    if name == "v":
        rx = [full[0], short[1], full[2]]
        return [mirror(0, False, -1.0, perp(0, rx)), mirror(0, True, -1.0, perp(0, rx)),
                const(1, planes_lo(N - 1, 3), 0.0, perp(1, full)),
                const(1, planes_hi(N + nj - 1, 3), 0.0, perp(1, full)),
                mirror(2, False, -1.0, perp(2, rx)), mirror(2, True, 1.0, perp(2, rx))]
    if name == "w":
        rx = [full[0], full[1], short[2]]
        return [mirror(0, False, -1.0, perp(0, rx)), mirror(0, True, -1.0, perp(0, rx)),
                mirror(1, False, 1.0, perp(1, rx)), mirror(1, True, 1.0, perp(1, rx)),
                const(2, planes_lo(N - 1, 3), 0.0, perp(2, full)),
                const(2, planes_hi(N + nk - 1, 3), 0.0, perp(2, full))]
    if name == "vis":
        # one layer; at the W, E and B walls the first real plane is set to
        # 1/Re and then mirrored (eddyvis_wale ghost section)
        p = lambda ax: perp(ax, full)  # noqa: E731
        return [const(0, [N], rrey, p(0)), mirror(0, False, 1.0, p(0), 1),
                const(0, [N + ni - 1], rrey, p(0)), mirror(0, True, 1.0, p(0), 1),
                mirror(1, False, 1.0, p(1), 1), mirror(1, True, 1.0, p(1), 1),
                const(2, [N], rrey, p(2)), mirror(2, False, 1.0, p(2), 1),
                mirror(2, True, 1.0, p(2), 1)]
    raise ValueError(name)


def _covers(rule, c):
    kind, ax, where, _, rng = rule
    on_plane = c[ax] in where  # const: a plane of the set; mirror: a ghost plane
    return on_plane and all(lo <= c[a] <= hi for a, (lo, hi) in rng.items())


def _compose(rules, c):
    """(sign, source cell or None, constant): the value at cell c after all
    rules is sign * F[source], or the constant when source is None."""
    sign, last = 1.0, len(rules)
    while True:
        hit = next((q for q in range(last - 1, -1, -1) if _covers(rules[q], c)), None)
        if hit is None:
            return sign, c, None
        kind, ax, where, val, _ = rules[hit]
        if kind == "const":
            return sign, None, val
        c = tuple(where[c[ax]] if a == ax else c[a] for a in range(3))
        sign *= val
        last = hit


@pytest.mark.parametrize("n3", [(6, 6, 6), (5, 7, 6)])
@pytest.mark.parametrize("name", ["u", "v", "w", "vis"])
def test_edge_and_corner_ghosts_compose_face_rules(name, n3):
    rules = _face_rules(name, n3)
    shape = tuple(x + 2 * N for x in n3)
    f0 = RNG.normal(size=shape)
    got = f0.copy()
    {"u": lambda a: O.apply_u(a, n3, 0.75), "v": lambda a: O.apply_v(a, n3),
     "w": lambda a: O.apply_w(a, n3), "vis": lambda a: O.apply_vis(a, n3, 1.0 / 400.0)}[name](got)
    exp = np.empty(shape)
    chains = {0: 0, 1: 0, 2: 0, 3: 0}
    for c in np.ndindex(*shape):
        sign, src, val = _compose(rules, c)
        exp[c] = sign * (val if src is None else f0[src])
        ghosts = sum(not (N <= c[a] < N + n3[a]) for a in range(3))
        if ghosts >= 2 and (src is None or src != c):
            chains[ghosts] += 1
    # bit for bit, signed zeros included
    assert np.array_equal(got.view(np.int64), exp.view(np.int64)), \
        [tuple(int(x) for x in i) for i in np.argwhere(got.view(np.int64) != exp.view(np.int64))[:5]]
    # the fills do reach edges (two ghost axes) and corners (three)
    assert chains[2] > 0 and chains[3] > 0, chains


# --------------------------------------------------------------------------
# The fused momentum passes reorder the reference's fill/conv sequence
# (solver.py:254-265).  What makes that exact, checked on the oracle:
#   first stage: u, v, w were filled since their last change, and a second
#   fill changes only deep edge ghosts that no other component's conv4th
#   stencil reads;
#   second stage: after calvel the other components' ghosts do matter (so
#   conv_u must see v and w before their fills), and conv_v reads w's fill
#   only at the cells outside w's operator box in rows -1, nj and columns
#   -2, -1, nk-1, nk -- the cells the v/w pass takes from its saved copy.
# --------------------------------------------------------------------------
def _fill_all(f, n, ulid=0.75):
    O.apply_u(f["u"], n, ulid)
    O.apply_v(f["v"], n)
    O.apply_w(f["w"], n)


def _conv(comp, f, n, dims):
    star = np.zeros_like(f["u"])
    oo = np.ones_like(f["u"])
    h = 1.0 / n
    getattr(O, "conv4th_" + comp)(f["u"], f["v"], f["w"], oo, star, h, h, h, 1e-3, 0.5, O.NPL,
                                  dims[comp])
    return star


@pytest.mark.parametrize("n", [8, 13])
def test_refill_changes_no_cell_a_cross_stencil_reads(n):
    N = O.NPL
    rng = np.random.default_rng(31)
    shape = (n + 2 * N,) * 3
    dims = {"u": (n - 1, n, n), "v": (n, n - 1, n), "w": (n, n, n - 1)}
    filled = {k: rng.standard_normal(shape) for k in "uvw"}  # random ghosts before
    _fill_all(filled, n)
    refilled = {k: a.copy() for k, a in filled.items()}
    _fill_all(refilled, n)
    # the refill does change cells (edges whose rule reads a later side's ghost)
    assert any(not np.array_equal(filled[k], refilled[k]) for k in "uvw")
    for comp in "uvw":
        own = {k: (refilled[k] if k == comp else filled[k]) for k in "uvw"}
        assert np.array_equal(_conv(comp, own, n, dims), _conv(comp, refilled, n, dims)), comp
    # nor any conv4th + diff4th result at all: the first stage's pass runs on
    # the once-filled fields and one fill after it reaches the twice-filled
    # state, which a third fill no longer changes
    vis = 0.002 + 0.001 * np.abs(rng.standard_normal(shape))
    h = 1.0 / n
    for comp in "uvw":
        stars = []
        for F in (filled, refilled):
            star = _conv(comp, F, n, dims)
            getattr(O, "diff4th_" + comp)(F[comp], vis, star, h, h, h, 1e-3, 0.5, N, dims[comp])
            stars.append(star)
        assert np.array_equal(stars[0], stars[1]), comp
    third = {k: a.copy() for k, a in refilled.items()}
    _fill_all(third, n)
    for k in "uvw":
        assert np.array_equal(third[k], refilled[k]), k


@pytest.mark.parametrize("n", [8, 13])
def test_second_stage_cross_reads(n):
    N = O.NPL
    rng = np.random.default_rng(32)
    shape = (n + 2 * N,) * 3
    dims = {"u": (n - 1, n, n), "v": (n, n - 1, n), "w": (n, n, n - 1)}
    f = {k: rng.standard_normal(shape) for k in "uvw"}
    _fill_all(f, n)
    for k in "uvw":  # calvel: new interiors, ghosts left as they were
        f[k][N:N + dims[k][0], N:N + dims[k][1], N:N + dims[k][2]] += \
            rng.standard_normal(dims[k])
    fresh = {k: a.copy() for k, a in f.items()}
    _fill_all(fresh, n)
    # conv_u reads v's and w's fills (so it must run before them) ...
    mixed_u = {"u": fresh["u"], "v": fresh["v"], "w": fresh["w"]}
    stale_u = {"u": fresh["u"], "v": f["v"], "w": f["w"]}
    assert not np.array_equal(_conv("u", mixed_u, n, dims), _conv("u", stale_u, n, dims))
    # ... and conv_v reads w's fill exactly at the cells the v/w pass saves
    w_view = fresh["w"].copy()
    ni = nj = nk = n
    for J in (-1, nj):
        w_view[N:N + ni, N + J, N - 2:N + nk + 1] = f["w"][N:N + ni, N + J, N - 2:N + nk + 1]
    for K in (-2, -1, nk - 1, nk):
        w_view[N:N + ni, N:N + nj, N + K] = f["w"][N:N + ni, N:N + nj, N + K]
    ref_v = _conv("v", {"u": fresh["u"], "v": fresh["v"], "w": f["w"]}, n, dims)
    assert np.array_equal(_conv("v", {"u": fresh["u"], "v": fresh["v"], "w": w_view}, n, dims),
                          ref_v)
    assert not np.array_equal(_conv("v", fresh, n, dims), ref_v)
    # ... while of v's and w's fills conv_u reads only the x-ghost planes -1
    # and ni: with those two planes from before the fills the fresh fields
    # give conv_u's values, and with none of them only u's planes 0 and ni-2
    # (the ones the edge pass runs before the fills) differ
    ref_u = _conv("u", stale_u, n, dims)
    view = {"u": fresh["u"], "v": fresh["v"].copy(), "w": fresh["w"].copy()}
    for k in "vw":
        for I in (-1, ni):
            view[k][N + I] = f[k][N + I]
    assert np.array_equal(_conv("u", view, n, dims), ref_u)
    got = _conv("u", mixed_u, n, dims)
    diff_planes = sorted(set((np.argwhere(got != ref_u)[:, 0] - N).tolist()))
    assert diff_planes == [0, ni - 2], diff_planes
