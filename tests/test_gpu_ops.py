"""GPU parity of every momentum / projection / SGS operator and every ghost
fill against the CPU oracle, on seeded random fields (ghosts included).

Bar: integer and copy work (noise hash, fills, add, fill) bit-exact; floating
point operators within 1e-12 relative (for the ``star = oo - dt*alpha*N``
operators the increment ``star - oo`` is compared, relative to its max-abs,
so the tolerance is not hidden behind ``oo``)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import np_oracle as O

pytestmark = pytest.mark.gpu
N = O.NPL
TOL = 1e-12
SHAPES = [(32, 32, 32), (20, 24, 28)]


def rand_fields(n3, names, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    shape = tuple(n + 2 * N for n in n3)
    return {nm: scale * rng.standard_normal(shape) for nm in names}


def up(a):
    from warpflow.device import array
    return array(a, np.float64)


@pytest.mark.parametrize("n3", SHAPES)
@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_conv_diff_convdiff(gpu, n3, comp):
    from warpflow import kernels as K
    h = tuple(1.0 / n for n in n3)
    dt, alpha = 3e-4, 0.5
    f = rand_fields(n3, ["u", "v", "w", "coo", "vis", "star0"], seed=1)
    f["vis"] = 0.0025 + 0.001 * np.abs(f["vis"])
    dims = {"u": (n3[0] - 1, n3[1], n3[2]), "v": (n3[0], n3[1] - 1, n3[2]),
            "w": (n3[0], n3[1], n3[2] - 1)}[comp]
    d = {k: up(v) for k, v in f.items()}
    # conv
    star_ref = f["star0"].copy()
    getattr(O, "conv4th_" + comp)(f["u"], f["v"], f["w"], f["coo"], star_ref, *h, dt, alpha, N, dims)
    star = up(f["star0"])
    getattr(K, "conv4th_" + comp)(dims, d["u"], d["v"], d["w"], d["coo"], star, *h, dt, alpha, N)
    got = star.numpy()
    assert rel_err(got - f["coo"], star_ref - f["coo"]) < TOL
    ghost = np.ones(got.shape, bool)
    ghost[tuple(slice(N, N + x) for x in dims)] = False
    assert np.array_equal(got[ghost], f["star0"][ghost])  # writes only its box
    # diff on top of conv
    c = {"u": f["u"], "v": f["v"], "w": f["w"]}[comp]
    diff_ref = star_ref.copy()
    getattr(O, "diff4th_" + comp)(c, f["vis"], diff_ref, *h, dt, alpha, N, dims)
    getattr(K, "diff4th_" + comp)(dims, d[comp], d["vis"], star, *h, dt, alpha, N)
    got2 = star.numpy()
    assert rel_err(got2 - star_ref, diff_ref - star_ref) < TOL
    # fused pair == conv then diff
    fused = up(f["star0"])
    getattr(K, "convdiff4th_" + comp)(dims, d["u"], d["v"], d["w"], d["coo"], d["vis"], fused,
                                      *h, dt, alpha, N)
    assert rel_err(fused.numpy() - f["coo"], got2 - f["coo"]) < TOL


@pytest.mark.parametrize("n3", SHAPES + [(9, 17, 70), (41, 10, 33)])
def test_convdiff3_save_equals_single_passes(gpu, n3):
    """The first stage's three-component pass equals the three single-component
    save passes bit for bit: ustar, vstar, wstar and the saved bases, each on
    its operator box only (random ghosts, ragged tiles and chunks)."""
    from warpflow import kernels as K
    h = (0.031, 0.027, 0.019)
    dt, alpha = 3e-4, 0.5
    names = ["u", "v", "w", "vis", "us", "vs", "ws", "uo", "vo", "wo"]
    f = rand_fields(n3, names, seed=11)
    f["vis"] = 0.0025 + 0.001 * np.abs(f["vis"])
    a = {k: up(v) for k, v in f.items()}
    b = {k: up(v) for k, v in f.items()}
    ops = {"u": (n3[0] - 1, n3[1], n3[2]), "v": (n3[0], n3[1] - 1, n3[2]),
           "w": (n3[0], n3[1], n3[2] - 1)}
    for c in "uvw":
        getattr(K, "convdiff4th_save_" + c)(ops[c], a["u"], a["v"], a["w"], a[c + "o"],
                                            a["vis"], a[c + "s"], *h, dt, alpha, N)
    K.convdiff3_save(n3, b["u"], b["v"], b["w"], b["vis"], b["us"], b["vs"], b["ws"], b["uo"],
                     b["vo"], b["wo"], *h, dt, alpha, N)
    for k in names:
        assert np.array_equal(a[k].numpy(), b[k].numpy()), k
    # tuned chunk lengths change nothing
    from warpflow import autotune
    try:
        for ich in (1, 7, 64):
            autotune.set_shape(autotune.CD3_ICH, n3, ich)
            c = {k: up(v) for k, v in f.items()}
            K.convdiff3_save(n3, c["u"], c["v"], c["w"], c["vis"], c["us"], c["vs"], c["ws"],
                             c["uo"], c["vo"], c["wo"], *h, dt, alpha, N)
            for k in ("us", "vs", "ws", "uo", "vo", "wo"):
                assert np.array_equal(c[k].numpy(), b[k].numpy()), (ich, k)
    finally:
        autotune.set_shape(autotune.CD3_ICH, n3, 0)


@pytest.mark.parametrize("n3", SHAPES + [(9, 17, 70), (41, 10, 33)])
def test_convdiff_vw_equals_single_passes(gpu, n3):
    """The second stage's v/w pass equals the single v and w passes in the
    reference's order bit for bit: v's pass reads w before w's fill (w0), w's
    pass after it (w1 = w0 with every cell outside w's operator box changed,
    as a fill may); the v/w pass reads w1 and, where v's z-term reaches those
    cells, the copy save_cross_ghosts took from w0."""
    from warpflow import kernels as K
    from warpflow.device import zeros
    h = (0.031, 0.027, 0.019)
    dt, alpha = 3e-4, 1.0
    names = ["u", "v", "w0", "vo", "wo", "vis", "vs", "ws"]
    f = rand_fields(n3, names, seed=13)
    f["vis"] = 0.0025 + 0.001 * np.abs(f["vis"])
    ni, nj, nk = n3
    box = np.zeros(f["w0"].shape, bool)
    box[N:N + ni, N:N + nj, N:N + nk - 1] = True  # w's operator box
    w1 = f["w0"].copy()
    w1[~box] = np.random.default_rng(14).standard_normal(int((~box).sum()))
    a = {k: up(v) for k, v in f.items()}
    K.convdiff4th_v((ni, nj - 1, nk), a["u"], a["v"], a["w0"], a["vo"], a["vis"], a["vs"], *h,
                    dt, alpha, N)
    aw1 = up(w1)
    K.convdiff4th_w((ni, nj, nk - 1), a["u"], a["v"], aw1, a["wo"], a["vis"], a["ws"], *h, dt,
                    alpha, N)
    b = {k: up(v) for k, v in f.items()}
    w_pre = zeros(w1.shape)
    K.save_cross_ghosts(n3, b["w0"], w_pre, N)
    bw1 = up(w1)
    K.convdiff_vw(n3, b["u"], b["v"], bw1, b["vo"], b["wo"], b["vis"], w_pre, b["vs"], b["ws"],
                  *h, dt, alpha, N)
    assert np.array_equal(a["vs"].numpy(), b["vs"].numpy())
    assert np.array_equal(a["ws"].numpy(), b["ws"].numpy())
    # without the copy v's z-term would see w1's cells: the test has teeth
    c = {k: up(v) for k, v in f.items()}
    K.convdiff_vw(n3, c["u"], c["v"], up(w1), c["vo"], c["wo"], c["vis"], up(w1), c["vs"],
                  c["ws"], *h, dt, alpha, N)
    assert not np.array_equal(c["vs"].numpy(), a["vs"].numpy())


@pytest.mark.parametrize("n3", SHAPES)
def test_calvel_calvel3_calmas(gpu, n3):
    from warpflow import kernels as K
    h = tuple(1.0 / n for n in n3)
    dt = 1e-3
    f = rand_fields(n3, ["us", "vs", "ws", "p", "u", "v", "w", "su"], seed=2)
    d = {k: up(v) for k, v in f.items()}
    ref = {c: f[c].copy() for c in "uvw"}
    ops = {"u": (n3[0] - 1, n3[1], n3[2]), "v": (n3[0], n3[1] - 1, n3[2]),
           "w": (n3[0], n3[1], n3[2] - 1)}
    for c, a in (("u", 0), ("v", 1), ("w", 2)):
        getattr(O, "calvel_" + c)(f[c + "s"], f["p"], ref[c], h[a], dt, 0.5, N, ops[c])
    K.calvel3(n3, d["us"], d["vs"], d["ws"], d["p"], d["u"], d["v"], d["w"], *h, dt, 0.5, N)
    for c in "uvw":
        assert rel_err(d[c].numpy() - f[c + "s"], ref[c] - f[c + "s"]) < TOL, c
        single = up(f[c])
        getattr(K, "calvel_" + c)(ops[c], d[c + "s"], d["p"], single, h["uvw".index(c)], dt,
                                  0.5, N)
        assert np.array_equal(single.numpy(), d[c].numpy()), c
    # calmas on the calvel output
    su_ref = f["su"].copy()
    rmax_ref = O.calmas(ref["u"], ref["v"], ref["w"], su_ref, *h, dt, N, n3)
    from warpflow.device import zeros
    rbuf = zeros(1)
    K.calmas(n3, d["u"], d["v"], d["w"], d["su"], rbuf, *h, dt, N)
    assert rel_err(d["su"].numpy(), su_ref) < TOL
    assert abs(rbuf.numpy()[0] - rmax_ref) <= TOL * rmax_ref


@pytest.mark.parametrize("n3", SHAPES + [(21, 10, 70)])
@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_convdiff_first_stage_saves_the_base(gpu, n3, comp):
    """First-stage conv+diff: the base value is the component itself and is
    stored into c_oo on the operator box -- the same star bits as the plain
    kernel given c_oo = c, c_oo = c on the box, every other cell untouched."""
    from warpflow import kernels as K
    h = tuple(1.0 / n for n in n3)
    f = rand_fields(n3, ["u", "v", "w", "coo", "vis", "star0"], seed=7)
    f["vis"] = 0.0025 + 0.001 * np.abs(f["vis"])
    dims = {"u": (n3[0] - 1, n3[1], n3[2]), "v": (n3[0], n3[1] - 1, n3[2]),
            "w": (n3[0], n3[1], n3[2] - 1)}[comp]
    d = {k: up(v) for k, v in f.items()}
    plain = up(f["star0"])
    getattr(K, "convdiff4th_" + comp)(dims, d["u"], d["v"], d["w"], d[comp], d["vis"], plain,
                                      *h, 3e-4, 0.5, N)
    star = up(f["star0"])
    getattr(K, "convdiff4th_save_" + comp)(dims, d["u"], d["v"], d["w"], d["coo"], d["vis"],
                                           star, *h, 3e-4, 0.5, N)
    assert np.array_equal(star.numpy(), plain.numpy())
    box = tuple(slice(N, N + x) for x in dims)
    want = f["coo"].copy()
    want[box] = f[comp][box]
    assert np.array_equal(d["coo"].numpy(), want)


@pytest.mark.parametrize("n3", SHAPES + [(20, 19, 37), (9, 17, 70), (33, 8, 33)])
def test_calvel3_calmas_one_pass(gpu, n3):
    """The fused calvel/calmas pass equals calvel3 then calmas bit for bit (u, v,
    w incl. ghosts, su, rmax), on random fields with random ghosts, with
    partial (j, k) tiles, and matches the oracle's calvel + calmas."""
    from warpflow import kernels as K
    from warpflow.device import zeros
    h = (0.031, 0.027, 0.019)
    dt = 7e-4
    f = rand_fields(n3, ["us", "vs", "ws", "p", "u", "v", "w", "su"], seed=5)
    a = {k: up(v) for k, v in f.items()}
    b = {k: up(v) for k, v in f.items()}
    ra, rb = zeros(1), zeros(1)
    K.calvel3(n3, a["us"], a["vs"], a["ws"], a["p"], a["u"], a["v"], a["w"], *h, dt, 1.0, N)
    K.calmas(n3, a["u"], a["v"], a["w"], a["su"], ra, *h, dt, N)
    K.calvel3_calmas(n3, b["us"], b["vs"], b["ws"], b["p"], b["u"], b["v"], b["w"], b["su"],
                     rb, *h, dt, 1.0, N)
    for nm in ("u", "v", "w", "su"):
        assert np.array_equal(a[nm].numpy(), b[nm].numpy()), nm
    assert ra.numpy()[0] == rb.numpy()[0] > 0.0
    # the first outer iteration's form: the same su and rmax, u, v, w untouched
    c = {k: up(v) for k, v in f.items()}
    rc = zeros(1)
    K.calvel3_calmas(n3, c["us"], c["vs"], c["ws"], c["p"], c["u"], c["v"], c["w"], c["su"],
                     rc, *h, dt, 1.0, N, store_uvw=False)
    assert np.array_equal(c["su"].numpy(), b["su"].numpy())
    assert rc.numpy()[0] == rb.numpy()[0]
    for nm in ("u", "v", "w"):
        assert np.array_equal(c[nm].numpy(), f[nm]), nm
    # oracle
    ref = {c: f[c].copy() for c in "uvw"}
    ops = {"u": (n3[0] - 1, n3[1], n3[2]), "v": (n3[0], n3[1] - 1, n3[2]),
           "w": (n3[0], n3[1], n3[2] - 1)}
    for c, ax in (("u", 0), ("v", 1), ("w", 2)):
        getattr(O, "calvel_" + c)(f[c + "s"], f["p"], ref[c], h[ax], dt, 1.0, N, ops[c])
    su_ref = f["su"].copy()
    rmax_ref = O.calmas(ref["u"], ref["v"], ref["w"], su_ref, *h, dt, N, n3)
    for c in "uvw":
        assert rel_err(b[c].numpy() - f[c + "s"], ref[c] - f[c + "s"]) < TOL, c
    assert rel_err(b["su"].numpy(), su_ref) < TOL
    assert abs(rb.numpy()[0] - rmax_ref) <= TOL * rmax_ref


@pytest.mark.parametrize("n3", [(20, 19, 37), (32, 32, 32)])
def test_calvel3_calmas_writes_the_fp32_rhs(gpu, n3):
    """With an FP32 level array the fused pass also stores (float)su on its real
    cells (the mixed solve's first defect) and leaves its ring alone; u, v, w,
    su and rmax are unchanged by it."""
    from warpflow import kernels as K
    from warpflow.device import array, zeros
    h = (0.031, 0.027, 0.019)
    f = rand_fields(n3, ["us", "vs", "ws", "p", "u", "v", "w", "su"], seed=11)
    a = {k: up(v) for k, v in f.items()}
    b = {k: up(v) for k, v in f.items()}
    r0 = np.random.default_rng(3).standard_normal(tuple(n + 2 for n in n3)).astype(np.float32)
    rhs = array(r0, np.float32)
    ra, rb = zeros(1), zeros(1)
    K.calvel3_calmas(n3, a["us"], a["vs"], a["ws"], a["p"], a["u"], a["v"], a["w"], a["su"],
                     ra, *h, 7e-4, 1.0, N)
    K.calvel3_calmas(n3, b["us"], b["vs"], b["ws"], b["p"], b["u"], b["v"], b["w"], b["su"],
                     rb, *h, 7e-4, 1.0, N, rhs)
    for nm in ("u", "v", "w", "su"):
        assert np.array_equal(a[nm].numpy(), b[nm].numpy()), nm
    assert ra.numpy()[0] == rb.numpy()[0]
    want = r0.copy()
    want[1:-1, 1:-1, 1:-1] = b["su"].numpy()[N:-N, N:-N, N:-N].astype(np.float32)
    assert np.array_equal(rhs.numpy(), want)


@pytest.mark.parametrize("n3", SHAPES + [(20, 19, 37), (9, 17, 70)])
def test_wale_checkdt_one_pass(gpu, n3):
    """checkdt's reductions and WALE in one pass equal checkdt then WALE bit for
    bit (out4 and vis incl. ghosts), with the vis each cell had before."""
    from warpflow import kernels as K
    from warpflow.device import zeros
    h = (0.031, 0.027, 0.019)
    f = rand_fields(n3, ["u", "v", "w", "vis"], seed=9, scale=0.3)
    f["vis"] = 0.0025 + np.abs(f["vis"]) * 0.01
    a = {k: up(v) for k, v in f.items()}
    b = {k: up(v) for k, v in f.items()}
    oa, ob = zeros(4), zeros(4)
    K.checkdt(n3, a["u"], a["v"], a["w"], a["vis"], oa, *h, N)
    K.wale_viscosity(n3, a["u"], a["v"], a["w"], a["vis"], *h, 1 / 400, 0.46, N)
    K.wale_checkdt(n3, b["u"], b["v"], b["w"], b["vis"], ob, *h, 1 / 400, 0.46, N)
    assert np.array_equal(oa.numpy(), ob.numpy())
    assert np.array_equal(a["vis"].numpy(), b["vis"].numpy())
    assert ob.numpy()[3] < 1e10 and ob.numpy()[0] > 0.0
    # the dt bound of the vis each cell had before, bit for bit: the kernels
    # divide with one FMA correction and take 1 / max instead of min of 1 / x
    exp = O.dtvisc_reduce(f["u"], f["v"], f["w"], f["vis"], *h, N, n3)
    assert ob.numpy()[3] == exp


@pytest.mark.parametrize("n3", SHAPES)
def test_wale_and_checkdt(gpu, n3):
    from warpflow import kernels as K
    from warpflow.device import zeros
    h = tuple(1.0 / n for n in n3)
    f = rand_fields(n3, ["u", "v", "w", "vis"], seed=3, scale=0.3)
    f["vis"] = 0.0025 + np.abs(f["vis"]) * 0.01
    d = {k: up(v) for k, v in f.items()}
    vis_ref = f["vis"].copy()
    O.wale_viscosity(f["u"], f["v"], f["w"], vis_ref, *h, 1 / 400, 0.46, N, n3)
    K.wale_viscosity(n3, d["u"], d["v"], d["w"], d["vis"], *h, 1 / 400, 0.46, N)
    assert rel_err(d["vis"].numpy(), vis_ref) < TOL
    # reductions: separate and fused
    out4 = zeros(4)
    K.checkdt(n3, d["u"], d["v"], d["w"], d["vis"], out4, *h, N)
    got = out4.numpy()
    exp = [O.maxabs_reduce(f[c], N, n3) for c in "uvw"]
    exp.append(O.dtvisc_reduce(f["u"], f["v"], f["w"], vis_ref, *h, N, n3))
    assert np.array_equal(got[:3], exp[:3])
    assert abs(got[3] - exp[3]) <= 1e-14 * exp[3]
    # from the same vis the dt bound is the oracle's bit for bit
    exact = O.dtvisc_reduce(f["u"], f["v"], f["w"], d["vis"].numpy(), *h, N, n3)
    assert got[3] == exact
    b = zeros(1)
    K.maxabs_reduce(n3, d["w"], b, N)
    assert b.numpy()[0] == exp[2]
    b.fill_(1e10)
    K.dtvisc_reduce(n3, d["u"], d["v"], d["w"], d["vis"], b, *h, N)
    assert b.numpy()[0] == exact


@pytest.mark.parametrize("n3", SHAPES)
def test_noise_add_fill_bitwise(gpu, n3):
    from warpflow import kernels as K
    from warpflow.device import zeros
    shape = tuple(n + 2 * N for n in n3)
    ref = [np.zeros(shape) for _ in range(3)]
    O.init_noise(*ref, 0.01, 7, N, n3)
    dev = [zeros(shape) for _ in range(3)]
    K.init_noise(n3, *dev, 0.01, 7, N)
    for a, b in zip(dev, ref):
        assert np.array_equal(a.numpy(), b)
    f = rand_fields(n3, ["p", "pp"], seed=4)
    p_ref = f["p"].copy()
    O.add_inplace(p_ref, f["pp"], N, tuple(n + 2 for n in n3))
    p, pp = up(f["p"]), up(f["pp"])
    K.add_inplace(tuple(n + 2 for n in n3), p, pp, N)
    assert np.array_equal(p.numpy(), p_ref)
    O.fill_const(p_ref, 0.125, N, n3)
    K.fill_const(n3, p, 0.125, N)
    assert np.array_equal(p.numpy(), p_ref)


@pytest.mark.parametrize("n3", SHAPES + [(9, 17, 70), (41, 10, 33)])
def test_convdiff3_stage2_equals_single_passes(gpu, n3):
    """The second stage's three-component pass with u's edge-plane pass equals
    the single u, v and w passes in the reference's fill/conv order bit for
    bit, with the cavity's own fills: u, v, w start filled (the previous
    step's fills), get new interiors (as calvel writes them), then conv_u
    reads u after its fill and v, w before theirs, conv_v reads v after its
    fill and w before, conv_w reads all three filled.  Ragged tiles, and the
    pass's chunk length changes nothing."""
    from warpflow import autotune
    from warpflow import kernels as K
    from warpflow.bcs import BoundaryKernels
    from warpflow.device import zeros
    bnd = {"west": "wall", "east": "wall", "south": "slip", "north": "slip",
           "bottom": "wall", "top": "moving_lid"}
    B = BoundaryKernels(bnd)
    h = (0.031, 0.027, 0.019)
    dt, alpha, ulid = 3e-4, 1.0, 1.0
    ni, nj, nk = n3
    ops = {"u": (ni - 1, nj, nk), "v": (ni, nj - 1, nk), "w": (ni, nj, nk - 1)}
    f = rand_fields(n3, ["u", "v", "w", "uo", "vo", "wo", "vis", "us", "vs", "ws"], seed=17)
    f["vis"] = 0.0025 + 0.001 * np.abs(f["vis"])
    d = {k: up(v) for k, v in f.items()}
    B.apply_uvw(d["u"], d["v"], d["w"], ni, ulid)  # the previous fills
    rng = np.random.default_rng(18)
    pre = {}
    for c in "uvw":  # new interiors, ghosts as the fills left them
        a = d[c].numpy()
        box = tuple(slice(N, N + m) for m in ops[c])
        a[box] += rng.standard_normal(ops[c])
        pre[c] = a
    fill = {"u": lambda x: B.apply_u(x, ni, ulid), "v": lambda x: B.apply_v(x, ni),
            "w": lambda x: B.apply_w(x, ni)}
    post = {}
    for c in "uvw":
        x = up(pre[c])
        fill[c](x)
        post[c] = x.numpy()
    # the reference's order
    a = {k: up(f[k]) for k in ("uo", "vo", "wo", "vis", "us", "vs", "ws")}
    K.convdiff4th_u(ops["u"], up(post["u"]), up(pre["v"]), up(pre["w"]), a["uo"], a["vis"],
                    a["us"], *h, dt, alpha, N)
    K.convdiff4th_v(ops["v"], up(post["u"]), up(post["v"]), up(pre["w"]), a["vo"], a["vis"],
                    a["vs"], *h, dt, alpha, N)
    K.convdiff4th_w(ops["w"], up(post["u"]), up(post["v"]), up(post["w"]), a["wo"], a["vis"],
                    a["ws"], *h, dt, alpha, N)
    want = {k: a[k].numpy() for k in ("us", "vs", "ws")}
    try:
        for ich in (0, 7, 64):
            autotune.set_shape(autotune.CD3S2_ICH, n3, ich)
            b = {k: up(f[k]) for k in ("uo", "vo", "wo", "vis", "us", "vs", "ws")}
            u1, v0, w0 = up(post["u"]), up(pre["v"]), up(pre["w"])
            K.convdiff4th_u_edges(n3, u1, v0, w0, b["uo"], b["vis"], b["us"], *h, dt, alpha, N)
            w_pre = zeros(w0.shape)
            K.save_cross_ghosts(n3, w0, w_pre, N)
            B.apply_vw(v0, w0, ni)  # v0, w0 now hold v's and w's fills
            K.convdiff3(n3, u1, v0, w0, b["uo"], b["vo"], b["wo"], b["vis"], w_pre, b["us"],
                        b["vs"], b["ws"], *h, dt, alpha, N)
            for k in ("us", "vs", "ws"):
                assert np.array_equal(b[k].numpy(), want[k]), (ich, k)
    finally:
        autotune.set_shape(autotune.CD3S2_ICH, n3, 0)
    # the edge pass has teeth: conv_u on the filled v and w differs from the
    # reference's on u's planes 0 and ni-2 and nowhere else
    c = {k: up(f[k]) for k in ("uo", "vis", "us")}
    K.convdiff4th_u(ops["u"], up(post["u"]), up(post["v"]), up(post["w"]), c["uo"], c["vis"],
                    c["us"], *h, dt, alpha, N)
    planes = sorted(set((np.argwhere(c["us"].numpy() != want["us"])[:, 0] - N).tolist()))
    assert planes == [0, ni - 2], planes


# (2, 2, 2): the high side's sources reach the low side's ghosts, so those
# op lists keep the op-by-op walk; the others load every source first.
@pytest.mark.parametrize("n3", SHAPES + [(2, 2, 2), (3, 5, 4)])
def test_ghost_fills_bitwise(gpu, n3):
    from warpflow.bcs import BoundaryKernels, bc_const, bc_lin
    bnd = {"west": "wall", "east": "wall", "south": "slip", "north": "slip",
           "bottom": "wall", "top": "moving_lid"}
    B = BoundaryKernels(bnd)
    f = rand_fields(n3, ["u", "v", "w", "vis", "p", "q"], seed=5)
    d = {k: up(v) for k, v in f.items()}
    O.apply_u(f["u"], n3, 1.0)
    B.apply_u(d["u"], n3[0], 1.0)
    O.apply_v(f["v"], n3)
    B.apply_v(d["v"], n3[0])
    O.apply_w(f["w"], n3)
    B.apply_w(d["w"], n3[0])
    O.apply_vis(f["vis"], n3, 0.0025)
    B.apply_vis(d["vis"], n3[0], 0.0025)
    O.apply_pbound(f["p"], n3)
    B.apply_pbound(d["p"], n3[0])
    for k in ("u", "v", "w", "vis", "p"):
        assert np.array_equal(d[k].numpy(), f[k]), k
    O.apply_p_mirror(f["q"], n3, 3)
    B.apply_p_mirror(d["q"], n3[0], 3)
    assert np.array_equal(d["q"].numpy(), f["q"])
    # u, v, w in one call (one x launch and one y-then-z launch for the three)
    h = rand_fields(n3, ["u", "v", "w"], seed=7)
    dh = {k: up(v) for k, v in h.items()}
    O.apply_u(h["u"], n3, 0.75)
    O.apply_v(h["v"], n3)
    O.apply_w(h["w"], n3)
    B.apply_uvw(dh["u"], dh["v"], dh["w"], n3[0], 0.75)
    for k in ("u", "v", "w"):
        assert np.array_equal(dh[k].numpy(), h[k]), ("uvw", k)
    # a refill (fields filled since their last change): only the band lines
    # run, and every cell equals a second full fill
    full = {k: up(dh[k].numpy()) for k in ("u", "v", "w")}
    part = {k: up(dh[k].numpy()) for k in ("u", "v", "w")}
    B.apply_uvw(full["u"], full["v"], full["w"], n3[0], 0.75)
    B.refill_uvw(part["u"], part["v"], part["w"], n3[0], 0.75)
    for k in ("u", "v", "w"):
        assert np.array_equal(part[k].numpy(), full[k].numpy()), ("refill", k)
    # v, w in one call (u absent)
    h = rand_fields(n3, ["v", "w"], seed=8)
    dh = {k: up(v) for k, v in h.items()}
    O.apply_v(h["v"], n3)
    O.apply_w(h["w"], n3)
    B.apply_vw(dh["v"], dh["w"], n3[0])
    for k in ("v", "w"):
        assert np.array_equal(dh[k].numpy(), h[k]), ("vw", k)
    # single _mk_lin / _mk_const launches
    g = rand_fields(n3, ["a"], seed=6)["a"]
    da = up(g)
    O.bc_lin(g, 1, True, N + n3[1], N + n3[1] - 1, -1.0, 1, n3[0] + 3, 2, n3[2] + 1, 3)
    bc_lin(da, 1, True, N + n3[1], N + n3[1] - 1, -1.0, 1, n3[0] + 3, 2, n3[2] + 1, 3)
    O.bc_const(g, 2, False, N - 1, 0.7, 0, n3[0] + 6, 0, n3[1] + 6, 2)
    bc_const(da, 2, False, N - 1, 0.7, 0, n3[0] + 6, 0, n3[1] + 6, 2)
    assert np.array_equal(da.numpy(), g)


def test_host_checks_reject_bad_boxes(gpu):
    from warpflow import _hip
    from warpflow import kernels as K
    from warpflow.device import zeros
    n3 = (8, 8, 8)
    a = [zeros((14, 14, 14)) for _ in range(5)]
    with pytest.raises(_hip.HipError, match="invalid argument"):
        K.conv4th_u((9, 8, 8), *a, 0.1, 0.1, 0.1, 1e-3, 0.5, 3)  # box + reach outside
    with pytest.raises(_hip.HipError, match="invalid argument"):
        from warpflow.bcs import bc_lin
        bc_lin(a[0], 0, False, 2, 3, 1.0, 0, 15, 0, 14, 1)  # perpendicular overrun
    K.conv4th_u(n3, *a, 0.1, 0.1, 0.1, 1e-3, 0.5, 3)  # the full box is fine
