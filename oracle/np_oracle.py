"""NumPy FP64 restatement of the warpflow-les per-step hot path (TEST ORACLE).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  The product never
imports this module; tests, ``smoke()`` and ``bench.py``'s cpu_baseline leg use
it as the checker.

Every function restates one reference kernel or method and cites the
reference ``file:line`` it follows (paths relative to the reference repo
``dr-chunshui/warpflow-les``).  Conventions kept from the reference:

* fields are C-ordered ``[i, j, k]`` float64 arrays of shape ``(n+2*NPL)`` per
  axis, ``k`` contiguous; every staggered component shares the cell index
  range (``warpflow/kernels.py:9-12``);
* a Warp launch of ``dim = (d0, d1, d2)`` whose kernel adds ``npl`` to
  ``wp.tid()`` is a vectorised operation over the full-array index box
  ``[npl, npl+d0) x [npl, npl+d1) x [npl, npl+d2)``; ``_Box.at(f, a, b, c)``
  returns ``f[i+a, j+b, k+c]`` for every thread ``(i, j, k)`` of that box;
* arithmetic is written in the reference's evaluation order (left to right,
  as Warp's code generator emits it), so operator results agree with any
  faithful IEEE FP64 implementation to a few ulps.

Generalisation beyond the reference: the reference hard-wires one cell count
``n = ncells[0]`` for every axis in the ghost fills and the multigrid
coefficients (``warpflow/solver.py:91``, ``warpflow/multigrid.py:252``).  Here
each axis uses its own count; with ``nx == ny == nz`` every range and
coefficient is exactly the reference's.
"""
from __future__ import annotations

import numpy as np

NPL = 3          # ghost layers of every field (warpflow/topology.py:14)
NPL_MG = 1       # ghost layers of every multigrid level (warpflow/multigrid.py:20)
ALFARK = (0.5, 1.0)  # Heun RK2 stage coefficients (warpflow/solver.py:26)


class _Box:
    """Index box of one Warp launch: thread (i, j, k) -> full index
    (lo0 + i, lo1 + j, lo2 + k)."""

    def __init__(self, lo, dims):
        self.lo = tuple(int(x) for x in lo)
        self.dims = tuple(int(x) for x in dims)

    def at(self, f, a=0, b=0, c=0):
        (l0, l1, l2), (d0, d1, d2) = self.lo, self.dims
        return f[l0 + a:l0 + a + d0, l1 + b:l1 + b + d1, l2 + c:l2 + c + d2]

    def sl(self):
        (l0, l1, l2), (d0, d1, d2) = self.lo, self.dims
        return (slice(l0, l0 + d0), slice(l1, l1 + d1), slice(l2, l2 + d2))


def _box(npl, dims):
    return _Box((npl, npl, npl), dims)


# ===========================================================================
# warpflow/kernels.py
# ===========================================================================

def _ix(B, f, a, b, c):
    """kernels.py:31-34  I(x) = (-f[i+1] + 9 f[i] + 9 f[i-1] - f[i-2]) / 16"""
    return (-B.at(f, a + 1, b, c) + 9.0 * B.at(f, a, b, c)
            + 9.0 * B.at(f, a - 1, b, c) - B.at(f, a - 2, b, c)) / 16.0


def _iy(B, f, a, b, c):
    """kernels.py:37-40"""
    return (-B.at(f, a, b + 1, c) + 9.0 * B.at(f, a, b, c)
            + 9.0 * B.at(f, a, b - 1, c) - B.at(f, a, b - 2, c)) / 16.0


def _iz(B, f, a, b, c):
    """kernels.py:43-46"""
    return (-B.at(f, a, b, c + 1) + 9.0 * B.at(f, a, b, c)
            + 9.0 * B.at(f, a, b, c - 1) - B.at(f, a, b, c - 2)) / 16.0


def _d4(sp, sn, sm, smm, h):
    """The staggered 4-point derivative shared by conv4th and calmas:
    (-s[+] + 27 s[0] - 27 s[-] + s[--]) / (24 h)  (kernels.py:69, :80, ...)."""
    return (-sp + 27.0 * sn - 27.0 * sm + smm) / (24.0 * h)


def conv4th_u(u, v, w, uoo, ustar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:53-89  ustar = uoo - dt*alpha*(du2/dx + duv/dy + duw/dz)."""
    B = _box(npl, dims)
    sp2 = _ix(B, u, 2, 0, 0) * _ix(B, u, 2, 0, 0)
    sp1 = _ix(B, u, 1, 0, 0) * _ix(B, u, 1, 0, 0)
    sn = _ix(B, u, 0, 0, 0) * _ix(B, u, 0, 0, 0)
    sm1 = _ix(B, u, -1, 0, 0) * _ix(B, u, -1, 0, 0)
    du2dx = _d4(sp2, sp1, sn, sm1, dx)
    # cross terms: interpolation point j+1 in the derivative direction
    # (kernels.py:72-80, the REPORT.md half-cell fix)
    sp1 = _iy(B, u, 0, 2, 0) * _ix(B, v, 1, 1, 0)
    sn = _iy(B, u, 0, 1, 0) * _ix(B, v, 1, 0, 0)
    sm1 = _iy(B, u, 0, 0, 0) * _ix(B, v, 1, -1, 0)
    sm2 = _iy(B, u, 0, -1, 0) * _ix(B, v, 1, -2, 0)
    duvdy = _d4(sp1, sn, sm1, sm2, dy)
    sp1 = _iz(B, u, 0, 0, 2) * _ix(B, w, 1, 0, 1)
    sn = _iz(B, u, 0, 0, 1) * _ix(B, w, 1, 0, 0)
    sm1 = _iz(B, u, 0, 0, 0) * _ix(B, w, 1, 0, -1)
    sm2 = _iz(B, u, 0, 0, -1) * _ix(B, w, 1, 0, -2)
    duwdz = _d4(sp1, sn, sm1, sm2, dz)
    ustar[B.sl()] = B.at(uoo) - dt * alpha * (du2dx + duvdy + duwdz)


def conv4th_v(u, v, w, voo, vstar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:92-129"""
    B = _box(npl, dims)
    sp2 = _iy(B, v, 0, 2, 0) * _iy(B, v, 0, 2, 0)
    sp1 = _iy(B, v, 0, 1, 0) * _iy(B, v, 0, 1, 0)
    sn = _iy(B, v, 0, 0, 0) * _iy(B, v, 0, 0, 0)
    sm1 = _iy(B, v, 0, -1, 0) * _iy(B, v, 0, -1, 0)
    dv2dy = _d4(sp2, sp1, sn, sm1, dy)
    sp1 = _ix(B, v, 2, 0, 0) * _iy(B, u, 1, 1, 0)
    sn = _ix(B, v, 1, 0, 0) * _iy(B, u, 0, 1, 0)
    sm1 = _ix(B, v, 0, 0, 0) * _iy(B, u, -1, 1, 0)
    sm2 = _ix(B, v, -1, 0, 0) * _iy(B, u, -2, 1, 0)
    duvdx = _d4(sp1, sn, sm1, sm2, dx)
    sp1 = _iz(B, v, 0, 0, 2) * _iy(B, w, 0, 1, 1)
    sn = _iz(B, v, 0, 0, 1) * _iy(B, w, 0, 1, 0)
    sm1 = _iz(B, v, 0, 0, 0) * _iy(B, w, 0, 1, -1)
    sm2 = _iz(B, v, 0, 0, -1) * _iy(B, w, 0, 1, -2)
    dvwdz = _d4(sp1, sn, sm1, sm2, dz)
    vstar[B.sl()] = B.at(voo) - dt * alpha * (duvdx + dv2dy + dvwdz)


def conv4th_w(u, v, w, woo, wstar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:132-168"""
    B = _box(npl, dims)
    sp2 = _iz(B, w, 0, 0, 2) * _iz(B, w, 0, 0, 2)
    sp1 = _iz(B, w, 0, 0, 1) * _iz(B, w, 0, 0, 1)
    sn = _iz(B, w, 0, 0, 0) * _iz(B, w, 0, 0, 0)
    sm1 = _iz(B, w, 0, 0, -1) * _iz(B, w, 0, 0, -1)
    dw2dz = _d4(sp2, sp1, sn, sm1, dz)
    sp1 = _ix(B, w, 2, 0, 0) * _iz(B, u, 1, 0, 1)
    sn = _ix(B, w, 1, 0, 0) * _iz(B, u, 0, 0, 1)
    sm1 = _ix(B, w, 0, 0, 0) * _iz(B, u, -1, 0, 1)
    sm2 = _ix(B, w, -1, 0, 0) * _iz(B, u, -2, 0, 1)
    duwdx = _d4(sp1, sn, sm1, sm2, dx)
    sp1 = _iy(B, w, 0, 2, 0) * _iz(B, v, 0, 1, 1)
    sn = _iy(B, w, 0, 1, 0) * _iz(B, v, 0, 0, 1)
    sm1 = _iy(B, w, 0, 0, 0) * _iz(B, v, 0, -1, 1)
    sm2 = _iy(B, w, 0, -1, 0) * _iz(B, v, 0, -2, 1)
    dvwdy = _d4(sp1, sn, sm1, sm2, dy)
    wstar[B.sl()] = B.at(woo) - dt * alpha * (duwdx + dvwdy + dw2dz)


def _lap4(B, f, dx, dy, dz):
    """kernels.py:177-186  4th-order Laplacian, summed x + y + z."""
    lx = (-B.at(f, 2, 0, 0) + 16.0 * B.at(f, 1, 0, 0) - 30.0 * B.at(f)
          + 16.0 * B.at(f, -1, 0, 0) - B.at(f, -2, 0, 0)) / (12.0 * dx * dx)
    ly = (-B.at(f, 0, 2, 0) + 16.0 * B.at(f, 0, 1, 0) - 30.0 * B.at(f)
          + 16.0 * B.at(f, 0, -1, 0) - B.at(f, 0, -2, 0)) / (12.0 * dy * dy)
    lz = (-B.at(f, 0, 0, 2) + 16.0 * B.at(f, 0, 0, 1) - 30.0 * B.at(f)
          + 16.0 * B.at(f, 0, 0, -1) - B.at(f, 0, 0, -2)) / (12.0 * dz * dz)
    return lx + ly + lz


def _diff4th(c, vis, star, voff, dx, dy, dz, dt, alpha, npl, dims):
    B = _box(npl, dims)
    diff = B.at(vis, *voff) * _lap4(B, c, dx, dy, dz)
    star[B.sl()] = B.at(star) + dt * alpha * diff


def diff4th_u(u, vis, ustar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:189-199  viscosity sampled at vis[i, j, k]."""
    _diff4th(u, vis, ustar, (0, 0, 0), dx, dy, dz, dt, alpha, npl, dims)


def diff4th_v(v, vis, vstar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:202-212  viscosity sampled at vis[i, j+1, k]."""
    _diff4th(v, vis, vstar, (0, 1, 0), dx, dy, dz, dt, alpha, npl, dims)


def diff4th_w(w, vis, wstar, dx, dy, dz, dt, alpha, npl, dims):
    """kernels.py:215-225  viscosity sampled at vis[i, j, k+1]."""
    _diff4th(w, vis, wstar, (0, 0, 1), dx, dy, dz, dt, alpha, npl, dims)


def _calvel(star, p, c, off, d, dt, alpha, npl, dims):
    B = _box(npl, dims)
    c[B.sl()] = B.at(star) - dt * alpha * (B.at(p, *off) - B.at(p)) / d


def calvel_u(ustar, p, u, dx, dt, alpha, npl, dims):
    """kernels.py:232-242  u = ustar - dt*alpha*(p[i+1]-p[i])/dx"""
    _calvel(ustar, p, u, (1, 0, 0), dx, dt, alpha, npl, dims)


def calvel_v(vstar, p, v, dy, dt, alpha, npl, dims):
    """kernels.py:245-253"""
    _calvel(vstar, p, v, (0, 1, 0), dy, dt, alpha, npl, dims)


def calvel_w(wstar, p, w, dz, dt, alpha, npl, dims):
    """kernels.py:256-264"""
    _calvel(wstar, p, w, (0, 0, 1), dz, dt, alpha, npl, dims)


def calmas(u, v, w, su, dx, dy, dz, dt, npl, dims):
    """kernels.py:271-289  su = div/dt; returns rmax = max|div*dx*dy*dz|
    (the atomic_max into a zeroed 1-element buffer, solver.py:169-175)."""
    B = _box(npl, dims)
    dudx = _d4(B.at(u, 1, 0, 0), B.at(u), B.at(u, -1, 0, 0), B.at(u, -2, 0, 0), dx)
    dvdy = _d4(B.at(v, 0, 1, 0), B.at(v), B.at(v, 0, -1, 0), B.at(v, 0, -2, 0), dy)
    dwdz = _d4(B.at(w, 0, 0, 1), B.at(w), B.at(w, 0, 0, -1), B.at(w, 0, 0, -2), dz)
    fact = dudx + dvdy + dwdz
    su[B.sl()] = fact / dt
    return max(0.0, float(np.abs(fact * dx * dy * dz).max(initial=0.0)))


def wale_gradients(u, v, w, dx, dy, dz, npl, dims):
    """kernels.py:311-327  the 9 velocity-gradient entries at cell centres
    (diagonal: face differences; off-diagonal: 4-point averaged centred
    differences)."""
    B = _box(npl, dims)
    A = B.at
    g = {}
    g["dudx"] = (A(u) - A(u, -1, 0, 0)) / dx
    g["dudy"] = (0.25 * (A(u, 0, 1, 0) + A(u, -1, 1, 0)
                         - A(u, 0, -1, 0) - A(u, -1, -1, 0))) / dy
    g["dudz"] = (0.25 * (A(u, 0, 0, 1) + A(u, -1, 0, 1)
                         - A(u, 0, 0, -1) - A(u, -1, 0, -1))) / dz
    g["dvdx"] = (0.25 * (A(v, 1, 0, 0) + A(v, 1, -1, 0)
                         - A(v, -1, 0, 0) - A(v, -1, -1, 0))) / dx
    g["dvdy"] = (A(v) - A(v, 0, -1, 0)) / dy
    g["dvdz"] = (0.25 * (A(v, 0, 0, 1) + A(v, 0, -1, 1)
                         - A(v, 0, 0, -1) - A(v, 0, -1, -1))) / dz
    g["dwdx"] = (0.25 * (A(w, 1, 0, 0) + A(w, 1, 0, -1)
                         - A(w, -1, 0, 0) - A(w, -1, 0, -1))) / dx
    g["dwdy"] = (0.25 * (A(w, 0, 1, 0) + A(w, 0, 1, -1)
                         - A(w, 0, -1, 0) - A(w, 0, -1, -1))) / dy
    g["dwdz"] = (A(w) - A(w, 0, 0, -1)) / dz
    return g


def wale_nut_from_gradients(g, delta, cw):
    """kernels.py:330-365  WALE eddy viscosity from the gradient tensor.

    The g2zx / g2zy last terms (``dwdz*dudz`` and ``dwdz*dudy``,
    kernels.py:336, :338) are kept verbatim: they differ from the textbook
    ``g_zk g_kx`` product (which would end in ``dwdz*dwdx`` / ``dwdz*dwdy``)."""
    dudx, dudy, dudz = g["dudx"], g["dudy"], g["dudz"]
    dvdx, dvdy, dvdz = g["dvdx"], g["dvdy"], g["dvdz"]
    dwdx, dwdy, dwdz = g["dwdx"], g["dwdy"], g["dwdz"]
    g2xx = dudx * dudx + dudy * dvdx + dudz * dwdx
    g2yy = dvdx * dudy + dvdy * dvdy + dvdz * dwdy
    g2zz = dwdx * dudz + dwdy * dvdz + dwdz * dwdz
    g2xy = dudx * dudy + dudy * dvdy + dudz * dwdy
    g2yx = dvdx * dudx + dvdy * dvdx + dvdz * dwdx
    g2xz = dudx * dudz + dudy * dvdz + dudz * dwdz
    g2zx = dwdx * dudx + dwdy * dvdx + dwdz * dudz
    g2yz = dvdx * dudz + dvdy * dvdz + dvdz * dwdz
    g2zy = dwdx * dudy + dwdy * dvdy + dwdz * dudy
    tr = (g2xx + g2yy + g2zz) / 3.0
    sdxx = g2xx - tr
    sdyy = g2yy - tr
    sdzz = g2zz - tr
    sdxy = 0.5 * (g2xy + g2yx)
    sdxz = 0.5 * (g2xz + g2zx)
    sdyz = 0.5 * (g2yz + g2zy)
    sdsd = (sdxx * sdxx + sdyy * sdyy + sdzz * sdzz
            + 2.0 * (sdxy * sdxy + sdxz * sdxz + sdyz * sdyz))
    sxx, syy, szz = dudx, dvdy, dwdz
    sxy = 0.5 * (dudy + dvdx)
    sxz = 0.5 * (dudz + dwdx)
    syz = 0.5 * (dvdz + dwdy)
    ss = (sxx * sxx + syy * syy + szz * szz
          + 2.0 * (sxy * sxy + sxz * sxz + syz * syz))
    ls = cw * delta
    with np.errstate(divide="ignore", invalid="ignore"):
        sdsd15 = np.power(sdsd, 1.5)
        denom = np.power(np.abs(ss), 2.5) + np.power(sdsd, 1.25)
        nut = np.where(sdsd > 1.0e-8, ls * ls * sdsd15 / denom, 0.0)
    return nut


def wale_delta(dx, dy, dz):
    """kernels.py:360  delta = (dx*dy*dz)^(1/3)"""
    return float(np.power(dx * dy * dz, 1.0 / 3.0))


def wale_viscosity(u, v, w, vis, dx, dy, dz, rrey, cw, npl, dims):
    """kernels.py:297-367  vis = 1/Re + nut on real cells."""
    g = wale_gradients(u, v, w, dx, dy, dz, npl, dims)
    nut = wale_nut_from_gradients(g, wale_delta(dx, dy, dz), cw)
    vis[_box(npl, dims).sl()] = rrey + nut


def maxabs_reduce(f, npl, dims):
    """kernels.py:374-381 (atomic_max into a zeroed buffer, solver.py:116)"""
    return max(0.0, float(np.abs(_box(npl, dims).at(f)).max(initial=0.0)))


def dtvisc_reduce(u, v, w, vis, dx, dy, dz, npl, dims):
    """kernels.py:384-402 (atomic_min into a 1e10-filled buffer,
    solver.py:127)."""
    B = _box(npl, dims)
    uc = 0.5 * (B.at(u) + B.at(u, -1, 0, 0))
    vc = 0.5 * (B.at(v) + B.at(v, 0, -1, 0))
    wc = 0.5 * (B.at(w) + B.at(w, 0, 0, -1))
    val = 1.0 / (np.abs(uc) / dx + np.abs(vc) / dy + np.abs(wc) / dz
                 + 2.0 * B.at(vis)
                 * (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz))
                 + 1.0e-30)
    return min(1.0e10, float(val.min(initial=1.0e10)))


def add_inplace(dst, src, npl, dims):
    """kernels.py:405-414  dst += src over real cells plus one ring;
    launched with dims = ncells + 2 (solver.py:198-201)."""
    B = _Box((npl - 1,) * 3, dims)
    dst[B.sl()] = B.at(dst) + B.at(src)


def fill_const(f, val, npl, dims):
    """kernels.py:442-449"""
    f[_box(npl, dims).sl()] = val


def noise_hash(i, j, k, seed):
    """kernels.py:461-468  uint32 hash -> r in [-1, 1].  i, j, k are
    full-array indices (the kernel adds npl before hashing)."""
    ui = np.asarray(i, dtype=np.uint32)
    uj = np.asarray(j, dtype=np.uint32)
    uk = np.asarray(k, dtype=np.uint32)
    h = ((ui * np.uint32(73856093)) ^ (uj * np.uint32(19349663))
         ^ (uk * np.uint32(83492791)) ^ np.uint32(seed & 0xFFFFFFFF))
    h = h ^ (h >> np.uint32(13))
    h = h * np.uint32(1274126177)
    h = h ^ (h >> np.uint32(16))
    return (h.astype(np.float64) / 4294967295.0) * 2.0 - 1.0, h


def init_noise(u, v, w, amp, seed, npl, dims):
    """kernels.py:452-471  u = v = amp*r, w = -amp*r (one r per cell)."""
    B = _box(npl, dims)
    i, j, k = np.meshgrid(np.arange(npl, npl + dims[0]),
                          np.arange(npl, npl + dims[1]),
                          np.arange(npl, npl + dims[2]), indexing="ij")
    with np.errstate(over="ignore"):
        r, _ = noise_hash(i, j, k, seed)
    u[B.sl()] = amp * r
    v[B.sl()] = amp * r
    w[B.sl()] = -amp * r


# ===========================================================================
# warpflow/bcs.py
# ===========================================================================

def _face(fld, axis, g, a0, na, b0, nb):
    """fld[g, a-range, b-range] along ``axis`` (perpendicular axes in
    ascending order), as _mk_lin/_mk_const index it (bcs.py:42-163)."""
    a, b = slice(a0, a0 + na), slice(b0, b0 + nb)
    if axis == 0:
        return (g, a, b)
    if axis == 1:
        return (a, g, b)
    return (a, b, g)


def bc_lin(fld, axis, high, g0, s0, sign, a0, na, b0, nb, nlay):
    """bcs.py:42-106  fld[ghost(m)] = sign*fld[src(m)], m < nlay.
    low side: ghost g0-m, source s0+m; high side: ghost g0+m, source s0-m."""
    for m in range(nlay):
        g, s = (g0 + m, s0 - m) if high else (g0 - m, s0 + m)
        fld[_face(fld, axis, g, a0, na, b0, nb)] = \
            sign * fld[_face(fld, axis, s, a0, na, b0, nb)]


def bc_const(fld, axis, high, g0, val, a0, na, b0, nb, nlay):
    """bcs.py:109-163  fld[ghost(m)] = val, m < nlay."""
    for m in range(nlay):
        g = g0 + m if high else g0 - m
        fld[_face(fld, axis, g, a0, na, b0, nb)] = val


def _n3(n):
    return (n, n, n) if np.isscalar(n) else tuple(int(x) for x in n)


def apply_u(fld, n, ulid):
    """bcs.py:201-223  u fill: W const 0, E const 0 from the wall face
    NPL+n-1, S/N even mirror, B odd mirror, T lid."""
    ni, nj, nk = _n3(n)
    N = NPL
    bc_const(fld, 0, False, N - 1, 0.0, N - 1, nj + 2, N - 1, nk + 2, 3)
    bc_const(fld, 0, True, N + ni - 1, 0.0, N - 1, nj + 2, N - 1, nk + 2, 3)
    bc_lin(fld, 1, False, N - 1, N, 1.0, N - 1, ni + 1, N - 1, nk + 2, 3)
    bc_lin(fld, 1, True, N + nj, N + nj - 1, 1.0, N - 1, ni + 1, N - 1, nk + 2, 3)
    bc_lin(fld, 2, False, N - 1, N, -1.0, N - 1, ni + 1, N - 1, nj + 2, 3)
    bc_const(fld, 2, True, N + nk, ulid, N - 1, ni + 1, N - 1, nj + 2, 3)


def apply_v(fld, n):
    """bcs.py:225-244"""
    ni, nj, nk = _n3(n)
    N = NPL
    bc_lin(fld, 0, False, N - 1, N, -1.0, N - 1, nj + 1, N - 1, nk + 2, 3)
    bc_lin(fld, 0, True, N + ni, N + ni - 1, -1.0, N - 1, nj + 1, N - 1, nk + 2, 3)
    bc_const(fld, 1, False, N - 1, 0.0, N - 1, ni + 2, N - 1, nk + 2, 3)
    bc_const(fld, 1, True, N + nj - 1, 0.0, N - 1, ni + 2, N - 1, nk + 2, 3)
    bc_lin(fld, 2, False, N - 1, N, -1.0, N - 1, ni + 2, N - 1, nj + 1, 3)
    bc_lin(fld, 2, True, N + nk, N + nk - 1, 1.0, N - 1, ni + 2, N - 1, nj + 1, 3)


def apply_w(fld, n):
    """bcs.py:246-265"""
    ni, nj, nk = _n3(n)
    N = NPL
    bc_lin(fld, 0, False, N - 1, N, -1.0, N - 1, nj + 2, N - 1, nk + 1, 3)
    bc_lin(fld, 0, True, N + ni, N + ni - 1, -1.0, N - 1, nj + 2, N - 1, nk + 1, 3)
    bc_lin(fld, 1, False, N - 1, N, 1.0, N - 1, ni + 2, N - 1, nk + 1, 3)
    bc_lin(fld, 1, True, N + nj, N + nj - 1, 1.0, N - 1, ni + 2, N - 1, nk + 1, 3)
    bc_const(fld, 2, False, N - 1, 0.0, N - 1, ni + 2, N - 1, nj + 2, 3)
    bc_const(fld, 2, True, N + nk - 1, 0.0, N - 1, ni + 2, N - 1, nj + 2, 3)


def apply_vis(fld, n, rrey):
    """bcs.py:267-293  one layer; W, E, B walls force the first real plane
    to 1/Re before mirroring it."""
    ni, nj, nk = _n3(n)
    N = NPL
    bc_const(fld, 0, False, N, rrey, N - 1, nj + 2, N - 1, nk + 2, 1)
    bc_lin(fld, 0, False, N - 1, N, 1.0, N - 1, nj + 2, N - 1, nk + 2, 1)
    bc_const(fld, 0, True, N + ni - 1, rrey, N - 1, nj + 2, N - 1, nk + 2, 1)
    bc_lin(fld, 0, True, N + ni, N + ni - 1, 1.0, N - 1, nj + 2, N - 1, nk + 2, 1)
    bc_lin(fld, 1, False, N - 1, N, 1.0, N - 1, ni + 2, N - 1, nk + 2, 1)
    bc_lin(fld, 1, True, N + nj, N + nj - 1, 1.0, N - 1, ni + 2, N - 1, nk + 2, 1)
    bc_const(fld, 2, False, N, rrey, N - 1, ni + 2, N - 1, nj + 2, 1)
    bc_lin(fld, 2, False, N - 1, N, 1.0, N - 1, ni + 2, N - 1, nj + 2, 1)
    bc_lin(fld, 2, True, N + nk, N + nk - 1, 1.0, N - 1, ni + 2, N - 1, nj + 2, 1)


def apply_pbound(fld, n):
    """bcs.py:295-302  one Neumann layer per side, perpendicular range
    [NPL-1, NPL+n+1], order x-lo, x-hi, y-lo, y-hi, z-lo, z-hi."""
    n3 = _n3(n)
    N = NPL
    for axis in (0, 1, 2):
        pa, pb = [n3[x] for x in (0, 1, 2) if x != axis]
        bc_lin(fld, axis, False, N - 1, N, 1.0, N - 1, pa + 3, N - 1, pb + 3, 1)
        bc_lin(fld, axis, True, N + n3[axis], N + n3[axis] - 1, 1.0,
               N - 1, pa + 3, N - 1, pb + 3, 1)


def apply_p_mirror(fld, n, npl):
    """bcs.py:304-311  full-depth mirror over the whole perpendicular
    range (initial state)."""
    n3 = _n3(n)
    N = NPL
    for axis, high in ((0, False), (0, True), (1, False), (1, True),
                       (2, False), (2, True)):
        pa, pb = [n3[x] for x in (0, 1, 2) if x != axis]
        g0, s0 = (N - 1, N) if not high else (N + n3[axis], N + n3[axis] - 1)
        bc_lin(fld, axis, high, g0, s0, 1.0, 0, pa + 2 * N, 0, pb + 2 * N, npl)


def apply_all(fields, n, ulid, npl=NPL):
    """bcs.py:313-323  initial fill."""
    if "u" in fields:
        apply_u(fields["u"], n, ulid)
    if "v" in fields:
        apply_v(fields["v"], n)
    if "w" in fields:
        apply_w(fields["w"], n)
    if "p" in fields:
        apply_p_mirror(fields["p"], n, npl)


def neumann_mirror(fld, side, n_of, npl, na, nb):
    """bcs.py:326-360  faces only: perpendicular ranges cover real cells, so
    edge and corner ghosts are never written."""
    axis, high = {"west": (0, False), "east": (0, True),
                  "south": (1, False), "north": (1, True),
                  "bottom": (2, False), "top": (2, True)}[side]
    a, b = slice(npl, npl + na), slice(npl, npl + nb)
    for m in range(npl):
        g, s = ((npl + n_of + m, npl + n_of - 1 - m) if high
                else (npl - 1 - m, npl + m))
        if axis == 0:
            fld[g, a, b] = fld[s, a, b]
        elif axis == 1:
            fld[a, g, b] = fld[a, s, b]
        else:
            fld[a, b, g] = fld[a, b, s]


# ===========================================================================
# warpflow/multigrid.py
# ===========================================================================

SIDES = ("west", "east", "south", "north", "bottom", "top")


class Level:
    """multigrid.py:207-220  arrays of one level, shape n3 + 2."""

    def __init__(self, n3, dtype=np.float64):
        self.n3 = tuple(int(x) for x in n3)
        self.shape = tuple(x + 2 for x in self.n3)
        z = lambda: np.zeros(self.shape, dtype=dtype)  # noqa: E731
        self.phi, self.rhs = z(), z()
        self.aw, self.ae, self.as_, self.an = z(), z(), z(), z()
        self.ab, self.at, self.ap = z(), z(), z()


def mgcoeff(lev, phys, dxl, dyl, dzl):
    """multigrid.py:23-67  face coefficients 1/h^2 on real cells, zeroed
    across physical faces; ap = -(aw+ae+as+an+ab+at).  The reference tests
    every axis against n3[0]-1; here each axis uses its own extent."""
    n0, n1, n2 = lev.n3
    dt = lev.phi.dtype
    i, j, k = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2),
                          indexing="ij")
    # 1/h^2 formed in float64 as the reference does, rounded once to the level
    # precision; ap is then summed in the level precision
    awv = np.full(lev.n3, 1.0 / (dxl * dxl), dtype=dt)
    aev = awv.copy()
    asv = np.full(lev.n3, 1.0 / (dyl * dyl), dtype=dt)
    anv = asv.copy()
    abv = np.full(lev.n3, 1.0 / (dzl * dzl), dtype=dt)
    atv = abv.copy()
    if phys["west"]:
        awv[i == 0] = 0.0
    if phys["east"]:
        aev[i == n0 - 1] = 0.0
    if phys["south"]:
        asv[j == 0] = 0.0
    if phys["north"]:
        anv[j == n1 - 1] = 0.0
    if phys["bottom"]:
        abv[k == 0] = 0.0
    if phys["top"]:
        atv[k == n2 - 1] = 0.0
    s = (slice(1, n0 + 1), slice(1, n1 + 1), slice(1, n2 + 1))
    lev.aw[s], lev.ae[s], lev.as_[s], lev.an[s] = awv, aev, asv, anv
    lev.ab[s], lev.at[s] = abv, atv
    lev.ap[s] = -(awv + aev + asv + anv + abv + atv)


def _nb_sum(lev, phi, B):
    """aw*W + ae*E + as*S + an*N + ab*B + at*T in the reference's order."""
    return (B.at(lev.aw) * B.at(phi, -1, 0, 0) + B.at(lev.ae) * B.at(phi, 1, 0, 0)
            + B.at(lev.as_) * B.at(phi, 0, -1, 0) + B.at(lev.an) * B.at(phi, 0, 1, 0)
            + B.at(lev.ab) * B.at(phi, 0, 0, -1) + B.at(lev.at) * B.at(phi, 0, 0, 1))


def _parity(n3):
    i, j, k = np.meshgrid(np.arange(n3[0]), np.arange(n3[1]), np.arange(n3[2]),
                          indexing="ij")
    return (i + j + k) % 2


def rbgs(lev, color):
    """multigrid.py:70-90  half-sweep on cells with (i+j+k+color) % 2 == 1
    (0-based real indices).  A half-sweep only reads the other colour, so the
    vectorised update equals the sequential one."""
    B = _Box((1, 1, 1), lev.n3)
    new = (B.at(lev.rhs) - _nb_sum(lev, lev.phi, B)) / B.at(lev.ap)
    mask = (_parity(lev.n3) + color) % 2 == 1
    view = lev.phi[B.sl()]
    view[mask] = new[mask]


def residual(lev):
    """rhs - A phi on real cells, with ap*phi added LAST as in _restrict
    (multigrid.py:110-117)."""
    B = _Box((1, 1, 1), lev.n3)
    return B.at(lev.rhs) - (_nb_sum(lev, lev.phi, B) + B.at(lev.ap) * B.at(lev.phi))


def restrict(f, c):
    """multigrid.py:93-119  rhs_c = 0.125 * sum over the 2x2x2 children of
    the fine residual (loop order di, dj, dk)."""
    res = residual(f)
    acc = np.zeros(c.n3, dtype=res.dtype)
    for di in range(2):
        for dj in range(2):
            for dk in range(2):
                acc = acc + res[di::2, dj::2, dk::2]
    c.rhs[1:-1, 1:-1, 1:-1] = acc * res.dtype.type(0.125)


def _prolong_axis(n):
    idx = np.arange(n)
    c0 = np.where(idx % 2 == 0, idx // 2 - 1, idx // 2)
    w0 = np.where(idx % 2 == 0, 0.25, 0.75)
    return c0, w0


def prolong(c, f):
    """multigrid.py:122-146  phi_f += sum over the 8 surrounding coarse
    cells of wx*wy*wz*phi_c, per-axis weights 1/4, 3/4; reads the coarse
    ring (faces mirrored by the last refill, edges/corners still 0)."""
    c0x, wx0 = _prolong_axis(f.n3[0])
    c0y, wy0 = _prolong_axis(f.n3[1])
    c0z, wz0 = _prolong_axis(f.n3[2])
    acc = np.zeros(f.n3, dtype=f.phi.dtype)
    for di in range(2):
        wxi = (wx0 if di == 0 else 1.0 - wx0).astype(f.phi.dtype)
        for dj in range(2):
            wyj = (wy0 if dj == 0 else 1.0 - wy0).astype(f.phi.dtype)
            for dk in range(2):
                wzk = (wz0 if dk == 0 else 1.0 - wz0).astype(f.phi.dtype)
                wgt = (wxi[:, None, None] * wyj[None, :, None]) * wzk[None, None, :]
                pc = c.phi[np.ix_(c0x + di + 1, c0y + dj + 1, c0z + dk + 1)]
                acc = acc + wgt * pc
    s = (slice(1, -1),) * 3
    f.phi[s] = f.phi[s] + acc


def copy_in(src, lev, npl):
    """multigrid.py:149-154"""
    n0, n1, n2 = lev.n3
    lev.rhs[1:n0 + 1, 1:n1 + 1, 1:n2 + 1] = \
        src[npl:npl + n0, npl:npl + n1, npl:npl + n2]


def copy_out_ring(lev, dst, npl):
    """multigrid.py:157-163  level-0 phi including its ring -> dst at
    offset npl-1."""
    s0, s1, s2 = lev.shape
    dst[npl - 1:npl - 1 + s0, npl - 1:npl - 1 + s1, npl - 1:npl - 1 + s2] = lev.phi


class Multigrid:
    """multigrid.py:223-337  V-cycle RBGS multigrid, fixed 4 V-cycles of
    V(6, 6) with 300 coarsest sweeps, no mean removal."""

    def __init__(self, n3, nlevels, dx, dy, dz, dtype=np.float64,
                 n_pre=6, n_post=6, n_coarse=300, n_vcycles=4):
        self.phys = {s: True for s in SIDES}   # single domain, no periodic side
        self.h0 = (dx, dy, dz)
        self.levels = []
        for l in range(nlevels):
            lev = Level(tuple(x >> l for x in n3), dtype)
            mgcoeff(lev, self.phys, dx * 2.0 ** l, dy * 2.0 ** l, dz * 2.0 ** l)
            self.levels.append(lev)
        self.n_pre, self.n_post = n_pre, n_post
        self.n_coarse, self.n_vcycles = n_coarse, n_vcycles

    def _refill(self, lev):
        """multigrid.py:264-287 (physical sides only)."""
        n0, n1, n2 = lev.n3
        dims = {"west": (n0, n1, n2), "east": (n0, n1, n2),
                "south": (n1, n0, n2), "north": (n1, n0, n2),
                "bottom": (n2, n0, n1), "top": (n2, n0, n1)}
        for side in SIDES:
            if self.phys[side]:
                n_of, na, nb = dims[side]
                neumann_mirror(lev.phi, side, n_of, NPL_MG, na, nb)

    def _relax(self, l, nsweeps):
        """multigrid.py:289-299"""
        lev = self.levels[l]
        for _ in range(nsweeps):
            rbgs(lev, 0)
            rbgs(lev, 1)
            self._refill(lev)

    def _vcycle(self):
        """multigrid.py:301-322"""
        nlev = len(self.levels)
        for l in range(nlev):
            if l > 0:
                restrict(self.levels[l - 1], self.levels[l])
                self.levels[l].phi[...] = 0.0
            self._relax(l, self.n_pre if l < nlev - 1 else self.n_coarse)
        for l in range(nlev - 2, -1, -1):
            prolong(self.levels[l + 1], self.levels[l])
            self._refill(self.levels[l])
            self._relax(l, self.n_post)

    def solve(self, su3, pp3, npl):
        """multigrid.py:324-337"""
        lev0 = self.levels[0]
        copy_in(su3, lev0, npl)
        pp3[...] = 0.0
        lev0.phi[...] = 0.0
        for _ in range(self.n_vcycles):
            self._vcycle()
        copy_out_ring(lev0, pp3, npl)


def field_residual(su3, x3, lev0, npl):
    """FP64 defect su - A x on the real cells of an npl-ghost field, with the
    level-0 coefficients (ap*x added last, multigrid.py:110-117)."""
    B = _Box((npl,) * 3, lev0.n3)
    C = _Box((1, 1, 1), lev0.n3)
    nb = (C.at(lev0.aw).astype(np.float64) * B.at(x3, -1, 0, 0)
          + C.at(lev0.ae).astype(np.float64) * B.at(x3, 1, 0, 0)
          + C.at(lev0.as_).astype(np.float64) * B.at(x3, 0, -1, 0)
          + C.at(lev0.an).astype(np.float64) * B.at(x3, 0, 1, 0)
          + C.at(lev0.ab).astype(np.float64) * B.at(x3, 0, 0, -1)
          + C.at(lev0.at).astype(np.float64) * B.at(x3, 0, 0, 1))
    return B.at(su3) - (nb + C.at(lev0.ap).astype(np.float64) * B.at(x3))


def solve_mixed(mg32, su3, pp3, npl, ncycles, defect_mean=False, result_mean=True,
                coef64=None):
    """FP32 V-cycles inside FP64 defect correction (the mixed form; no
    reference counterpart).  mg32: a float32 Multigrid.  coef64: a float64
    Level with the level-0 coefficients for the FP64 defect (built from mg32's
    level-0 spacing if omitted)."""
    lev0 = mg32.levels[0]
    if coef64 is None:
        coef64 = Level(lev0.n3)
        h = mg32.h0
        mgcoeff(coef64, {s: True for s in SIDES}, *h)
    pp3[...] = 0.0
    B = _Box((npl,) * 3, lev0.n3)
    mean = float(B.at(su3).mean()) if defect_mean else 0.0
    for c in range(ncycles):
        r = B.at(su3).copy() if c == 0 else field_residual(su3, pp3, coef64, npl)
        if defect_mean:
            r = r - mean
        lev0.rhs[1:-1, 1:-1, 1:-1] = r.astype(np.float32)
        lev0.phi[...] = 0.0
        mg32._vcycle()
        pp3[B.sl()] = B.at(pp3) + lev0.phi[1:-1, 1:-1, 1:-1].astype(np.float64)
    if result_mean:
        pp3[B.sl()] = B.at(pp3) - B.at(pp3).mean()
    # first ghost layer: Neumann face mirror
    n0, n1, n2 = lev0.n3
    a, b, c_ = slice(npl, npl + n0), slice(npl, npl + n1), slice(npl, npl + n2)
    pp3[npl - 1, b, c_] = pp3[npl, b, c_]
    pp3[npl + n0, b, c_] = pp3[npl + n0 - 1, b, c_]
    pp3[a, npl - 1, c_] = pp3[a, npl, c_]
    pp3[a, npl + n1, c_] = pp3[a, npl + n1 - 1, c_]
    pp3[a, b, npl - 1] = pp3[a, b, npl]
    pp3[a, b, npl + n2] = pp3[a, b, npl + n2 - 1]


# ===========================================================================
# warpflow/solver.py
# ===========================================================================

class Case:
    """The subset of the case configuration (warpflow/config.py:15-98) the
    step reads.  ``from_config`` accepts any object with the CaseConfig
    attribute layout."""

    def __init__(self, n3=(32, 32, 32), L=(1.0, 1.0, 1.0), reynolds=400.0,
                 sgs_enabled=True, cw=0.46, noise=0.01, init_type="rest",
                 ulid=1.0, mg_levels=3, max_outer_iters=30, div_tolerance=1e-6,
                 variable_dt=True, dt_init=3.0e-4, safety_factor=0.8,
                 dt_growth=1.1, seed=1):
        self.n3 = tuple(int(x) for x in n3)
        self.L = tuple(float(x) for x in L)
        self.reynolds = float(reynolds)
        self.sgs_enabled, self.cw = bool(sgs_enabled), float(cw)
        self.noise, self.init_type = float(noise), init_type
        self.ulid = float(ulid)
        self.mg_levels = int(mg_levels)
        self.max_outer_iters = int(max_outer_iters)
        self.div_tolerance = float(div_tolerance)
        self.variable_dt = bool(variable_dt)
        self.dt_init = float(dt_init)
        self.safety_factor, self.dt_growth = float(safety_factor), float(dt_growth)
        self.seed = int(seed)

    @classmethod
    def from_config(cls, cfg, seed=1):
        return cls(n3=(cfg.grid.nx, cfg.grid.ny, cfg.grid.nz),
                   L=(cfg.domain.lx, cfg.domain.ly, cfg.domain.lz),
                   reynolds=cfg.physics.reynolds,
                   sgs_enabled=cfg.physics.sgs_enabled, cw=cfg.physics.sgs_cw,
                   noise=cfg.initial.noise, init_type=cfg.initial.type,
                   ulid=cfg.boundaries.lid_velocity,
                   mg_levels=cfg.numerics.mg_levels,
                   max_outer_iters=cfg.numerics.max_outer_iters,
                   div_tolerance=cfg.numerics.div_tolerance,
                   variable_dt=cfg.time.variable_dt, dt_init=cfg.time.dt_init,
                   safety_factor=cfg.time.safety_factor,
                   dt_growth=cfg.time.dt_growth, seed=seed)


FIELDS = ("u", "v", "w", "p", "ustar", "vstar", "wstar",
          "uoo", "voo", "woo", "vis", "su", "pp")


class Solver:
    """solver.py:29-271  one RK2 fractional step per ``step()``, in the
    reference's exact launch order (warpflow/solver.py:217-271)."""

    def __init__(self, case: Case, init=True):
        self.case = case
        self.ncells = case.n3
        nx, ny, nz = case.n3
        self.dx = case.L[0] / nx
        self.dy = case.L[1] / ny
        self.dz = case.L[2] / nz
        self.rey = 1.0 / case.reynolds
        self.ulid = case.ulid
        shape = tuple(n + 2 * NPL for n in case.n3)
        self.f = {name: np.zeros(shape) for name in FIELDS}
        self.mg = Multigrid(case.n3, case.mg_levels, self.dx, self.dy, self.dz)
        self.dt = case.dt_init
        self.ctime = 0.0
        self.itime = 0
        self.rmax = 0.0
        self.cfl = 0.0
        self.outer_iters = 0
        if init:
            self.initialise()

    def initialise(self):
        """solver.py:67-78"""
        amp = self.case.noise if self.case.init_type == "rest" else 0.0
        if amp > 0.0:
            init_noise(self.f["u"], self.f["v"], self.f["w"], amp,
                       self.case.seed, NPL, self.ncells)
        self.f["vis"][...] = self.rey
        apply_all(self.f, self.ncells, self.ulid, NPL)

    def fill_bc(self, name):
        """solver.py:88-97 (halo exchange is a no-op in one domain)."""
        if name == "u":
            apply_u(self.f["u"], self.ncells, self.ulid)
        elif name == "v":
            apply_v(self.f["v"], self.ncells)
        elif name == "w":
            apply_w(self.f["w"], self.ncells)

    def op_dim(self, comp):
        """solver.py:99-110"""
        ni, nj, nk = self.ncells
        return {"u": (ni - 1, nj, nk), "v": (ni, nj - 1, nk),
                "w": (ni, nj, nk - 1)}.get(comp, (ni, nj, nk))

    def compute_dt(self):
        """solver.py:122-150"""
        f = self.f
        umax = maxabs_reduce(f["u"], NPL, self.ncells)
        vmax = maxabs_reduce(f["v"], NPL, self.ncells)
        wmax = maxabs_reduce(f["w"], NPL, self.ncells)
        dtvisc = dtvisc_reduce(f["u"], f["v"], f["w"], f["vis"],
                               self.dx, self.dy, self.dz, NPL, self.ncells)
        big = 1.0e10
        dtcfl = big
        for d, um in zip((self.dx, self.dy, self.dz), (umax, vmax, wmax)):
            if um > 0.0:
                dtcfl = min(dtcfl, d / um)
        dtmax = min(dtvisc, dtcfl)
        if dtmax <= 0.0 or not (dtmax < big):
            raise RuntimeError("time step collapsed (dtmax = %.3e)" % dtmax)
        self.cfl = self.dt / dtmax
        if self.case.variable_dt:
            self.dt = min(self.dt * self.case.dt_growth,
                          self.case.safety_factor * dtmax)
        return umax, vmax, wmax, dtvisc

    def calvel(self, alpha):
        """solver.py:155-166"""
        f = self.f
        calvel_u(f["ustar"], f["p"], f["u"], self.dx, self.dt, alpha, NPL, self.op_dim("u"))
        calvel_v(f["vstar"], f["p"], f["v"], self.dy, self.dt, alpha, NPL, self.op_dim("v"))
        calvel_w(f["wstar"], f["p"], f["w"], self.dz, self.dt, alpha, NPL, self.op_dim("w"))

    def calmas(self):
        """solver.py:168-176"""
        f = self.f
        self.rmax = calmas(f["u"], f["v"], f["w"], f["su"], self.dx, self.dy,
                           self.dz, self.dt, NPL, self.ncells)
        return self.rmax

    def mg_solve(self):
        self.mg.solve(self.f["su"], self.f["pp"], NPL)

    def project(self):
        """solver.py:178-212"""
        f = self.f
        eps = self.case.div_tolerance
        maxcy = self.case.max_outer_iters
        self.calvel(1.0)
        iters = 0
        self.rmax_history = []
        for it in range(1, maxcy + 1):
            iters = it
            self.calmas()
            self.rmax_history.append(self.rmax)
            if self.rmax < eps and it > 1:
                break
            if self.rmax > 10.0:
                raise RuntimeError("projection diverged: rmax = %.6e at itime %d"
                                   % (self.rmax, self.itime))
            self.mg_solve()
            add_inplace(f["p"], f["pp"], NPL, tuple(n + 2 for n in self.ncells))
            apply_pbound(f["p"], self.ncells)
            self.calvel(1.0)
        self.outer_iters = iters
        for name in ("u", "v", "w"):
            self.fill_bc(name)

    def sgs(self):
        """solver.py:227-237"""
        f = self.f
        if self.case.sgs_enabled:
            wale_viscosity(f["u"], f["v"], f["w"], f["vis"], self.dx, self.dy,
                           self.dz, self.rey, self.case.cw, NPL, self.ncells)
            apply_vis(f["vis"], self.ncells, self.rey)
        else:
            fill_const(f["vis"], self.rey, NPL, self.ncells)

    def rk_stage(self, alpha):
        """solver.py:254-265 (one stage: conv for u, v, w each after its
        fill, then diff for u, v, w each after its fill)."""
        f = self.f
        d3 = (self.dx, self.dy, self.dz)
        dt = self.dt
        conv = {"u": conv4th_u, "v": conv4th_v, "w": conv4th_w}
        diff = {"u": diff4th_u, "v": diff4th_v, "w": diff4th_w}
        for c in ("u", "v", "w"):
            self.fill_bc(c)
            conv[c](f["u"], f["v"], f["w"], f[c + "oo"], f[c + "star"],
                    *d3, dt, alpha, NPL, self.op_dim(c))
        for c in ("u", "v", "w"):
            self.fill_bc(c)
            diff[c](f[c], f["vis"], f[c + "star"], *d3, dt, alpha, NPL,
                    self.op_dim(c))

    def step(self):
        """solver.py:217-271"""
        self.itime += 1
        self.compute_dt()
        self.ctime += self.dt
        f = self.f
        for c in ("u", "v", "w"):
            f[c + "oo"][...] = f[c]
        self.sgs()
        for stage, alpha in enumerate(ALFARK):
            self.rk_stage(alpha)
            if stage + 1 < len(ALFARK):
                self.calvel(alpha)
        self.project()
        return {"ctime": self.ctime, "dt": self.dt, "rmax": self.rmax,
                "iters": self.outer_iters, "cfl": self.cfl}
