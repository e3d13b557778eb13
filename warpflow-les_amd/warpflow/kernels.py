"""Kernel handles of the per-step hot path (the reference's
warpflow/kernels.py), dispatched to the gfx950 HIP library.

Each handle keeps the reference kernel's name and input list, so the
reference's call sites read the same::

    launch(K.conv4th_u, dim=sol._op_dim("u"),
           inputs=[u, v, w, uoo, ustar, dx, dy, dz, dt, alpha, NPL])

``dim`` is the launch box; like the reference kernels, every handle offsets
the thread index by ``npl`` (the last input).  Arrays are ``DeviceArray``s of
the same shape.  The fused handles have no single reference kernel:
``convdiff4th_{u,v,w}`` (the conv/diff pair of one RK stage),
``convdiff3_save`` (the three pairs of the first stage in one pass),
``convdiff_vw`` with ``save_cross_ghosts`` (v's and w's pairs of the second
stage in one pass), ``convdiff3`` with ``convdiff4th_u_edges`` (the three
pairs of the second stage in one pass, u's two x-edge planes apart),
``calvel3`` (the three calvel launches of ``FlowSolver._calvel``) and
``checkdt`` (the four reductions of ``FlowSolver.compute_dt``).
"""
from __future__ import annotations

import numpy as np

from . import _hip
from .device import DeviceArray, current_stream, grid_of

NPL = 3


class Kernel:
    """A launchable kernel: ``kernel(dim, *inputs)`` or ``launch(kernel, ...)``."""

    def __init__(self, name, fn):
        self.name = name
        self._fn = fn

    def __call__(self, dim, *inputs, **options):
        d = tuple(int(x) for x in dim)
        if len(d) != 3:
            raise ValueError("%s: dim must have 3 entries, got %s" % (self.name, dim))
        return self._fn(d, *inputs, **options)

    def __repr__(self):
        return "<warpflow kernel %s>" % self.name


def launch(kernel, dim, inputs, device=None, stream=None):
    """Stand-in for ``wp.launch(kernel, dim=..., inputs=[...], device=...)``."""
    return kernel(dim, *inputs)


def _same(*arrs):
    """One shape and one memory layout (row pitch, alignment shift): a kernel
    addresses every array it is given through one wf_grid."""
    a0 = arrs[0]
    for a in arrs:
        if (not isinstance(a, DeviceArray) or a.shape != a0.shape or a.pitch != a0.pitch
                or a.lead != a0.lead):
            raise ValueError("arrays must be DeviceArrays of one shape and layout: %r"
                             % (arrs,))


def _conv(comp):
    def f(dim, u, v, w, coo, cstar, dx, dy, dz, dt, alpha, npl):
        _same(u, v, w, coo, cstar)
        _hip.call("wf_conv4th", comp, u.ptr, v.ptr, w.ptr, coo.ptr, cstar.ptr,
                  grid_of(cstar, npl), *dim, dx, dy, dz, dt, alpha, current_stream())
    return f


def _diff(comp):
    def f(dim, c, vis, cstar, dx, dy, dz, dt, alpha, npl):
        _same(c, vis, cstar)
        _hip.call("wf_diff4th", comp, c.ptr, vis.ptr, cstar.ptr, grid_of(cstar, npl),
                  *dim, dx, dy, dz, dt, alpha, current_stream())
    return f


def _convdiff(comp):
    def f(dim, u, v, w, coo, vis, cstar, dx, dy, dz, dt, alpha, npl):
        _same(u, v, w, coo, vis, cstar)
        _hip.call("wf_convdiff4th", comp, u.ptr, v.ptr, w.ptr, coo.ptr, vis.ptr,
                  cstar.ptr, grid_of(cstar, npl), *dim, dx, dy, dz, dt, alpha,
                  current_stream())
    return f


def _convdiff_save(comp):
    def f(dim, u, v, w, coo_save, vis, cstar, dx, dy, dz, dt, alpha, npl):
        """convdiff4th_c of the first RK stage: the base value is the component
        itself, written to coo_save on the operator box (the copy of
        solver.py:223-225 restricted to the cells the second stage reads)."""
        _same(u, v, w, coo_save, vis, cstar)
        _hip.call("wf_convdiff4th_save", comp, u.ptr, v.ptr, w.ptr, coo_save.ptr, vis.ptr,
                  cstar.ptr, grid_of(cstar, npl), *dim, dx, dy, dz, dt, alpha,
                  current_stream())
    return f


def _convdiff3_save(dim, u, v, w, vis, ustar, vstar, wstar, uoo, voo, woo, dx, dy, dz, dt,
                    alpha, npl):
    """convdiff4th_save_u, _v and _w in one pass, each over its operator box,
    launched on the real cells (dim = ncells): the first RK stage with the
    three fills done first, valid when u, v, w were filled since their last
    change."""
    _same(u, v, w, vis, ustar, vstar, wstar, uoo, voo, woo)
    if tuple(dim) != tuple(n - 2 * npl for n in u.shape):
        raise ValueError("convdiff3_save: dim %s is not the real-cell extent" % (dim,))
    _hip.call("wf_convdiff3_save", u.ptr, v.ptr, w.ptr, vis.ptr, ustar.ptr, vstar.ptr,
              wstar.ptr, uoo.ptr, voo.ptr, woo.ptr, grid_of(u, npl), dx, dy, dz, dt, alpha,
              current_stream())


def _convdiff_vw(dim, u, v, w, voo, woo, vis, w_pre, vstar, wstar, dx, dy, dz, dt, alpha,
                 npl):
    """convdiff4th_v and _w of the second RK stage in one pass, each over its
    operator box, launched on the real cells (dim = ncells) after the fills
    of v and w; v's z-term reads w_pre where w's fill writes (filled by
    save_cross_ghosts from w just before that fill)."""
    _same(u, v, w, voo, woo, vis, w_pre, vstar, wstar)
    if tuple(dim) != tuple(n - 2 * npl for n in u.shape):
        raise ValueError("convdiff_vw: dim %s is not the real-cell extent" % (dim,))
    _hip.call("wf_convdiff_vw", u.ptr, v.ptr, w.ptr, voo.ptr, woo.ptr, vis.ptr, w_pre.ptr,
              vstar.ptr, wstar.ptr, grid_of(u, npl), dx, dy, dz, dt, alpha, current_stream())


def _convdiff3(dim, u, v, w, uoo, voo, woo, vis, w_pre, ustar, vstar, wstar, dx, dy, dz, dt,
               alpha, npl):
    """convdiff4th_u, _v and _w of the second RK stage in one pass, each over
    its operator box, launched on the real cells (dim = ncells) after the
    fills of u, v and w; v's z-term reads w_pre as in convdiff_vw.  u's planes
    0 and ni-2 are not written: convdiff4th_u_edges writes them after u's fill
    and before the fills of v and w, when conv_u reads v and w."""
    _same(u, v, w, uoo, voo, woo, vis, w_pre, ustar, vstar, wstar)
    if tuple(dim) != tuple(n - 2 * npl for n in u.shape):
        raise ValueError("convdiff3: dim %s is not the real-cell extent" % (dim,))
    _hip.call("wf_convdiff3", u.ptr, v.ptr, w.ptr, uoo.ptr, voo.ptr, woo.ptr, vis.ptr,
              w_pre.ptr, ustar.ptr, vstar.ptr, wstar.ptr, grid_of(u, npl), dx, dy, dz, dt,
              alpha, current_stream())


def _convdiff_u_edges(dim, u, v, w, uoo, vis, ustar, dx, dy, dz, dt, alpha, npl):
    """convdiff4th_u on planes 0 and ni-2 of u's operator box (dim = ncells)."""
    _same(u, v, w, uoo, vis, ustar)
    if tuple(dim) != tuple(n - 2 * npl for n in u.shape):
        raise ValueError("convdiff4th_u_edges: dim %s is not the real-cell extent" % (dim,))
    _hip.call("wf_convdiff4th_u_edges", u.ptr, v.ptr, w.ptr, uoo.ptr, vis.ptr, ustar.ptr,
              grid_of(u, npl), dx, dy, dz, dt, alpha, current_stream())


def _save_cross_ghosts(dim, w, w_pre, npl):
    """w_pre <- w at the cells convdiff_vw's v z-term reads outside w's
    operator box (rows -1 and nj, columns -2, -1, nk-1, nk of every plane)."""
    _same(w, w_pre)
    _hip.call("wf_save_cross_ghosts", w.ptr, w_pre.ptr, grid_of(w, npl), current_stream())


def _calvel(comp):
    def f(dim, cstar, p, c, h, dt, alpha, npl):
        _same(cstar, p, c)
        _hip.call("wf_calvel", comp, cstar.ptr, p.ptr, c.ptr, grid_of(c, npl), *dim, h,
                  dt, alpha, current_stream())
    return f


def _calvel3(dim, ustar, vstar, wstar, p, u, v, w, dx, dy, dz, dt, alpha, npl):
    _same(ustar, vstar, wstar, p, u, v, w)
    g = grid_of(u, npl)
    if dim != (g.ni, g.nj, g.nk):
        raise ValueError("calvel3 runs over the real cells (dim = ncells)")
    _hip.call("wf_calvel3", ustar.ptr, vstar.ptr, wstar.ptr, p.ptr, u.ptr, v.ptr, w.ptr, g,
              dx, dy, dz, dt, alpha, current_stream())


def _calmas(dim, u, v, w, su, rmax, dx, dy, dz, dt, npl):
    _same(u, v, w, su)
    _hip.call("wf_calmas", u.ptr, v.ptr, w.ptr, su.ptr, rmax.ptr, grid_of(su, npl), *dim,
              dx, dy, dz, dt, current_stream())


def _calvel3_calmas(dim, ustar, vstar, wstar, p, u, v, w, su, rmax, dx, dy, dz, dt, alpha,
                    npl, rhs32=None, store_uvw=True):
    """calvel3 then calmas over the real cells in one pass (``_project`` runs
    them back to back, solver.py:186-197); same results as the two kernels.
    rhs32: an FP32 (n+2)^3 multigrid level array that also receives (float)su
    on its real cells.  store_uvw=False: u, v, w are read but not written
    (su, rmax, rhs32 the same) -- for an outer iteration that cannot exit the
    loop, whose velocities the next iteration overwrites unread."""
    _same(ustar, vstar, wstar, p, u, v, w, su)
    g = grid_of(u, npl)
    if dim != (g.ni, g.nj, g.nk):
        raise ValueError("calvel3_calmas runs over the real cells (dim = ncells)")
    if rhs32 is not None and (rhs32.shape != tuple(n + 2 for n in dim)
                              or np.dtype(rhs32.dtype) != np.float32 or rhs32.pitch
                              != rhs32.shape[2] or rhs32.lead != 0):
        raise ValueError("rhs32 must be a dense FP32 (n+2)^3 level array")
    _hip.call("wf_calvel3_calmas" if store_uvw else "wf_calmas_of_calvel3", ustar.ptr,
              vstar.ptr, wstar.ptr, p.ptr, u.ptr, v.ptr, w.ptr, su.ptr, rmax.ptr,
              rhs32.ptr if rhs32 is not None else None, g, dx, dy, dz, dt, alpha,
              current_stream())


def wale_delta(dx, dy, dz):
    """(dx*dy*dz)^(1/3) (kernels.py:360)."""
    return (dx * dy * dz) ** (1.0 / 3.0)


def _wale(dim, u, v, w, vis, dx, dy, dz, rrey, cw, npl):
    _same(u, v, w, vis)
    _hip.call("wf_wale", u.ptr, v.ptr, w.ptr, vis.ptr, grid_of(vis, npl), *dim, dx, dy, dz,
              rrey, cw, wale_delta(dx, dy, dz), current_stream())


def _wale_checkdt(dim, u, v, w, vis, out4, dx, dy, dz, rrey, cw, npl):
    """checkdt's four reductions into out4, then WALE into vis, in one pass
    over the real cells (same values as the two kernels)."""
    _same(u, v, w, vis)
    g = grid_of(vis, npl)
    if dim != (g.ni, g.nj, g.nk) or out4.size < 4:
        raise ValueError("wale_checkdt runs over the real cells into a 4-element buffer")
    _hip.call("wf_wale_checkdt", u.ptr, v.ptr, w.ptr, vis.ptr, out4.ptr, g, dx, dy, dz, rrey,
              cw, wale_delta(dx, dy, dz), current_stream())


def _maxabs(dim, f, out, npl):
    _hip.call("wf_maxabs", f.ptr, out.ptr, grid_of(f, npl), *dim, current_stream())


def _dtvisc(dim, u, v, w, vis, out, dx, dy, dz, npl):
    _same(u, v, w, vis)
    _hip.call("wf_dtvisc", u.ptr, v.ptr, w.ptr, vis.ptr, out.ptr, grid_of(vis, npl), *dim,
              dx, dy, dz, current_stream())


def _checkdt(dim, u, v, w, vis, out4, dx, dy, dz, npl):
    _same(u, v, w, vis)
    g = grid_of(vis, npl)
    if dim != (g.ni, g.nj, g.nk) or out4.size < 4:
        raise ValueError("checkdt runs over the real cells into a 4-element buffer")
    _hip.call("wf_checkdt", u.ptr, v.ptr, w.ptr, vis.ptr, out4.ptr, g, dx, dy, dz,
              current_stream())


def _add_inplace(dim, dst, src, npl):
    _same(dst, src)
    g = grid_of(dst, npl)
    if dim != (g.ni + 2, g.nj + 2, g.nk + 2):
        raise ValueError("add_inplace covers the real cells plus one ring (dim = n+2)")
    _hip.call("wf_add_ring", dst.ptr, src.ptr, g, current_stream())


def _fill_const(dim, f, val, npl):
    _hip.call("wf_fill_interior", f.ptr, float(val), grid_of(f, npl), *dim,
              current_stream())


def _init_noise(dim, u, v, w, amp, seed, npl):
    _same(u, v, w)
    _hip.call("wf_init_noise", u.ptr, v.ptr, w.ptr, float(amp), int(seed),
              grid_of(u, npl), *dim, current_stream())


conv4th_u = Kernel("conv4th_u", _conv(0))
conv4th_v = Kernel("conv4th_v", _conv(1))
conv4th_w = Kernel("conv4th_w", _conv(2))
diff4th_u = Kernel("diff4th_u", _diff(0))
diff4th_v = Kernel("diff4th_v", _diff(1))
diff4th_w = Kernel("diff4th_w", _diff(2))
convdiff4th_u = Kernel("convdiff4th_u", _convdiff(0))
convdiff4th_v = Kernel("convdiff4th_v", _convdiff(1))
convdiff4th_w = Kernel("convdiff4th_w", _convdiff(2))
convdiff4th_save_u = Kernel("convdiff4th_save_u", _convdiff_save(0))
convdiff4th_save_v = Kernel("convdiff4th_save_v", _convdiff_save(1))
convdiff4th_save_w = Kernel("convdiff4th_save_w", _convdiff_save(2))
convdiff3_save = Kernel("convdiff3_save", _convdiff3_save)
convdiff_vw = Kernel("convdiff_vw", _convdiff_vw)
save_cross_ghosts = Kernel("save_cross_ghosts", _save_cross_ghosts)
convdiff3 = Kernel("convdiff3", _convdiff3)
convdiff4th_u_edges = Kernel("convdiff4th_u_edges", _convdiff_u_edges)
calvel_u = Kernel("calvel_u", _calvel(0))
calvel_v = Kernel("calvel_v", _calvel(1))
calvel_w = Kernel("calvel_w", _calvel(2))
calvel3 = Kernel("calvel3", _calvel3)
calmas = Kernel("calmas", _calmas)
calvel3_calmas = Kernel("calvel3_calmas", _calvel3_calmas)
wale_viscosity = Kernel("wale_viscosity", _wale)
wale_checkdt = Kernel("wale_checkdt", _wale_checkdt)
maxabs_reduce = Kernel("maxabs_reduce", _maxabs)
dtvisc_reduce = Kernel("dtvisc_reduce", _dtvisc)
checkdt = Kernel("checkdt", _checkdt)
add_inplace = Kernel("add_inplace", _add_inplace)
fill_const = Kernel("fill_const", _fill_const)
init_noise = Kernel("init_noise", _init_noise)
