"""Boundary-condition ghost fills for the lid-driven cavity (the reference's
warpflow/bcs.py): west/east/bottom no-slip walls, south/north slip planes,
top moving lid.

``BoundaryKernels`` keeps the reference's class, method names and
arguments.  Each ``apply_*`` is one C-ABI call that runs the reference's six
side fills (W, E, S, N, B, T) as three launches, one per axis, with the same
per-cell order of operations.  Ranges use each axis' own cell count; on a
cubic grid they are the reference's ranges.
"""
from __future__ import annotations

from . import _hip
from .device import DeviceArray, current_stream, grid_of
from .topology import NPL

_APPLY = {"u": 0, "v": 1, "w": 2, "vis": 3, "pbound": 4, "p_mirror": 5}


def _check_n(fld: DeviceArray, n):
    """The reference passes one cell count n (ncells[0]); the fill uses the
    array's own per-axis counts, which must include n on the x axis."""
    g = grid_of(fld, NPL)
    if n is not None and int(n if not hasattr(n, "__len__") else n[0]) != g.ni:
        raise ValueError("n=%s does not match the field's %d real cells" % (n, g.ni))
    return g


def bc_lin(fld, axis, high, g0, s0, sign, a0, na, b0, nb, nlay):
    """_mk_lin(axis, high) launch (bcs.py:42-106, :187-191)."""
    _hip.call("wf_bc_lin", fld.ptr, grid_of(fld, NPL), axis, int(bool(high)), g0, s0,
              float(sign), a0, na, b0, nb, nlay, current_stream())


def bc_const(fld, axis, high, g0, val, a0, na, b0, nb, nlay):
    """_mk_const(axis, high) launch (bcs.py:109-163, :193-196)."""
    _hip.call("wf_bc_const", fld.ptr, grid_of(fld, NPL), axis, int(bool(high)), g0,
              float(val), a0, na, b0, nb, nlay, current_stream())


class BoundaryKernels:
    """Cavity ghost fills for u, v, w, p and vis (bcs.py:166-323).

    Expected configuration: west/east/bottom = wall, south/north = slip,
    top = moving_lid; anything else raises ValueError (bcs.py:176-184)."""

    def __init__(self, bnd, device=None):
        self.device = device
        self.bnd = dict(bnd)
        ok = (self.bnd.get("south") == "slip" and self.bnd.get("north") == "slip"
              and self.bnd.get("top") == "moving_lid"
              and self.bnd.get("west") == "wall" and self.bnd.get("east") == "wall"
              and self.bnd.get("bottom") == "wall")
        if not ok:
            raise ValueError("unsupported boundary combination for the cavity "
                             "benchmark: %s" % self.bnd)

    def _fill(self, fld, which, param, n):
        g = _check_n(fld, n)
        _hip.call("wf_bc_fill", fld.ptr, g, _APPLY[which], float(param), current_stream())

    def apply_u(self, fld, n, ulid, alfabc=1):
        """bcs.py:201-223"""
        self._fill(fld, "u", ulid, n)

    def apply_v(self, fld, n, alfabc=1):
        """bcs.py:225-244"""
        self._fill(fld, "v", 0.0, n)

    def apply_w(self, fld, n, alfabc=1):
        """bcs.py:246-265"""
        self._fill(fld, "w", 0.0, n)

    def apply_uvw(self, u, v, w, n, ulid):
        """apply_u, apply_v, apply_w (bcs.py:201-265) back to back, as one
        launch per axis covering the three fields."""
        g = _check_n(u, n)
        key = lambda x: (x.ni, x.nj, x.nk, x.npl, x.sj, x.si)  # noqa: E731
        for f in (v, w):
            if key(grid_of(f, NPL)) != key(g) or f.lead != u.lead:
                raise ValueError("u, v and w must share one geometry")
        _hip.call("wf_bc_fill_uvw", u.ptr, v.ptr, w.ptr, g, float(ulid), current_stream())

    def refill_uvw(self, u, v, w, n, ulid):
        """apply_uvw on fields filled since their last change: the same result,
        computed on the lines of the ghost bands only."""
        g = _check_n(u, n)
        key = lambda x: (x.ni, x.nj, x.nk, x.npl, x.sj, x.si)  # noqa: E731
        for f in (v, w):
            if key(grid_of(f, NPL)) != key(g) or f.lead != u.lead:
                raise ValueError("u, v and w must share one geometry")
        _hip.call("wf_bc_refill_uvw", u.ptr, v.ptr, w.ptr, g, float(ulid), current_stream())

    def apply_vw(self, v, w, n):
        """apply_v, apply_w (bcs.py:225-265) back to back, as one launch per
        axis covering the two fields."""
        g = _check_n(v, n)
        key = lambda x: (x.ni, x.nj, x.nk, x.npl, x.sj, x.si)  # noqa: E731
        if key(grid_of(w, NPL)) != key(g) or w.lead != v.lead:
            raise ValueError("v and w must share one geometry")
        _hip.call("wf_bc_fill_uvw", None, v.ptr, w.ptr, g, 0.0, current_stream())

    def apply_vis(self, fld, n, rrey):
        """bcs.py:267-293"""
        self._fill(fld, "vis", rrey, n)

    def apply_pbound(self, fld, n):
        """bcs.py:295-302"""
        self._fill(fld, "pbound", 0.0, n)

    def apply_p_mirror(self, fld, n, npl):
        """bcs.py:304-311"""
        self._fill(fld, "p_mirror", npl, n)

    def apply(self, fields, ncells, ulid, npl=NPL):
        """bcs.py:313-323  initial fill of u, v, w and the p mirror."""
        n = ncells[0]
        if "u" in fields:
            self.apply_u(fields["u"], n, ulid, 1)
        if "v" in fields:
            self.apply_v(fields["v"], n, 1)
        if "w" in fields:
            self.apply_w(fields["w"], n, 1)
        if "p" in fields:
            self.apply_p_mirror(fields["p"], n, npl)
