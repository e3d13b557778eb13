"""Geometric multigrid Poisson solver (the reference's warpflow/multigrid.py).

Same class, attributes and methods as the reference -- ``Multigrid(cfg, topo,
device, dx, dy, dz)``, ``levels[l].{phi, rhs, aw..ap, n3, shape}``,
``_refill``, ``_relax``, ``_vcycle``, ``solve`` and the ``n_pre / n_post /
n_coarse / n_vcycles`` schedule -- with the V-cycle driven natively by the HIP
library (one C-ABI call per ``_vcycle`` / ``solve``).

The operator is matrix-free on the GPU; ``levels[l].aw`` etc. materialise the
coefficient arrays on first access for code that inspects them
(tests/test_mg.py:16-32 in the reference).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _hip
from .device import DeviceArray, current_stream, grid_of, zeros

NPL_MG = 1
SIDES = ("west", "east", "south", "north", "bottom", "top")
# coarsest levels solved as phi = M rhs: up to 4096 cells each row of M is
# summed serially by one thread (the form the batched solve shares); larger
# ones, up to 16384 cells (M = 2 GiB in FP64), in row x column chunks
COARSE_OPERATOR_MAX_CELLS = 16384
COARSE_OPERATOR_SERIAL_MAX_CELLS = 4096


class _Level:
    """multigrid.py:207-220: phi and rhs of shape n3 + 2 (one-cell ring)."""

    def __init__(self, n3, h3, device=None, dtype=np.float64):
        self.n3 = tuple(int(x) for x in n3)
        self.shape = tuple(x + 2 for x in self.n3)
        self.h3 = tuple(float(h) for h in h3)
        self.dtype = np.dtype(dtype)
        self.device = device
        self.phi = zeros(self.shape, self.dtype, device)
        self.rhs = zeros(self.shape, self.dtype, device)
        # second phi buffer for the fused ping-pong sweeps (never read as a
        # result: every relax leaves its result in phi)
        self.tmp = zeros(self.shape, self.dtype, device)
        self._coef = None
        self.cop = None        # coarse operator (coarsest level only)
        self.cop_sweeps = 0

    def desc(self) -> _hip.WfMgLevel:
        return _hip.WfMgLevel(self.phi.ptr, self.rhs.ptr, *self.n3,
                              self.dtype.itemsize, *self.h3, self.tmp.ptr,
                              self.cop.ptr if self.cop is not None else None,
                              self.cop_sweeps, 0)

    def build_coarse_operator(self, nsweeps):
        """Form M with phi = M rhs equal to nsweeps RBGS sweeps from phi = 0
        (the coarsest solve of every V-cycle), by running those sweeps on
        every unit rhs.  Returns False when the level is too large."""
        d = self.desc()
        nbytes = C.c_size_t(0)
        if _hip.lib.wf_mg_coarse_op_bytes(C.byref(d), C.byref(nbytes)) != 0:
            return False
        op = zeros(nbytes.value // self.dtype.itemsize, self.dtype, self.device)
        _hip.call("wf_mg_coarse_op_build", C.byref(d), int(nsweeps), op.ptr, current_stream())
        self.cop, self.cop_sweeps = op, int(nsweeps)
        return True

    def _coefficients(self):
        if self._coef is None:
            arrs = [zeros(self.shape, self.dtype, self.device) for _ in range(7)]
            d = self.desc()
            _hip.call("wf_mg_coeff", C.byref(d), *[a.ptr for a in arrs], current_stream())
            self._coef = arrs
        return self._coef

    aw = property(lambda self: self._coefficients()[0])
    ae = property(lambda self: self._coefficients()[1])
    as_ = property(lambda self: self._coefficients()[2])
    an = property(lambda self: self._coefficients()[3])
    ab = property(lambda self: self._coefficients()[4])
    at = property(lambda self: self._coefficients()[5])
    ap = property(lambda self: self._coefficients()[6])


class Multigrid:
    """multigrid.py:223-337  RBGS V-cycle multigrid, 7-point cell-centred
    Neumann Laplacian; fixed n_vcycles V(n_pre, n_post) cycles with n_coarse
    sweeps on the coarsest level, starting from zero, no mean removal."""

    def __init__(self, cfg, topo, device, dx, dy, dz, dtype=np.float64):
        self.cfg = cfg
        self.topo = topo
        self.device = device
        for i, s in enumerate(SIDES):
            if getattr(cfg.boundaries, s) == "periodic":
                raise NotImplementedError("periodic multigrid faces are not part of this build")
            if topo.neighbors["xyz"[i // 2]][i % 2] >= 0:
                raise NotImplementedError("multi-domain multigrid is not part of this build")
        self.phys = {s: True for s in SIDES}
        nb = topo.local_ncells
        self.levels = []
        for l in range(cfg.numerics.mg_levels):
            n3 = tuple(x >> l for x in nb)
            h3 = (dx * 2.0 ** l, dy * 2.0 ** l, dz * 2.0 ** l)
            self.levels.append(_Level(n3, h3, device, dtype))
        self.n_pre = 6
        self.n_post = 6
        self.n_coarse = 300
        self.n_vcycles = 4
        # the coarsest solve as a matrix-vector product (M = the n_coarse
        # sweeps from zero) instead of 2 * n_coarse dependent half-sweeps in
        # one workgroup: 80^3 on 3 levels has a 20^3 coarsest level, whose 300
        # sweeps took 0.67 ms per V-cycle against ~0.1 ms for the 512 MB M
        coarse = self.levels[-1]
        if len(self.levels) > 1 and np.prod(coarse.n3) <= COARSE_OPERATOR_MAX_CELLS:
            coarse.build_coarse_operator(self.n_coarse)

    def _descs(self):
        arr = (_hip.WfMgLevel * len(self.levels))()
        for i, lev in enumerate(self.levels):
            arr[i] = lev.desc()
        return arr

    def residual_max(self, l=0):
        """max|rhs - A phi| over the real cells of level l (the reference's
        residual norm, tests/test_mg.py:16-32)."""
        d = self.levels[l].desc()
        out = zeros(1, np.float64, self.device)
        _hip.call("wf_mg_residual_max", C.byref(d), out.ptr, current_stream())
        return float(out.numpy()[0])

    def _refill(self, lev):
        """multigrid.py:264-287  faces-only Neumann mirror of the ring."""
        d = lev.desc()
        _hip.call("wf_mg_refill", C.byref(d), current_stream())

    def _relax(self, l, nsweeps):
        """multigrid.py:289-299  nsweeps x (red, black, refill)."""
        d = self.levels[l].desc()
        _hip.call("wf_mg_relax", C.byref(d), int(nsweeps), current_stream())

    def _vcycle(self):
        """multigrid.py:301-322"""
        descs = self._descs()
        _hip.call("wf_mg_vcycle", descs, len(self.levels), self.n_pre, self.n_post,
                  self.n_coarse, current_stream())

    def solve(self, su3: DeviceArray, pp3: DeviceArray, npl: int):
        """multigrid.py:324-337  pp = MG(su): copy-in, zero pp and phi0,
        n_vcycles V-cycles, copy-out including the ring."""
        descs = self._descs()
        g = grid_of(pp3, npl)
        if su3.shape != pp3.shape or su3.pitch != pp3.pitch or su3.lead != pp3.lead:
            raise ValueError("su and pp must have one shape and memory layout")
        _hip.call("wf_mg_solve", descs, len(self.levels), su3.ptr, pp3.ptr, g, self.n_pre,
                  self.n_post, self.n_coarse, self.n_vcycles, current_stream())
