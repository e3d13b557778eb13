"""Mixed-precision pressure solve: FP32 V-cycles inside FP64 defect
correction, a drop-in for ``Multigrid`` (``solve(su, pp, npl)``).

Each cycle forms the FP64 defect ``r = su - A pp`` on the field, runs one
FP32 RBGS V-cycle (the same V(n_pre, n_post) schedule and n_coarse coarsest
sweeps as the plain form) on ``A e = r`` from ``e = 0``, and adds ``e`` to
``pp`` in FP64.  With ``rtol = 0`` it runs exactly ``max_cycles`` cycles
(default: ``numerics.mg_mixed_cycles`` of the case, 2); with as many cycles
as the plain form's ``n_vcycles`` (4) it reproduces the plain FP64 V-cycles
up to FP32 rounding of each correction.  With ``rtol > 0`` it stops once
``max|r| <= rtol * max|su|``.

Two cycles is the default because the projection only needs the 4th-order
divergence below ``div_tolerance``: at 256^3 it lands ~1e-9, far below 1e-6,
with any cycle count, and the velocity moves by ~6e-10 rms in 10 steps
relative to the 4-cycle FP64 solve (1 cycle: ~2e-8, 3 cycles: ~2e-11).

``defect_mean`` removes the FP64 mean of the defect before the FP32 cycle
(a consistent singular Neumann system); ``result_mean`` subtracts the FP64
mean of ``pp`` afterwards (a pressure gauge; velocities only see grad p).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _hip
from .device import DeviceArray, current_stream, grid_of, zeros
from .multigrid import Multigrid


class MixedMultigrid(Multigrid):
    def __init__(self, cfg, topo, device, dx, dy, dz, rtol=0.0, max_cycles=None,
                 defect_mean=False, result_mean=True):
        super().__init__(cfg, topo, device, dx, dy, dz, dtype=np.float32)
        self.rtol = float(rtol)
        if max_cycles is None:
            max_cycles = getattr(cfg.numerics, "mg_mixed_cycles", None)
        self.max_cycles = max_cycles
        self.defect_mean = bool(defect_mean)
        self.result_mean = bool(result_mean)
        self._work = None
        self.last_cycles = 0
        self.last_residual = 0.0

    def add_result(self, p3: DeviceArray, pp3: DeviceArray, npl: int):
        """p += pp over the real cells and the first ghost layer (add_inplace),
        with the gauge mean of a ``solve(defer_mean=True)`` subtracted on the
        way: the same p as a demeaned pp added, without rewriting pp."""
        g = grid_of(pp3, npl)
        if not getattr(self, "_deferred", False):
            _hip.call("wf_add_ring", p3.ptr, pp3.ptr, g, current_stream())
            return
        _hip.call("wf_add_ring_shift", p3.ptr, pp3.ptr, self._work.ptr + 2 * 8, g,
                  current_stream())

    def _workspace(self, g):
        nbytes = C.c_size_t(0)
        _hip.call("wf_mg_mixed_workspace_bytes", g, C.byref(nbytes))
        if self._work is None or self._work.nbytes < nbytes.value:
            self._work = zeros((nbytes.value + 7) // 8, np.float64, self.device)
        return self._work

    def solve(self, su3: DeviceArray, pp3: DeviceArray, npl: int, defer_mean: bool = False,
              rhs_ready: bool = False):
        """pp3 = solution for su3.  defer_mean (with result_mean): leave the
        gauge mean unsubtracted and apply it in ``add_result`` instead.
        rhs_ready: levels[0].rhs already holds (float)su3 on the real cells
        (written by calvel3_calmas), so the first defect is not formed again."""
        g = grid_of(pp3, npl)
        if su3.shape != pp3.shape or su3.pitch != pp3.pitch or su3.lead != pp3.lead:
            raise ValueError("su and pp must have one shape and memory layout")
        work = self._workspace(g)
        descs = self._descs()
        cycles = C.c_int(0)
        res = C.c_double(0.0)
        flags = (1 if self.defect_mean else 0) | (2 if self.result_mean else 0)
        self._deferred = bool(defer_mean and self.result_mean)
        if self._deferred:
            flags |= 4
        if rhs_ready:
            flags |= 8
        ncyc = self.n_vcycles if self.max_cycles is None else int(self.max_cycles)
        _hip.call("wf_mg_solve_mixed", descs, len(self.levels), su3.ptr, pp3.ptr, g,
                  self.n_pre, self.n_post, self.n_coarse, ncyc, self.rtol, flags,
                  work.ptr, work.nbytes, C.byref(cycles), C.byref(res), current_stream())
        self.last_cycles = cycles.value
        self.last_residual = res.value
