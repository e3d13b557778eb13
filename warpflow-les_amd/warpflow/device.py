"""Device arrays in GPU memory (HBM), owned by the HIP library.

``DeviceArray`` offers exactly the ``wp.array`` surface the reference uses:
``.numpy()`` (solver.py:119), ``.assign()`` (tests/test_mg.py:45), ``.zero_()``
(multigrid.py:312), ``.fill_()`` (solver.py:77) and ``.shape``; ``zeros`` and
``copy`` stand in for ``wp.zeros`` (solver.py:47) and ``wp.copy``
(solver.py:223).  Arrays are C-ordered with the last axis unit-stride.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _hip

_DTYPES = {np.dtype(np.float64), np.dtype(np.float32), np.dtype(np.int32),
           np.dtype(np.uint32), np.dtype(np.int64)}

# Stream all warpflow work is enqueued on (None = the null stream).
_stream = None


def current_stream():
    return _stream


def set_stream(stream) -> None:
    global _stream
    _stream = stream


class DeviceArray:
    """A C-ordered array in device memory, zero-initialised."""

    def __init__(self, shape, dtype=np.float64, device=None):
        _hip.ensure_device(device)
        self.shape = tuple(int(s) for s in (shape if np.ndim(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError("unsupported dtype %s" % self.dtype)
        self.size = int(np.prod(self.shape, dtype=np.int64))
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        _hip.call("wf_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value
        self.device = device

    # -- wp.array surface ------------------------------------------------ #
    @property
    def ndim(self):
        return len(self.shape)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            _hip.call("wf_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, _stream)
        return out

    def assign(self, src) -> "DeviceArray":
        if isinstance(src, DeviceArray):
            copy(self, src)
            return self
        a = np.ascontiguousarray(src, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError("assign: shape %s != %s" % (a.shape, self.shape))
        if self.nbytes:
            _hip.call("wf_memcpy_h2d", self.ptr, a.ctypes.data, self.nbytes, _stream)
        return self

    def zero_(self) -> "DeviceArray":
        return self.fill_(0)

    def fill_(self, value) -> "DeviceArray":
        if self.dtype == np.float64:
            _hip.call("wf_fill_f64", self.ptr, float(value), self.size, _stream)
        elif self.dtype == np.float32:
            _hip.call("wf_fill_f32", self.ptr, float(value), self.size, _stream)
        elif value == 0:
            _hip.call("wf_memset", self.ptr, 0, self.nbytes, _stream)
        else:
            self.assign(np.full(self.shape, value, dtype=self.dtype))
        return self

    def view(self, offset, shape) -> "DeviceArray":
        """A C-ordered array of `shape` over this array's memory from element
        `offset` on (no copy; keeps this array alive)."""
        shape = tuple(int(s) for s in shape)
        size = int(np.prod(shape, dtype=np.int64))
        if offset < 0 or offset + size > self.size:
            raise ValueError("view [%d, %d) outside %r" % (offset, offset + size, self))
        v = DeviceArray.__new__(DeviceArray)
        v.shape, v.dtype, v.size = shape, self.dtype, size
        v.nbytes = size * self.dtype.itemsize
        v.ptr = self.ptr + int(offset) * self.dtype.itemsize
        v.device = self.device
        v._base = self
        return v

    # -------------------------------------------------------------------- #
    def __del__(self):
        if getattr(self, "_base", None) is not None:
            return  # a view: the base array owns the memory
        p = getattr(self, "ptr", None)
        if p:
            try:
                _hip.lib.wf_free(p)
            except Exception:  # interpreter shutdown
                pass
            self.ptr = None

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype)


def zeros(shape, dtype=np.float64, device=None) -> DeviceArray:
    return DeviceArray(shape, dtype, device)


def empty(shape, dtype=np.float64, device=None) -> DeviceArray:
    return DeviceArray(shape, dtype, device)


def array(src, dtype=None, device=None) -> DeviceArray:
    a = np.asarray(src)
    d = DeviceArray(a.shape, dtype or a.dtype, device)
    return d.assign(a)


def copy(dst: DeviceArray, src: DeviceArray) -> None:
    """Whole-array device copy, ghosts included (wp.copy, solver.py:223-225)."""
    if dst.shape != src.shape or dst.dtype != src.dtype:
        raise ValueError("copy: %r <- %r" % (dst, src))
    if dst.nbytes:
        _hip.call("wf_copy", dst.ptr, src.ptr, dst.nbytes, _stream)


def synchronize() -> None:
    _hip.call("wf_synchronize")


def grid_of(arr: DeviceArray, npl: int) -> _hip.WfGrid:
    """wf_grid of a 3-D field with ``npl`` ghost layers per side."""
    if arr.ndim != 3 or arr.dtype != np.float64:
        raise ValueError("expected a 3-D float64 field, got %r" % arr)
    a, b, c = arr.shape
    ni, nj, nk = a - 2 * npl, b - 2 * npl, c - 2 * npl
    if min(ni, nj, nk) <= 0:
        raise ValueError("field %r too small for %d ghost layers" % (arr, npl))
    return _hip.WfGrid(ni, nj, nk, npl, c, b * c)
