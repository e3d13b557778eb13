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
    """A C-ordered array in device memory, zero-initialised.

    ``aligned=True`` (3-D arrays with ghost layers) pads every k-row to a
    whole number of 128-byte lines and shifts the array so that element
    k = ``aligned_k`` of every row starts a line: stencil kernels reading the
    real cells of a row then touch no partial line.  The logical view
    (``shape``, ``.numpy()``, ``.assign()``) is unchanged; kernels see the
    row pitch through ``wf_grid`` (``grid_of``)."""

    LINE_BYTES = 128

    def __init__(self, shape, dtype=np.float64, device=None, aligned=False, aligned_k=3):
        _hip.ensure_device(device)
        self.shape = tuple(int(s) for s in (shape if np.ndim(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError("unsupported dtype %s" % self.dtype)
        self.size = int(np.prod(self.shape, dtype=np.int64))
        self.nbytes = self.size * self.dtype.itemsize
        self.pitch = self.shape[-1] if self.shape else 1
        self.lead = 0
        if aligned and len(self.shape) == 3:
            line = self.LINE_BYTES // self.dtype.itemsize
            self.pitch = -(-self.shape[2] // line) * line
            self.lead = (line - aligned_k % line) % line
        rows = int(np.prod(self.shape[:-1], dtype=np.int64)) if self.shape else 1
        self.phys_size = self.lead + rows * self.pitch if self.shape else 1
        if self.padded:  # whole lines, so that copies run 16-byte vectors
            line = self.LINE_BYTES // self.dtype.itemsize
            self.phys_size = -(-self.phys_size // line) * line
        self.phys_nbytes = self.phys_size * self.dtype.itemsize
        p = C.c_void_p()
        _hip.call("wf_malloc", C.byref(p), max(self.phys_nbytes, 1))
        self.base = p.value
        self.ptr = self.base + self.lead * self.dtype.itemsize
        self.device = device

    @property
    def padded(self):
        return self.lead != 0 or (bool(self.shape) and self.pitch != self.shape[-1])

    # -- wp.array surface ------------------------------------------------ #
    @property
    def ndim(self):
        return len(self.shape)

    def numpy(self) -> np.ndarray:
        if not self.padded:
            out = np.empty(self.shape, dtype=self.dtype)
            if self.nbytes:
                _hip.call("wf_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, _stream)
            return out
        raw = np.empty(self.phys_size, dtype=self.dtype)
        _hip.call("wf_memcpy_d2h", raw.ctypes.data, self.base, self.phys_nbytes, _stream)
        a, b, c = self.shape
        rows = raw[self.lead:self.lead + a * b * self.pitch]
        return rows.reshape(a, b, self.pitch)[:, :, :c].copy()

    def assign(self, src) -> "DeviceArray":
        if isinstance(src, DeviceArray):
            copy(self, src)
            return self
        a = np.ascontiguousarray(src, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError("assign: shape %s != %s" % (a.shape, self.shape))
        if not self.padded:
            if self.nbytes:
                _hip.call("wf_memcpy_h2d", self.ptr, a.ctypes.data, self.nbytes, _stream)
            return self
        raw = np.zeros(self.phys_size, dtype=self.dtype)
        n0, n1, n2 = self.shape
        raw[self.lead:self.lead + n0 * n1 * self.pitch].reshape(n0, n1, self.pitch)[:, :, :n2] = a
        _hip.call("wf_memcpy_h2d", self.base, raw.ctypes.data, self.phys_nbytes, _stream)
        return self

    def zero_(self) -> "DeviceArray":
        return self.fill_(0)

    def fill_(self, value) -> "DeviceArray":
        """Every element, row padding included."""
        if self.dtype == np.float64:
            _hip.call("wf_fill_f64", self.base, float(value), self.phys_size, _stream)
        elif self.dtype == np.float32:
            _hip.call("wf_fill_f32", self.base, float(value), self.phys_size, _stream)
        elif value == 0:
            _hip.call("wf_memset", self.base, 0, self.phys_nbytes, _stream)
        else:
            self.assign(np.full(self.shape, value, dtype=self.dtype))
        return self

    def view(self, offset, shape) -> "DeviceArray":
        """A C-ordered array of `shape` over this array's memory from element
        `offset` on (no copy; keeps this array alive).  Unpadded arrays only."""
        if self.padded:
            raise ValueError("view of a padded array")
        shape = tuple(int(s) for s in shape)
        size = int(np.prod(shape, dtype=np.int64))
        if offset < 0 or offset + size > self.size:
            raise ValueError("view [%d, %d) outside %r" % (offset, offset + size, self))
        v = DeviceArray.__new__(DeviceArray)
        v.shape, v.dtype, v.size = shape, self.dtype, size
        v.nbytes = size * self.dtype.itemsize
        v.pitch, v.lead = (shape[-1] if shape else 1), 0
        v.phys_size, v.phys_nbytes = size, v.nbytes
        v.ptr = self.ptr + int(offset) * self.dtype.itemsize
        v.base = v.ptr
        v.device = self.device
        v._base = self
        return v

    # -------------------------------------------------------------------- #
    def __del__(self):
        if getattr(self, "_base", None) is not None:
            return  # a view: the base array owns the memory
        p = getattr(self, "base", None)
        if p:
            try:
                _hip.lib.wf_free(p)
            except Exception:  # interpreter shutdown
                pass
            self.base = None
            self.ptr = None

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s%s)" % (
            self.shape, self.dtype, ", pitch=%d" % self.pitch if self.padded else "")


def zeros(shape, dtype=np.float64, device=None, aligned=False) -> DeviceArray:
    return DeviceArray(shape, dtype, device, aligned=aligned)


def zeros_like(arr: DeviceArray) -> DeviceArray:
    """A zeroed array of `arr`'s shape, dtype and memory layout."""
    return DeviceArray(arr.shape, arr.dtype, arr.device, aligned=arr.padded)


def empty(shape, dtype=np.float64, device=None) -> DeviceArray:
    return DeviceArray(shape, dtype, device)


def array(src, dtype=None, device=None) -> DeviceArray:
    a = np.asarray(src)
    d = DeviceArray(a.shape, dtype or a.dtype, device)
    return d.assign(a)


def copy(dst: DeviceArray, src: DeviceArray) -> None:
    """Whole-array device copy, ghosts included (wp.copy, solver.py:223-225)."""
    if (dst.shape != src.shape or dst.dtype != src.dtype or dst.pitch != src.pitch
            or dst.lead != src.lead):
        raise ValueError("copy: %r <- %r" % (dst, src))
    if dst.phys_nbytes:
        _hip.call("wf_copy", dst.base, src.base, dst.phys_nbytes, _stream)


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
    return _hip.WfGrid(ni, nj, nk, npl, arr.pitch, b * arr.pitch)
