"""ctypes wrapper of the C oracle (oracle/c/wf_oracle.c).  TEST
INFRASTRUCTURE ONLY -- see oracle/__init__.py."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "libwf_oracle.so"
FIELD_IDS = {n: i for i, n in enumerate(
    ("u", "v", "w", "p", "ustar", "vstar", "wstar", "uoo", "voo", "woo", "vis", "su", "pp"))}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from . import build
            build.build()
        L = C.CDLL(str(LIB))
        D, I, P = C.c_double, C.c_int, C.c_void_p
        L.wfo_create.argtypes = [I, I, I, D, D, D, D, I, D, D, I, D, I, I, D, I, D, D, D]
        L.wfo_create.restype = P
        L.wfo_destroy.argtypes = [P]
        L.wfo_step.argtypes = [P, C.POINTER(D)]
        L.wfo_step.restype = I
        L.wfo_get_field.argtypes = [P, I, P]
        L.wfo_set_field.argtypes = [P, I, P]
        L.wfo_rmax_history.argtypes = [P, P, I]
        L.wfo_rmax_history.restype = I
        L.wfo_threads.restype = I
        L.wfo_set_threads.argtypes = [I]
        _lib = L
    return _lib


def threads() -> int:
    return lib().wfo_threads()


def set_threads(n: int) -> None:
    """OpenMP threads of the following steps."""
    lib().wfo_set_threads(int(n))


class CSolver:
    """The reference step schedule in C; same case fields as np_oracle.Case."""

    def __init__(self, case):
        self.case = case
        self.shape = tuple(n + 6 for n in case.n3)
        self._h = lib().wfo_create(*case.n3, *case.L, case.reynolds, int(case.sgs_enabled),
                                   case.cw, case.noise if case.init_type == "rest" else 0.0,
                                   case.seed, case.ulid, case.mg_levels, case.max_outer_iters,
                                   case.div_tolerance, int(case.variable_dt), case.dt_init,
                                   case.safety_factor, case.dt_growth)
        if not self._h:
            raise RuntimeError("wfo_create failed")

    def step(self):
        info = (C.c_double * 5)()
        r = lib().wfo_step(self._h, info)
        if r == -1:
            raise RuntimeError("time step collapsed")
        if r == -2:
            raise RuntimeError("projection diverged")
        return {"ctime": info[0], "dt": info[1], "rmax": info[2], "iters": int(info[3]),
                "cfl": info[4]}

    def field(self, name):
        out = np.empty(self.shape)
        lib().wfo_get_field(self._h, FIELD_IDS[name], out.ctypes.data)
        return out

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape
        lib().wfo_set_field(self._h, FIELD_IDS[name], a.ctypes.data)

    def rmax_history(self):
        buf = (C.c_double * 64)()
        n = lib().wfo_rmax_history(self._h, buf, 64)
        return [buf[i] for i in range(n)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().wfo_destroy(h)
            self._h = None
