"""ctypes binding of ``libwarpflow_hip.so`` (C ABI in include/warpflow_hip.h).

The library is the only compute path: importing this module without the
built library, or using it on a machine without a GPU, raises -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("WARPFLOW_HIP_LIB", _PKG_ROOT / "lib" / "libwarpflow_hip.so"))

WF_EINVAL = 10001
WF_ESCRATCH = 10002
WF_STATIC_SCRATCH_BYTES = 1 << 20


class WfGrid(C.Structure):
    """wf_grid: real cells, ghost layers, element strides of j and i."""
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32),
                ("npl", C.c_int32), ("sj", C.c_int64), ("si", C.c_int64)]


class WfMember(C.Structure):
    """wf_member: one cavity of an ensemble."""
    _fields_ = [("n0", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32),
                ("seed", C.c_uint32), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double), ("rey", C.c_double), ("ulid", C.c_double),
                ("ls", C.c_double), ("noise", C.c_double), ("dt", C.c_double)]


class WfMgLevel(C.Structure):
    """wf_mg_level: one multigrid level (arrays with a one-cell ring)."""
    _fields_ = [("phi", C.c_void_p), ("rhs", C.c_void_p),
                ("n0", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32),
                ("prec", C.c_int32),
                ("h0", C.c_double), ("h1", C.c_double), ("h2", C.c_double),
                ("tmp", C.c_void_p), ("cop", C.c_void_p),
                ("cop_sweeps", C.c_int32), ("reserved", C.c_int32)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_Z = C.c_size_t
_G = WfGrid
_L = C.POINTER(WfMgLevel)

# name -> argtypes (all return int)
_SIGS = {
    "wf_device_count": [C.POINTER(C.c_int)],
    "wf_set_device": [_I],
    "wf_get_device_name": [C.c_char_p, _I],
    "wf_get_device_attr": [C.POINTER(C.c_int)] * 3,
    "wf_malloc": [C.POINTER(C.c_void_p), _Z],
    "wf_free": [_P],
    "wf_memcpy_h2d": [_P, _P, _Z, _P],
    "wf_memcpy_d2h": [_P, _P, _Z, _P],
    "wf_memcpy_d2d": [_P, _P, _Z, _P],
    "wf_copy": [_P, _P, _Z, _P],
    "wf_memset": [_P, _I, _Z, _P],
    "wf_fill_f64": [_P, _D, _Z, _P],
    "wf_fill_f32": [_P, C.c_float, _Z, _P],
    "wf_synchronize": [],
    "wf_batch_create": [_P, _I, _I, C.POINTER(C.c_void_p)],
    "wf_batch_destroy": [_P],
    "wf_batch_layout": [_P, _P, _P, _P, _P, _P, _P],
    "wf_batch_init_noise": [_P, _P, _P, _P, _P],
    "wf_batch_fill_vis": [_P, _P, _P],
    "wf_batch_wale": [_P, _P, _P, _P, _P, _P],
    "wf_batch_checkdt": [_P, _P, _P, _P, _P, _P, _P],
    "wf_batch_bc_fill": [_P, _P, _I, _P, _P],
    "wf_batch_convdiff": [_P, _I, _P, _P, _P, _P, _P, _P, C.c_double, _P],
    "wf_batch_convdiff_save": [_P, _I, _P, _P, _P, _P, _P, _P, C.c_double, _P],
    "wf_batch_calvel3": [_P, _P, _P, _P, _P, _P, _P, _P, C.c_double, _P, _P],
    "wf_batch_calmas": [_P, _P, _P, _P, _P, _P, _P, _P],
    "wf_batch_calvel3_calmas": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_double, _P, _P],
    "wf_batch_add_ring": [_P, _P, _P, _P, _P],
    "wf_batch_mg_solve_mixed": [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P],
    "wf_batch_project_mask": [_P, _P, C.c_double, _I, _P, _P, _P, _P],
    "wf_batch_set_int": [_P, _I, _I, _P],
    "wf_tune_set": [_I, _I, _I, _I, _I],
    "wf_tune_get": [_I, _I, _I, _I, C.POINTER(C.c_int)],
    "wf_tune_reset": [],
    "wf_graph_begin": [_P],
    "wf_graph_end": [_P, C.POINTER(C.c_void_p)],
    "wf_graph_launch": [_P, _P],
    "wf_graph_destroy": [_P],
    "wf_stream_create": [C.POINTER(C.c_void_p)],
    "wf_stream_destroy": [_P],
    "wf_stream_synchronize": [_P],
    "wf_event_create": [C.POINTER(C.c_void_p)],
    "wf_event_destroy": [_P],
    "wf_event_record": [_P, _P],
    "wf_event_elapsed_ms": [C.POINTER(C.c_float), _P, _P],
    # kernels.py
    "wf_conv4th": [_I, _P, _P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _D, _P],
    "wf_diff4th": [_I, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _D, _P],
    "wf_convdiff4th": [_I, _P, _P, _P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _D, _P],
    "wf_convdiff4th_save": [_I, _P, _P, _P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _D, _P],
    "wf_convdiff3_save": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_convdiff_vw": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_save_cross_ghosts": [_P, _P, _G, _P],
    "wf_convdiff3": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_convdiff4th_u_edges": [_P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_batch_convdiff3_save": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P],
    "wf_batch_convdiff3": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P],
    "wf_batch_convdiff4th_u_edges": [_P, _P, _P, _P, _P, _P, _P, _D, _P],
    "wf_batch_save_cross_ghosts": [_P, _P, _P, _P],
    "wf_batch_bc_fill_uvw": [_P, _P, _P, _P, _P, _P],
    "wf_calvel": [_I, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _P],
    "wf_calvel3": [_P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_calmas": [_P, _P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _P],
    "wf_calvel3_calmas": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_calmas_of_calvel3": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _P],
    "wf_wale": [_P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _D, _D, _D, _P],
    "wf_wale_checkdt": [_P, _P, _P, _P, _P, _G, _D, _D, _D, _D, _D, _D, _P],
    "wf_maxabs": [_P, _P, _G, _I, _I, _I, _P],
    "wf_dtvisc": [_P, _P, _P, _P, _P, _G, _I, _I, _I, _D, _D, _D, _P],
    "wf_checkdt": [_P, _P, _P, _P, _P, _G, _D, _D, _D, _P],
    "wf_add_ring": [_P, _P, _G, _P],
    "wf_add_ring_shift": [_P, _P, _P, _G, _P],
    "wf_fill_interior": [_P, _D, _G, _I, _I, _I, _P],
    "wf_init_noise": [_P, _P, _P, _D, C.c_int32, _G, _I, _I, _I, _P],
    # bcs.py
    "wf_bc_lin": [_P, _G, _I, _I, _I, _I, _D, _I, _I, _I, _I, _I, _P],
    "wf_bc_const": [_P, _G, _I, _I, _I, _D, _I, _I, _I, _I, _I, _P],
    "wf_bc_fill": [_P, _G, _I, _D, _P],
    "wf_bc_fill_uvw": [_P, _P, _P, _G, _D, _P],
    "wf_bc_refill_uvw": [_P, _P, _P, _G, _D, _P],
    # multigrid.py
    "wf_mg_coeff": [_L, _P, _P, _P, _P, _P, _P, _P, _P],
    "wf_mg_rbgs": [_L, _I, _P],
    "wf_mg_refill": [_L, _P],
    "wf_mg_relax": [_L, _I, _P],
    "wf_mg_restrict": [_L, _L, _P],
    "wf_mg_prolong": [_L, _L, _P],
    "wf_mg_copy_in": [_P, _G, _L, _P],
    "wf_mg_copy_out_ring": [_L, _P, _G, _P],
    "wf_mg_vcycle": [_L, _I, _I, _I, _I, _P],
    "wf_mg_solve": [_L, _I, _P, _P, _G, _I, _I, _I, _I, _P],
    "wf_mg_residual_max": [_L, _P, _P],
    "wf_mg_coarse_op_bytes": [_L, C.POINTER(C.c_size_t)],
    "wf_mg_coarse_op_build": [_L, _I, _P, _P],
    "wf_mg_coarse_op_apply": [_L, _P],
    # mixed-precision solve
    "wf_mg_mixed_workspace_bytes": [_G, C.POINTER(C.c_size_t)],
    "wf_mg_solve_mixed": [_L, _I, _P, _P, _G, _I, _I, _I, _I, _D, _I, _P, _Z,
                          C.POINTER(C.c_int), C.POINTER(C.c_double), _P],
    "wf_mg_solve_mixed_static": [_L, _I, _P, _P, _G, _I, _I, _I, _I, _D, _I, _P, _P, _P],
    "wf_field_mean": [_P, _G, _P, _Z, _P, _P],
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            "warpflow: HIP library %s is not built; run "
            "`python warpflow-les_amd/build.py` (hipcc, gfx950)" % LIB_PATH)
    lib = C.CDLL(str(LIB_PATH))
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.wf_error_string.argtypes = [C.c_int]
    lib.wf_error_string.restype = C.c_char_p
    return lib


lib = _load()


class HipError(RuntimeError):
    pass


def check(code: int, what: str = "") -> None:
    if code != 0:
        msg = lib.wf_error_string(code).decode(errors="replace")
        raise HipError("%s failed: %s (code %d)" % (what or "warpflow call", msg, code))


def call(name: str, *args) -> None:
    """Call C-ABI entry ``name`` and raise HipError on a non-zero return."""
    check(getattr(lib, name)(*args), name)


_device_ready = False


def ensure_device(device=None) -> None:
    """Select GPU ``device`` (an int, 'hip:N', 'cuda:N' or None for 0) once.
    Raises when no GPU is visible: the solver has no CPU path."""
    global _device_ready
    if _device_ready:
        return
    idx = 0
    if device not in (None, "auto"):
        s = str(device)
        if s == "cpu":
            raise RuntimeError("warpflow runs on an AMD GPU only (device='cpu' requested)")
        if ":" in s:
            idx = int(s.split(":", 1)[1])
        elif s.isdigit():
            idx = int(s)
    n = C.c_int(0)
    code = lib.wf_device_count(C.byref(n))
    if code != 0 or n.value == 0:
        raise RuntimeError("warpflow: no HIP GPU visible (hipGetDeviceCount -> %d, code %d)"
                           % (n.value, code))
    check(lib.wf_set_device(idx), "wf_set_device")
    _device_ready = True


def exported_symbols():
    """Names of every C-ABI entry point bound here (for the ABI tests)."""
    return sorted(list(_SIGS) + ["wf_error_string"])
