"""Launch-shape autotuner for the hot kernels.

The fused momentum kernels (``wf_convdiff4th``, the first stage's
three-component ``wf_convdiff3_save`` and the second stage's ``wf_convdiff3``
or ``wf_convdiff_vw``),
the fused calvel/calmas pass
(``wf_calvel3_calmas``) and the RBGS sweep kernels have launch-shape
parameters that change no result: the i-planes each workgroup streams
through (``cd_ich``, ``cd3_ich``, ``cdvw_ich``, ``cd3s2_ich``, ``proj_ich``, ``sweep_ich``,
``sweep2_ich``) and the (j, k) tile of the one-sweep kernel (``sweep_tile``:
1 = 8 x 64, 2 = 16 x 32, 3 = 16 x 64, 4 = 32 x 32) and of the two-sweep
kernel (``sweep2_tile``: 1 = 8 x 64, 2 = 16 x 32, and in FP32 also
3 = 16 x 64, 4 = 32 x 32); a negative ``sweep_ich`` (-2, -4) runs the one-sweep
kernel's one-pass variant, chunks of 2 or 4 planes whose loads are all issued
at once (for small levels); ``sweep_pairs`` picks, per level and precision,
one-sweep launches (1) or two-sweep launches (2) for pairs of sweeps.  The
library keeps one setting per (parameter, launch shape); ``tune_solver``
times every candidate with HIP events on a solver's own arrays -- the
momentum kernels on the solver's grid, the sweeps on each multigrid level
that uses them -- checks that each candidate's output equals the built-in
choice bit for bit (two-sweep launches against one-sweep launches), restores
the arrays and records the fastest.  Results can be cached in a JSON file
keyed by device and grid.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _hip
from .device import copy, current_stream, zeros, zeros_like

# knob ids (include/warpflow_hip.h)
(CD_ICH, SWEEP_TILE_F64, SWEEP_TILE_F32, SWEEP_ICH_F64, SWEEP_ICH_F32, SWEEP2_TILE_F64,
 SWEEP2_TILE_F32, SWEEP2_ICH_F64, SWEEP2_ICH_F32, SWEEP_PAIRS_F64, PROJ_ICH,
 SWEEP_PAIRS_F32, CD3_ICH, CDVW_ICH, CD3S2_ICH) = range(15)
CD_ICH_CANDIDATES = (0, 8, 16, 24, 32, 48, 64)
CD3_ICH_CANDIDATES = (0, 32, 43, 64, 86, 128, 256)
CDVW_ICH_CANDIDATES = (0, 16, 32, 43, 64, 86, 128)
CD3S2_ICH_CANDIDATES = (0, 32, 43, 64, 86, 128, 256)
SWEEP_TILE_CANDIDATES = (1, 2, 3, 4)
SWEEP_ICH_CANDIDATES = (0, 1, 2, 4, 8, 16, 32, -2, -4)  # negative: one-pass chunks
SWEEP2_TILE_CANDIDATES = {True: (1, 2), False: (1, 2, 3, 4)}  # by "is FP64"
# 43: six chunks of 256 planes, one round of three workgroups per CU for the
# FP32 tiles of one pair per thread at 256^3
SWEEP2_ICH_CANDIDATES = (0, 2, 4, 8, 16, 32, 43, 64)
PROJ_ICH_CANDIDATES = (0, 32, 64, 86, 128, 172, 256)
SWEEP_MIN_CELLS = 4096  # smaller levels relax in one LDS-resident workgroup


def set_shape(knob, n3, value) -> None:
    _hip.call("wf_tune_set", int(knob), *[int(x) for x in n3], int(value))


def get_shape(knob, n3) -> int:
    v = C.c_int(0)
    _hip.call("wf_tune_get", int(knob), *[int(x) for x in n3], C.byref(v))
    return v.value


def reset() -> None:
    _hip.call("wf_tune_reset")


def apply(entries) -> None:
    """entries: [(knob, (n0, n1, n2), value), ...] as tune_solver returns."""
    for knob, n3, value in entries:
        set_shape(knob, n3, value)


class _Timer:
    def __init__(self):
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        _hip.call("wf_event_create", C.byref(self.e0))
        _hip.call("wf_event_create", C.byref(self.e1))

    def us(self, fn, reps):
        fn()
        _hip.call("wf_event_record", self.e0, current_stream())
        for _ in range(reps):
            fn()
        _hip.call("wf_event_record", self.e1, current_stream())
        ms = C.c_float(0.0)
        _hip.call("wf_event_elapsed_ms", C.byref(ms), self.e0, self.e1)
        return 1000.0 * ms.value / reps

    def close(self):
        _hip.call("wf_event_destroy", self.e0)
        _hip.call("wf_event_destroy", self.e1)


def _device_name():
    buf = C.create_string_buffer(256)
    _hip.call("wf_get_device_name", buf, 256)
    return buf.value.decode(errors="replace")


def _key(sol):
    return "%s|%dx%dx%d|%s|L%d" % (_device_name(), *sol.ncells,
                                    np.dtype(sol.mg.levels[0].dtype).name, len(sol.mg.levels))


def _tune_convdiff(sol, T, reps, out):
    from . import kernels as K
    from .topology import NPL
    f = sol.f
    h = (sol.dx, sol.dy, sol.dz)
    opu = sol._op_dim("u")
    n3 = sol.ncells
    saved = zeros_like(f["ustar"])
    copy(saved, f["ustar"])

    def cd():
        K.convdiff4th_u(opu, f["u"], f["v"], f["w"], f["uoo"], f["vis"], f["ustar"], *h, sol.dt,
                        0.5, NPL)
    set_shape(CD_ICH, n3, 0)
    cd()
    want = f["ustar"].numpy()
    best = None
    for ich in CD_ICH_CANDIDATES:
        set_shape(CD_ICH, n3, ich)
        t = T.us(cd, reps)
        if not np.array_equal(f["ustar"].numpy(), want):
            raise RuntimeError("convdiff result depends on cd_ich=%d" % ich)
        out["timings_us"]["convdiff %s ich=%d" % ("x".join(map(str, n3)), ich)] = round(t, 2)
        if best is None or t < best[0]:
            best = (t, ich)
    set_shape(CD_ICH, n3, best[1])
    out["entries"].append((CD_ICH, tuple(n3), best[1]))
    copy(f["ustar"], saved)


def _tune_convdiff3(sol, T, reps, out):
    """The first stage's three-component pass: every output it writes is
    restored afterwards."""
    from . import kernels as K
    from .topology import NPL
    f = sol.f
    h = (sol.dx, sol.dy, sol.dz)
    n3 = sol.ncells
    names = ("ustar", "vstar", "wstar", "uoo", "voo", "woo")
    saved = {nm: zeros_like(f[nm]) for nm in names}
    for nm in names:
        copy(saved[nm], f[nm])

    def cd3():
        K.convdiff3_save(n3, f["u"], f["v"], f["w"], f["vis"], f["ustar"], f["vstar"],
                         f["wstar"], f["uoo"], f["voo"], f["woo"], *h, sol.dt, 0.5, NPL)
    set_shape(CD3_ICH, n3, 0)
    cd3()
    want = [f[nm].numpy() for nm in names[:3]]
    best = None
    for ich in CD3_ICH_CANDIDATES:
        set_shape(CD3_ICH, n3, ich)
        t = T.us(cd3, reps)
        if not all(np.array_equal(f[nm].numpy(), a) for nm, a in zip(names, want)):
            raise RuntimeError("convdiff3 result depends on cd3_ich=%d" % ich)
        out["timings_us"]["convdiff3 %s ich=%d" % ("x".join(map(str, n3)), ich)] = round(t, 2)
        if best is None or t < best[0]:
            best = (t, ich)
    set_shape(CD3_ICH, n3, best[1])
    out["entries"].append((CD3_ICH, tuple(n3), best[1]))
    for nm in names:
        copy(f[nm], saved[nm])


def _tune_convdiff_vw(sol, T, reps, out):
    """The second stage's v/w pass (w_pre as the solver holds it): vstar and
    wstar are restored afterwards."""
    from . import kernels as K
    from .topology import NPL
    f = sol.f
    if "w_pre" not in f:
        return
    h = (sol.dx, sol.dy, sol.dz)
    n3 = sol.ncells
    names = ("vstar", "wstar")
    saved = {nm: zeros_like(f[nm]) for nm in names}
    for nm in names:
        copy(saved[nm], f[nm])

    def vw():
        K.convdiff_vw(n3, f["u"], f["v"], f["w"], f["voo"], f["woo"], f["vis"], f["w_pre"],
                      f["vstar"], f["wstar"], *h, sol.dt, 1.0, NPL)
    set_shape(CDVW_ICH, n3, 0)
    vw()
    want = [f[nm].numpy() for nm in names]
    best = None
    for ich in CDVW_ICH_CANDIDATES:
        set_shape(CDVW_ICH, n3, ich)
        t = T.us(vw, reps)
        if not all(np.array_equal(f[nm].numpy(), a) for nm, a in zip(names, want)):
            raise RuntimeError("convdiff_vw result depends on cdvw_ich=%d" % ich)
        out["timings_us"]["convdiff_vw %s ich=%d" % ("x".join(map(str, n3)), ich)] = round(t, 2)
        if best is None or t < best[0]:
            best = (t, ich)
    set_shape(CDVW_ICH, n3, best[1])
    out["entries"].append((CDVW_ICH, tuple(n3), best[1]))
    for nm in names:
        copy(f[nm], saved[nm])


def _tune_convdiff3s2(sol, T, reps, out):
    """The second stage's three-component pass (w_pre as the solver holds
    it): ustar, vstar and wstar are restored afterwards."""
    from . import kernels as K
    from .topology import NPL
    f = sol.f
    if "w_pre" not in f:
        return
    h = (sol.dx, sol.dy, sol.dz)
    n3 = sol.ncells
    names = ("ustar", "vstar", "wstar")
    saved = {nm: zeros_like(f[nm]) for nm in names}
    for nm in names:
        copy(saved[nm], f[nm])

    def uvw():
        K.convdiff3(n3, f["u"], f["v"], f["w"], f["uoo"], f["voo"], f["woo"], f["vis"],
                    f["w_pre"], f["ustar"], f["vstar"], f["wstar"], *h, sol.dt, 1.0, NPL)
    set_shape(CD3S2_ICH, n3, 0)
    uvw()
    want = [f[nm].numpy() for nm in names]
    best = None
    for ich in CD3S2_ICH_CANDIDATES:
        set_shape(CD3S2_ICH, n3, ich)
        t = T.us(uvw, reps)
        if not all(np.array_equal(f[nm].numpy(), a) for nm, a in zip(names, want)):
            raise RuntimeError("convdiff3 (second stage) result depends on cd3s2_ich=%d" % ich)
        out["timings_us"]["convdiff3s2 %s ich=%d" % ("x".join(map(str, n3)), ich)] = round(t, 2)
        if best is None or t < best[0]:
            best = (t, ich)
    set_shape(CD3S2_ICH, n3, best[1])
    out["entries"].append((CD3S2_ICH, tuple(n3), best[1]))
    for nm in names:
        copy(f[nm], saved[nm])


def _tune_proj(sol, T, reps, out):
    from . import kernels as K
    from .topology import NPL
    f = sol.f
    n3 = sol.ncells
    names = ("u", "v", "w", "su")
    saved = {nm: zeros_like(f[nm]) for nm in names}
    for nm in names:
        copy(saved[nm], f[nm])

    def run():
        for nm in ("u", "v", "w"):
            copy(f[nm], saved[nm])
        sol._rbuf.zero_()
        K.calvel3_calmas(n3, f["ustar"], f["vstar"], f["wstar"], f["p"], f["u"], f["v"],
                         f["w"], f["su"], sol._rbuf, sol.dx, sol.dy, sol.dz, sol.dt, 1.0, NPL)
    t_copy = T.us(lambda: [copy(f[nm], saved[nm]) for nm in ("u", "v", "w")], reps)
    set_shape(PROJ_ICH, n3, 0)
    run()
    want = [f[nm].numpy() for nm in names] + [sol._rbuf.numpy()]
    best = None
    for ich in PROJ_ICH_CANDIDATES:
        set_shape(PROJ_ICH, n3, ich)
        t = T.us(run, reps) - t_copy
        got = [f[nm].numpy() for nm in names] + [sol._rbuf.numpy()]
        if not all(np.array_equal(a, b) for a, b in zip(got, want)):
            raise RuntimeError("calvel3_calmas result depends on proj_ich=%d" % ich)
        out["timings_us"]["calvel3_calmas %s ich=%d" % ("x".join(map(str, n3)), ich)] = round(t, 2)
        if best is None or t < best[0]:
            best = (t, ich)
    set_shape(PROJ_ICH, n3, best[1])
    out["entries"].append((PROJ_ICH, tuple(n3), best[1]))
    for nm in names:
        copy(f[nm], saved[nm])


def _tune_level(lev, T, reps, out):
    n3 = lev.n3
    f64 = np.dtype(lev.dtype) == np.float64
    ktile = SWEEP_TILE_F64 if f64 else SWEEP_TILE_F32
    kich = SWEEP_ICH_F64 if f64 else SWEEP_ICH_F32
    kpairs = SWEEP_PAIRS_F64 if f64 else SWEEP_PAIRS_F32
    keep_phi, keep_tmp = zeros_like(lev.phi), zeros_like(lev.tmp)
    copy(keep_phi, lev.phi)
    copy(keep_tmp, lev.tmp)
    d = lev.desc()

    def relax():
        copy(lev.phi, keep_phi)
        _hip.call("wf_mg_relax", C.byref(d), 2, current_stream())
    set_shape(ktile, n3, 0)
    set_shape(kich, n3, 0)
    set_shape(kpairs, n3, 0)
    relax()
    want = lev.phi.numpy()
    t_copy = T.us(lambda: copy(lev.phi, keep_phi), reps)
    best = None
    for tile in SWEEP_TILE_CANDIDATES:
        for ich in SWEEP_ICH_CANDIDATES:
            set_shape(ktile, n3, tile)
            set_shape(kich, n3, ich)
            t = (T.us(relax, reps) - t_copy) / 2
            if not np.array_equal(lev.phi.numpy(), want):
                raise RuntimeError("sweep result depends on tile=%d ich=%d" % (tile, ich))
            out["timings_us"]["sweep %s %s tile=%d ich=%d" % (
                "x".join(map(str, n3)), "f64" if f64 else "f32", tile, ich)] = round(t, 2)
            if best is None or t < best[0]:
                best = (t, tile, ich)
    set_shape(ktile, n3, best[1])
    set_shape(kich, n3, best[2])
    out["entries"] += [(ktile, tuple(n3), best[1]), (kich, tuple(n3), best[2])]

    # four sweeps: two one-sweep pairs against two two-sweep launches
    k2tile = SWEEP2_TILE_F64 if f64 else SWEEP2_TILE_F32
    k2ich = SWEEP2_ICH_F64 if f64 else SWEEP2_ICH_F32

    def relax4():
        copy(lev.phi, keep_phi)
        _hip.call("wf_mg_relax", C.byref(d), 4, current_stream())
    set_shape(kpairs, n3, 1)
    relax4()
    want4 = lev.phi.numpy()
    t_single = T.us(relax4, reps) - t_copy
    out["timings_us"]["4 sweeps %s %s one-sweep launches" % (
        "x".join(map(str, n3)), "f64" if f64 else "f32")] = round(t_single, 2)
    set_shape(kpairs, n3, 2)
    best2 = None
    try:
        for tile in SWEEP2_TILE_CANDIDATES[f64]:
            for ich in SWEEP2_ICH_CANDIDATES:
                set_shape(k2tile, n3, tile)
                set_shape(k2ich, n3, ich)
                t = T.us(relax4, reps) - t_copy
                if not np.array_equal(lev.phi.numpy(), want4):
                    raise RuntimeError("two-sweep result differs from one-sweep launches "
                                       "(tile=%d ich=%d)" % (tile, ich))
                out["timings_us"]["4 sweeps %s %s two-sweep tile=%d ich=%d" % (
                    "x".join(map(str, n3)), "f64" if f64 else "f32", tile, ich)] = round(t, 2)
                if best2 is None or t < best2[0]:
                    best2 = (t, tile, ich)
    finally:
        set_shape(k2tile, n3, 0)
        set_shape(k2ich, n3, 0)
    if best2[0] < t_single:
        set_shape(k2tile, n3, best2[1])
        set_shape(k2ich, n3, best2[2])
        out["entries"] += [(k2tile, tuple(n3), best2[1]), (k2ich, tuple(n3), best2[2]),
                           (kpairs, tuple(n3), 2)]
    else:
        set_shape(kpairs, n3, 1)
        out["entries"].append((kpairs, tuple(n3), 1))
    copy(lev.phi, keep_phi)
    copy(lev.tmp, keep_tmp)


def tune_solver(sol, reps=10, cache=None):
    """Tune on FlowSolver `sol` (its arrays are restored afterwards); returns
    {"entries": [(knob, n3, value), ...], "timings_us": {...}} and leaves
    the entries set."""
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            hit = json.load(fh).get(_key(sol))
        if hit:
            hit["entries"] = [(k, tuple(n3), v) for k, n3, v in hit["entries"]]
            apply(hit["entries"])
            return hit
    out = {"entries": [], "timings_us": {}}
    T = _Timer()
    try:
        _tune_convdiff(sol, T, reps, out)
        _tune_convdiff3(sol, T, reps, out)
        if getattr(sol, "fused_uvw", False):
            _tune_convdiff3s2(sol, T, reps, out)
        else:
            _tune_convdiff_vw(sol, T, reps, out)
        _tune_proj(sol, T, reps, out)
        for lev in sol.mg.levels:
            if lev.tmp is not None and int(np.prod(lev.n3)) > SWEEP_MIN_CELLS:
                _tune_level(lev, T, reps, out)
        _hip.call("wf_synchronize")
    except Exception:
        for knob, n3, _ in out["entries"]:
            set_shape(knob, n3, 0)
        raise
    finally:
        T.close()
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[_key(sol)] = out
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return out
