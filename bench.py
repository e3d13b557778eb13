#!/usr/bin/env python
"""RK2 step throughput of the lid-driven-cavity LES on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--levels 6]

Workload: unit cube, 256^3 cells, Re 1000, WALE (cw 0.46), rest + 1% seeded
noise, fixed dt 1e-4, projection to rmax < 1e-6 (at most 30 outer
iterations), RBGS V(6,6) multigrid with 300 coarsest sweeps and 4 V-cycles
per solve on 6 levels (8^3 coarsest).  Two forms run back to back on the same
GPU from the same initial state:

  mixed -- FP32 V-cycles inside FP64 defect correction, 2 cycles per solve
           by default (--mixed-cycles) (the headline value)
  fp64  -- the reference's plain FP64 V-cycles, 4 per solve

W untimed steps, then K steps timed between device synchronisations.
value = mixed-form Mcell/s = nx*ny*nz / (seconds per step) / 1e6.  The JSON
line also carries the FP64 form's rate, the mixed/FP64 speed-up, the velocity
rms deviation between the forms, a per-kernel table (HIP-event timings of
each hot kernel on the solver's own arrays, algorithmic bytes per launch),
the roofline of the dominant kernel and a CPU baseline (the C oracle, the
same FP64 schedule, one step of the same 256^3 case on the host cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "warpflow-les_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "RK2 step Mcell/s at 256^3, div<1e-6: FP64 vs FP32-MG+FP64 refine"


def parse(argv=None):
    ap = argparse.ArgumentParser(description="warpflow cavity RK2 step benchmark")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--reynolds", type=float, default=1000.0)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-kernels", action="store_true", help="skip the kernel table")
    ap.add_argument("--reps", type=int, default=20, help="launches per kernel timing")
    ap.add_argument("--workload", default="cavity256", choices=("cavity256", "ensemble64"),
                    help="cavity256: the headline 256^3 step; ensemble64: 64 uneven cavities "
                         "batched (mixed MG, graph-captured step) against sequential solves")
    ap.add_argument("--seq-steps", type=int, default=3,
                    help="ensemble64: timed steps of the sequential single-cavity baselines")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of tuning them first")
    ap.add_argument("--tune-cache", default=None,
                    help="JSON file of tuned launch shapes: read if it holds this grid and "
                         "device, else tuned and written (e.g. to profile a run with the "
                         "shapes an unprofiled run chose)")
    ap.add_argument("--probe", action="store_true",
                    help="only launch each hot kernel --reps times (hardware-counter runs)")
    ap.add_argument("--mixed-cycles", type=int, default=2,
                    help="FP64-refined FP32 V-cycles per pressure solve in the mixed form")
    return ap.parse_args(argv)


def make_cfg(a, precision):
    from warpflow.config import CaseConfig, validate
    c = CaseConfig(name="cavity%d_re%d" % (a.n, a.reynolds))
    c.grid.nx = c.grid.ny = c.grid.nz = a.n
    c.physics.reynolds = a.reynolds
    c.physics.sgs_enabled = True
    c.physics.sgs_cw = 0.46
    c.initial.noise = 0.01
    c.numerics.mg_levels = a.levels
    c.numerics.div_tolerance = 1e-6
    c.numerics.max_outer_iters = 30
    c.numerics.mg_precision = precision
    c.numerics.mg_mixed_cycles = a.mixed_cycles
    c.time.variable_dt = False
    c.time.dt_init = a.dt
    validate(c)
    return c


def run_form(a, precision):
    from warpflow import FlowSolver, build_topology
    from warpflow import autotune
    from warpflow.device import synchronize
    cfg = make_cfg(a, precision)
    sol = FlowSolver(cfg, build_topology(cfg), "hip:0")
    tuned = None
    if not a.no_autotune:  # launch shapes for this grid and form, untimed, arrays restored
        tuned = [[k, list(n3), v] for k, n3, v in
                 autotune.tune_solver(sol, reps=10, cache=a.tune_cache)["entries"]]
    for _ in range(a.warmup):
        sol.step()
    synchronize()
    t0 = time.perf_counter()
    infos = [sol.step() for _ in range(a.steps)]
    synchronize()
    sec = time.perf_counter() - t0
    sol.tuned = tuned
    return sol, infos, sec


class Timer:
    """HIP events on the stream the kernels are launched on (the null
    stream): average duration of `reps` back-to-back launches."""

    def __init__(self):
        from warpflow import _hip
        self.h = _hip
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        _hip.call("wf_event_create", C.byref(self.e0))
        _hip.call("wf_event_create", C.byref(self.e1))

    def avg_us(self, fn, reps):
        from warpflow.device import current_stream
        fn()  # warm
        self.h.call("wf_event_record", self.e0, current_stream())
        for _ in range(reps):
            fn()
        self.h.call("wf_event_record", self.e1, current_stream())
        ms = C.c_float(0.0)
        self.h.call("wf_event_elapsed_ms", C.byref(ms), self.e0, self.e1)
        return 1000.0 * ms.value / reps


def hot_kernels(a, s64, smx):
    """The hot kernels of one step on the solvers' own arrays, as
    (name, launch, algorithmic bytes per launch, launches per step FP64,
    launches per step mixed).  bytes = distinct arrays read + written (RMW
    twice) x element size x launch cells; launches per step are those of
    the schedule with 2 outer iterations (one pressure solve: 4 V-cycles in
    the FP64 form, --mixed-cycles in the mixed form).  The fused RBGS entry
    is one call of two full sweeps (phi -> tmp -> phi), counted per sweep."""
    from warpflow import _hip
    from warpflow import kernels as K
    from warpflow.device import copy, current_stream
    from warpflow.topology import NPL
    n = a.n
    cells = float(n) ** 3
    f = s64.f
    h = (s64.dx, s64.dy, s64.dz)
    dt = s64.dt
    st = current_stream()
    opu = s64._op_dim("u")
    out = [
        # 6 per step: 3 of them the first stage's variant, same bytes (its c_oo
        # read becomes a c_oo write)
        ("convdiff4th (u)", lambda: K.convdiff4th_u(opu, f["u"], f["v"], f["w"], f["uoo"],
                                                    f["vis"], f["ustar"], *h, dt, 0.5, NPL),
         48 * opu[0] * opu[1] * opu[2], 6, 6, 1),
        ("calvel3", lambda: K.calvel3(s64.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"],
                                      f["u"], f["v"], f["w"], *h, dt, 1.0, NPL),
         56 * cells, 1, 1, 1),
        ("calvel3+calmas", lambda: K.calvel3_calmas(s64.ncells, f["ustar"], f["vstar"],
                                                    f["wstar"], f["p"], f["u"], f["v"], f["w"],
                                                    f["su"], s64._buf, *h, dt, 1.0, NPL),
         64 * cells, 2, 2, 1),
        ("calmas", lambda: K.calmas(s64.ncells, f["u"], f["v"], f["w"], f["su"], s64._buf,
                                    *h, dt, NPL), 32 * cells, 0, 0, 1),
        ("checkdt", lambda: K.checkdt(s64.ncells, f["u"], f["v"], f["w"], f["vis"], s64._buf,
                                      *h, NPL), 32 * cells, 0, 0, 1),
        ("wale", lambda: K.wale_viscosity(s64.ncells, f["u"], f["v"], f["w"], f["vis"], *h,
                                          s64.rey, 0.46, NPL), 32 * cells, 0, 0, 1),
        ("wale+checkdt", lambda: K.wale_checkdt(s64.ncells, f["u"], f["v"], f["w"], f["vis"],
                                                s64._buf, *h, s64.rey, 0.46, NPL),
         40 * cells, 1, 1, 1),
        ("add_ring", lambda: K.add_inplace(tuple(x + 2 for x in s64.ncells), f["p"], f["pp"],
                                           NPL), 24 * (n + 2) ** 3, 1, 1, 1),
        # the plain copy: the step no longer copies u -> uoo (the first stage's
        # conv+diff stores uoo on its box); kept as the streaming reference
        ("copy (u -> uoo)", lambda: copy(f["uoo"], f["u"]), 2 * f["u"].phys_nbytes, 0, 0, 1),
        ("bc_fill (u)", lambda: s64.fill_bc("u"), 0, 15, 15, 1),
    ]
    nv = s64.mg.n_vcycles
    nm = a.mixed_cycles
    sweeps = s64.mg.n_pre + s64.mg.n_post
    for tag, mg, eb, cyc64, cycmx in (("fp64", s64.mg, 8, nv, 0), ("fp32", smx.mg, 4, 0, nm)):
        L0, L1 = mg.levels[0].desc(), mg.levels[1].desc()
        out.append(("rbgs full sweep L0 (%s)" % tag,
                    lambda L0=L0: _hip.call("wf_mg_relax", C.byref(L0), 2, st),
                    3 * eb * cells, sweeps * cyc64, sweeps * cycmx, 2))
        out.append(("restrict L0 (%s)" % tag,
                    lambda L0=L0, L1=L1: _hip.call("wf_mg_restrict", C.byref(L0), C.byref(L1),
                                                   st),
                    (2 * eb + 2 * eb / 8) * cells, cyc64, cycmx, 1))
        out.append(("prolong L0 (%s)" % tag,
                    lambda L0=L0, L1=L1: _hip.call("wf_mg_prolong", C.byref(L1), C.byref(L0),
                                                   st),
                    (2 * eb + eb / 8) * cells, cyc64, cycmx, 1))
    return out


def load_traffic():
    """HBM bytes per launch of each hot kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    same counter's reading on the plain copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_cavity256.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh).get("kernels", {})


def kernel_table(a, s64, smx):
    """HIP-event timing of every hot kernel (hot_kernels) on the null stream,
    the stream the solver launches on."""
    T = Timer()
    traffic = load_traffic()
    rows = []
    for name, fn, nbytes, ps64, psmx, per_call in hot_kernels(a, s64, smx):
        us = T.avg_us(fn, a.reps) / per_call
        row = {"kernel": name, "us": round(us, 2), "bytes": int(nbytes),
               "GBps": round(nbytes / us / 1e3, 1),
               "hbm_frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 3),
               "per_step_fp64": ps64, "per_step_mixed": psmx}
        if name in traffic:
            row["traffic"] = traffic[name]["bytes_per_launch"]
        rows.append(row)
    return rows


def probe(a):
    """Short run for hardware-counter passes: build both solvers (no steps),
    then launch every hot kernel a.reps times, in order, on 256^3 arrays."""
    from warpflow import FlowSolver, build_topology
    from warpflow.device import synchronize
    s64 = FlowSolver(make_cfg(a, "fp64"), build_topology(make_cfg(a, "fp64")), "hip:0")
    smx = FlowSolver(make_cfg(a, "mixed"), build_topology(make_cfg(a, "mixed")), "hip:0")
    synchronize()
    for name, fn, nbytes, _, _, per_call in hot_kernels(a, s64, smx):
        print("# probe %s: %d calls, %d algorithmic bytes per kernel launch"
              % (name, a.reps, nbytes), file=sys.stderr, flush=True)
        for _ in range(a.reps):
            fn()
        synchronize()
    return 0


def dominant(rows, key):
    best = max(rows, key=lambda r: r["us"] * r[key])
    return best


def cpu_baseline(a):
    from oracle import c_oracle as CO
    from oracle import np_oracle as O
    case = O.Case(n3=(a.n,) * 3, reynolds=a.reynolds, sgs_enabled=True, cw=0.46, noise=0.01,
                  mg_levels=a.levels, variable_dt=False, dt_init=a.dt)
    sol = CO.CSolver(case)
    sol.step()  # untimed: first touch of every array
    t0 = time.perf_counter()
    sol.step()
    sec = time.perf_counter() - t0
    return {"value": round(a.n ** 3 / sec / 1e6, 3), "unit": "Mcell/s", "cores": CO.threads(),
            "kind": "port",
            "sample": "step 2 of the same %d^3 Re-%d cavity (plain FP64 schedule, %d MG levels) "
                      "in the C oracle (oracle/c/wf_oracle.c), %.2f s"
                      % (a.n, a.reynolds, a.levels, sec)}


def deviation(sa, sb):
    out = {}
    for nm in ("u", "v", "w"):
        x = sa.f[nm].numpy()[3:-3, 3:-3, 3:-3]
        y = sb.f[nm].numpy()[3:-3, 3:-3, 3:-3]
        out[nm] = float(np.sqrt(np.mean((x - y) ** 2)))
    p = sa.f["p"].numpy()[3:-3, 3:-3, 3:-3]
    q = sb.f["p"].numpy()[3:-3, 3:-3, 3:-3]
    out["p_demeaned"] = float(np.sqrt(np.mean(((p - p.mean()) - (q - q.mean())) ** 2)))
    return out


def run_ensemble(a):
    """64 cavities, nx, ny, nz each from {32, 48, 64, 96} and Re uniform in
    [100, 1000] (seeded), WALE on, 1% noise (seed = member + 1), fixed dt
    5e-4, 4 multigrid levels: the batched ensemble (mixed MG, graph-captured
    step) against the same members solved one after another by single-cavity
    solvers (plain FP64, and mixed)."""
    from warpflow import FlowSolver, build_topology
    from warpflow.device import synchronize
    from warpflow.ensemble import Ensemble, draw_members, member_config
    specs = draw_members(64)
    dt = 5e-4
    ens = Ensemble(specs, dt, mg_levels=4, graph=True)
    cells = ens.cells()
    for _ in range(a.warmup):
        ens.step()
    t0 = time.perf_counter()
    iters = [ens.step()["iters"] for _ in range(a.steps)]
    sec = time.perf_counter() - t0
    ens.close()
    out = {"batched": cells * a.steps / sec / 1e6, "batched_ms": 1000.0 * sec / a.steps,
           "iters": [int(x) for x in np.max(iters, axis=0)]}
    for form in ("fp64", "mixed"):
        sols = []
        for sp in specs:
            cfg = member_config(sp, dt, mg_levels=4, precision=form)
            sols.append(FlowSolver(cfg, build_topology(cfg), "hip:0"))
        for sol in sols:
            sol.step()
        synchronize()
        t0 = time.perf_counter()
        for _ in range(a.seq_steps):
            for sol in sols:
                sol.step()
        synchronize()
        sec = time.perf_counter() - t0
        out["sequential_" + form] = cells * a.seq_steps / sec / 1e6
        out["sequential_%s_ms" % form] = 1000.0 * sec / a.seq_steps
        del sols
    print("# ensemble64: %.1f Mcells, batched %.3f ms/step (%.1f Mcell/s); sequential fp64 %.3f "
          "ms/step (%.1f Mcell/s), mixed %.3f ms/step (%.1f Mcell/s)"
          % (cells / 1e6, out["batched_ms"], out["batched"], out["sequential_fp64_ms"],
             out["sequential_fp64"], out["sequential_mixed_ms"], out["sequential_mixed"]),
          file=sys.stderr, flush=True)
    line = {"metric": "ensemble aggregate cell-updates/s, 64 uneven cavities",
            "value": round(out["batched"], 3), "unit": "Mcell/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(out["batched_ms"], 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 (MG V-cycle f32)",
            "data": "synthetic (seeded member draw and hash-noise initial states)",
            "config": {"workload": "ensemble64: 64 cavities, n from {32,48,64,96}, Re in "
                                   "[100,1000], WALE, fixed dt 5e-4, 4 MG levels",
                       "cells": cells, "graph": True},
            "sequential_fp64_value": round(out["sequential_fp64"], 3),
            "sequential_mixed_value": round(out["sequential_mixed"], 3),
            "speedup_vs_sequential_fp64": round(out["batched"] / out["sequential_fp64"], 3),
            "speedup_vs_sequential_mixed": round(out["batched"] / out["sequential_mixed"], 3),
            "max_outer_iters": out["iters"]}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    a = parse(argv)
    if a.gpus != 1:
        print("bench.py: this build runs one cavity on one GPU; --gpus %d is not supported "
              "(multi-GPU domain decomposition is out of scope)" % a.gpus, file=sys.stderr)
        return 2
    if a.probe:
        return probe(a)
    if a.workload == "ensemble64":
        return run_ensemble(a)
    forms = [f.strip() for f in a.forms.split(",") if f.strip()]
    res = {}
    for form in forms:
        sol, infos, sec = run_form(a, form)
        res[form] = {"sol": sol, "infos": infos, "sec": sec, "tuned": sol.tuned,
                     "mcell": a.n ** 3 * a.steps / sec / 1e6,
                     "ms": 1000.0 * sec / a.steps}
        print("# %s: %.3f ms/step, %.1f Mcell/s, iters %s, rmax %.3e" %
              (form, res[form]["ms"], res[form]["mcell"], [i["iters"] for i in infos],
               infos[-1]["rmax"]), file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else forms[0]
    line = {"metric": METRIC, "value": round(res[head]["mcell"], 3), "unit": "Mcell/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(res[head]["ms"], 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 (MG V-cycle f32)" if head == "mixed" else "f64",
            "data": "synthetic (seeded hash-noise initial state)",
            "config": {"workload": "lid-driven cavity %d^3, Re %g, WALE cw 0.46, fixed dt %g, "
                                   "rmax<1e-6, %s form" % (a.n, a.reynolds, a.dt, head),
                       "n": a.n, "mg_levels": a.levels,
                       "vcycles_per_solve": {"fp64": 4, "mixed": a.mixed_cycles},
                       "sweeps": "V(6,6)+300 coarsest"}}
    line["config"]["launch_shapes"] = {f: res[f]["tuned"] for f in res}  # [knob, n3, value]
    if "fp64" in res:
        line["fp64_value"] = round(res["fp64"]["mcell"], 3)
        line["fp64_ms_per_step"] = round(res["fp64"]["ms"], 4)
        if head == "mixed":
            line["speedup_mixed_vs_fp64"] = round(res["mixed"]["mcell"] / res["fp64"]["mcell"], 4)
            line["deviation_mixed_vs_fp64"] = deviation(res["mixed"]["sol"], res["fp64"]["sol"])
            line["rmax_trajectory"] = {f: [float("%.4g" % i["rmax"]) for i in res[f]["infos"]]
                                       for f in ("mixed", "fp64")}
    if not a.no_kernels and "fp64" in res and "mixed" in res:
        rows = kernel_table(a, res["fp64"]["sol"], res["mixed"]["sol"])
        line["kernels"] = rows
        dom = dominant(rows, "per_step_mixed" if head == "mixed" else "per_step_fp64")
        achieved = dom["bytes"] / dom["us"] / 1e3
        line["roofline"] = {"bound": "hbm", "kernel": dom["kernel"],
                            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": dom.get("traffic"),
                            "algorithmic_bytes": dom["bytes"]}
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
