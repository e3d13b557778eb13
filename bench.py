#!/usr/bin/env python
"""RK2 step throughput of the lid-driven-cavity LES on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--levels 6]

Workload: unit cube, 256^3 cells, Re 1000, WALE (cw 0.46), rest + 1% seeded
noise, fixed dt 1e-4, projection to rmax < 1e-6 (at most 30 outer
iterations), RBGS V(6,6) multigrid with 300 coarsest sweeps on 6 levels
(8^3 coarsest).  Forms run back to back on the same GPU from the same
initial state:

  mixed  -- FP32 V-cycles inside FP64 defect correction, 2 cycles per solve
            (--mixed-cycles; the headline value)
  fp64   -- the reference's plain FP64 V-cycles, 4 per solve
  mixed4 -- the mixed form at the FP64 form's 4 cycles (the FP32 gain alone)

W untimed steps, then K steps timed between device synchronisations.
value = mixed-form Mcell/s = nx*ny*nz / (seconds per step) / 1e6.  The JSON
line also carries the other forms' rates and speed-ups, the velocity rms
deviation between the forms, a per-kernel table, the roofline of the kernel
that takes the most time per mixed step, the 80^3 production case
(cases/cavity_serial.yaml) in both forms, and CPU baselines (the C and NumPy
oracles: the same FP64 schedule on the host cores).

Kernel table: each hot kernel (or launch group: a cycle's level-0
smoothing, the three momentum launches of an RK stage) is timed with HIP
events on the stream the solver uses, after a 1 GiB streaming copy that
evicts the 256 MB Infinity Cache and the L2 (back-to-back repetitions on the
same arrays would time cache-resident data), and back to back.  The momentum
launches are also timed inside two real steps (events around each launch the
step makes): there they find part of their inputs left in the Infinity Cache
by the passes just before, so their in-step time is shorter than either
isolated timing, and that is the time the table ranks and the roofline uses;
a cycle's smoothing started cold matches its in-step time.  Algorithmic bytes per cell follow the
reference schedule's accounting (distinct arrays read + written x element
size; an RBGS sweep 32 B/cell in FP64, 16 in FP32; restriction and
prolongation 17 B per fine cell in FP64); traffic is HBM read + write bytes
per launch from rocprofv3 counters (profiles/pmc_traffic_cavity256.json).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "warpflow-les_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "RK2 step Mcell/s at 256^3, div<1e-6: FP64 vs FP32-MG+FP64 refine"
REPORT_80_MCELLS = 2.3  # 80^3 FP64 step on an RTX 5060 Laptop GPU (REPORT.md:101-103)


def parse(argv=None):
    ap = argparse.ArgumentParser(description="warpflow cavity RK2 step benchmark")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--reynolds", type=float, default=1000.0)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--forms", default="mixed,fp64,mixed4")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-kernels", action="store_true", help="skip the kernel table")
    ap.add_argument("--no-80", action="store_true", help="skip the 80^3 production case")
    ap.add_argument("--reps", type=int, default=10, help="timed launches per kernel")
    ap.add_argument("--workload", default="cavity256",
                    choices=("cavity256", "ensemble64", "poisson256"),
                    help="cavity256: the headline 256^3 step; ensemble64: 64 uneven cavities "
                         "batched (mixed MG, graph-captured step) against sequential solves; "
                         "poisson256: the standalone 256^3 Poisson solve to 1e-10 in both "
                         "forms with its multigrid kernels' GB/s")
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
    ap.add_argument("--summarize-trace", metavar="DB",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace database "
                         "(no GPU needed); --steps: the steps it covers")
    ap.add_argument("--summarize-sq", nargs="+", metavar="DB",
                    help="print per-wave SQ counters of the table rows from rocprofv3 --pmc "
                         "runs of --probe (no GPU needed)")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("READ_DB", "WRITE_DB", "OUT_JSON"),
                    help="write HBM bytes per launch from the two rocprofv3 --pmc runs of "
                         "--probe (no GPU needed)")
    return ap.parse_args(argv)


def make_cfg(a, precision, cycles=None):
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
    c.numerics.mg_mixed_cycles = a.mixed_cycles if cycles is None else cycles
    c.time.variable_dt = False
    c.time.dt_init = a.dt
    validate(c)
    return c


FORMS = {"mixed": ("mixed", None), "fp64": ("fp64", None), "mixed4": ("mixed", 4)}


def tune(a, sol):
    from warpflow import autotune
    if a.no_autotune:
        return None
    return [[k, list(n3), v] for k, n3, v in
            autotune.tune_solver(sol, reps=10, cache=a.tune_cache)["entries"]]


def run_form(a, form):
    from warpflow import FlowSolver, build_topology
    from warpflow.device import synchronize
    precision, cycles = FORMS[form]
    cfg = make_cfg(a, precision, cycles)
    sol = FlowSolver(cfg, build_topology(cfg), "hip:0")
    tuned = tune(a, sol)  # launch shapes for this grid and form, untimed, arrays restored
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
    stream)."""

    def __init__(self):
        from warpflow import _hip
        self.h = _hip
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        _hip.call("wf_event_create", C.byref(self.e0))
        _hip.call("wf_event_create", C.byref(self.e1))

    def _elapsed_us(self):
        ms = C.c_float(0.0)  # (waits for e1)
        self.h.call("wf_event_elapsed_ms", C.byref(ms), self.e0, self.e1)
        return 1000.0 * ms.value

    def warm_us(self, fn, reps):
        """Average of `reps` back-to-back calls (cache-warm)."""
        from warpflow.device import current_stream
        fn()
        self.h.call("wf_event_record", self.e0, current_stream())
        for _ in range(reps):
            fn()
        self.h.call("wf_event_record", self.e1, current_stream())
        return self._elapsed_us() / reps

    def cold_us(self, fn, reps, flush):
        """Average of `reps` calls, each timed alone after `flush` has
        streamed 1 GiB through the memory system."""
        from warpflow.device import current_stream
        fn()
        tot = 0.0
        for _ in range(reps):
            flush()
            self.h.call("wf_event_record", self.e0, current_stream())
            fn()
            self.h.call("wf_event_record", self.e1, current_stream())
            tot += self._elapsed_us()
        return tot / reps


class Flush:
    """Copy 512 MiB between two device buffers: 1 GiB of HBM traffic, which
    leaves none of a kernel's arrays in the Infinity Cache or the L2."""

    def __init__(self, nbytes=512 << 20):
        from warpflow.device import zeros
        self.a = zeros(nbytes // 8)
        self.b = zeros(nbytes // 8)

    def __call__(self):
        from warpflow.device import copy
        copy(self.b, self.a)


def hot_kernels(a, s64, smx):
    """The hot kernels and launch groups of one step on the solvers' own
    arrays: dicts with the launch (fn), the kernel it runs, launches and
    algorithmic bytes per call, and calls per step of each form with two
    outer iterations (one pressure solve: 4 V-cycles in the FP64 form,
    --mixed-cycles in the mixed form)."""
    from warpflow import _hip, autotune
    from warpflow import kernels as K
    from warpflow.device import copy, current_stream
    from warpflow.topology import NPL
    n = a.n
    cells = float(n) ** 3
    f = s64.f
    h = (s64.dx, s64.dy, s64.dz)
    dt = s64.dt
    st = current_stream()
    opd = {c: s64._op_dim(c) for c in "uvw"}

    n3 = s64.ncells

    def row(name, kernel, fn, bytes_per_call, per_step_fp64, per_step_mixed, launches=1):
        return {"name": name, "kernel": kernel, "fn": fn, "bytes": float(bytes_per_call),
                "launches": launches, "per_step_fp64": per_step_fp64,
                "per_step_mixed": per_step_mixed}

    out = [
        # the momentum launches of one step: the first stage's three-component
        # pass (reads u, v, w, vis; writes ustar, vstar, wstar, uoo, voo, woo),
        # then the second stage's pass of u's planes 0 and ni-2 (two launches;
        # reads u, v, w, uoo, vis, writes ustar on those planes) and its
        # three-component pass (reads u, v, w, uoo, voo, woo, vis; writes
        # ustar, vstar, wstar)
        row("convdiff3 stage 1 (u, v, w)", "convdiff3_save_kernel (conv4th + diff4th of u, v, w)",
            lambda: K.convdiff3_save(n3, f["u"], f["v"], f["w"], f["vis"], f["ustar"], f["vstar"],
                                     f["wstar"], f["uoo"], f["voo"], f["woo"], *h, dt, 0.5, NPL),
            80 * cells, 1, 1),
        row("convdiff u edges stage 2", "convdiff_kernel<0, false> (u's planes 0 and ni-2)",
            lambda: K.convdiff4th_u_edges(n3, f["u"], f["v"], f["w"], f["uoo"], f["vis"],
                                          f["ustar"], *h, dt, 1.0, NPL),
            48 * 2 * opd["u"][1] * opd["u"][2], 1, 1, 2),
        row("convdiff3 stage 2 (u, v, w)", "convdiff3_kernel (conv4th + diff4th of u, v, w)",
            lambda: K.convdiff3(n3, f["u"], f["v"], f["w"], f["uoo"], f["voo"], f["woo"],
                                f["vis"], f["w_pre"], f["ustar"], f["vstar"], f["wstar"], *h,
                                dt, 1.0, NPL),
            80 * cells, 1, 1),
        row("calvel3", "calvel3_kernel",
            lambda: K.calvel3(s64.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"], f["u"],
                              f["v"], f["w"], *h, dt, 1.0, NPL), 56 * cells, 1, 1),
        # the second outer iteration's pass stores u, v, w; the first's cannot
        # exit the loop and forms su and rmax only
        row("calvel3+calmas", "calvel3_calmas_kernel<true>",
            lambda: K.calvel3_calmas(s64.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"],
                                     f["u"], f["v"], f["w"], f["su"], s64._buf, *h, dt, 1.0,
                                     NPL), 64 * cells, 1, 1),
        row("calmas of calvel3 (no store)", "calvel3_calmas_kernel<false>",
            lambda: K.calvel3_calmas(s64.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"],
                                     f["u"], f["v"], f["w"], f["su"], s64._buf, *h, dt, 1.0,
                                     NPL, store_uvw=False), 40 * cells, 1, 1),
        row("wale+checkdt", "wale_checkdt_kernel",
            lambda: K.wale_checkdt(s64.ncells, f["u"], f["v"], f["w"], f["vis"], s64._buf, *h,
                                   s64.rey, 0.46, NPL), 40 * cells, 1, 1),
        row("add_ring", "add_kernel",
            lambda: K.add_inplace(tuple(x + 2 for x in s64.ncells), f["p"], f["pp"], NPL),
            24 * (n + 2) ** 3, 1, 1),
        # the plain copy: the step makes no whole-field copy any more (the first
        # stage's conv+diff stores uoo); kept as the streaming reference
        row("copy (u -> uoo)", "copy16_kernel", lambda: copy(f["uoo"], f["u"]),
            2 * f["u"].phys_nbytes, 0, 0),
        row("bc_fill (u)", "bc_axis_kernel", lambda: s64.fill_bc("u"), 0, 2, 2, 3),
    ]
    nv = s64.mg.n_vcycles
    nm = a.mixed_cycles
    nsw = s64.mg.n_pre + s64.mg.n_post  # level-0 sweeps per cycle
    for tag, mg, eb, cyc64, cycmx in (("fp64", s64.mg, 8, nv, 0), ("fp32", smx.mg, 4, 0, nm)):
        lv0, lv1 = mg.levels[0], mg.levels[1]
        L0, L1 = lv0.desc(), lv1.desc()
        knob = autotune.SWEEP_PAIRS_F64 if eb == 8 else autotune.SWEEP_PAIRS_F32
        pairs = autotune.get_shape(knob, lv0.n3)
        two = pairs == 2 or (pairs == 0 and eb == 8 and cells >= 2 ** 23)
        ctype = "double" if eb == 8 else "float"
        kern = ("sweep2_kernel<%s> (two RBGS sweeps per launch)" % ctype if two
                else "sweep_kernel<%s> (one RBGS sweep per launch)" % ctype)
        # one cycle's level-0 smoothing (6 pre + 6 post sweeps) as one group of
        # launches: the first reads arrays another kernel touched long before,
        # the others read what the previous sweep wrote, as in a V-cycle
        out.append(row("rbgs %d sweeps L0 (%s)" % (nsw, tag), kern,
                       lambda L0=L0: _hip.call("wf_mg_relax", C.byref(L0), nsw, st),
                       nsw * (4 * eb) * cells, cyc64, cycmx, nsw // 2 if two else nsw))
        out.append(row("restrict L0 (%s)" % tag, "restrict_kernel<%s>" % ctype,
                       lambda L0=L0, L1=L1: _hip.call("wf_mg_restrict", C.byref(L0),
                                                      C.byref(L1), st),
                       (17 * eb / 8) * cells, cyc64, cycmx))
        out.append(row("prolong L0 (%s)" % tag, "prolong8_kernel<%s>" % ctype,
                       lambda L0=L0, L1=L1: _hip.call("wf_mg_prolong", C.byref(L1), C.byref(L0),
                                                      st),
                       (17 * eb / 8) * cells, cyc64, cycmx))
    return out


def load_traffic():
    """HBM bytes per launch of each hot kernel from the committed rocprofv3
    --pmc summary (read requests by size, write requests by size)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_cavity256.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh).get("kernels", {})


# table row of each momentum entry point the step calls
INSTEP_ROWS = {"convdiff3_save": "convdiff3 stage 1 (u, v, w)",
               "convdiff4th_u_edges": "convdiff u edges stage 2",
               "convdiff3": "convdiff3 stage 2 (u, v, w)"}


def instep_momentum_us(sol, nsteps=2):
    """Average duration per call of each momentum (conv4th + diff4th)
    entry point inside `nsteps` real steps of `sol`, keyed by table row:
    each call the step makes is bracketed by HIP events on the stream it
    is issued on (the kernels module's entry points are wrapped for these
    steps only).  Inside a step these kernels find part of their inputs in
    the Infinity Cache, left there by the passes just before them, which
    neither a cold nor a back-to-back timing reproduces."""
    from warpflow import _hip
    from warpflow import kernels as K
    from warpflow.device import current_stream, synchronize
    orig = {n: getattr(K, n) for n in INSTEP_ROWS}
    pairs = {n: [] for n in INSTEP_ROWS}

    def wrap(name, fn):
        def timed(*args, **kw):
            e0, e1 = C.c_void_p(), C.c_void_p()
            _hip.call("wf_event_create", C.byref(e0))
            _hip.call("wf_event_create", C.byref(e1))
            _hip.call("wf_event_record", e0, current_stream())
            fn(*args, **kw)
            _hip.call("wf_event_record", e1, current_stream())
            pairs[name].append((e0, e1))
        return timed
    try:
        for n in INSTEP_ROWS:
            setattr(K, n, wrap(n, orig[n]))
        for _ in range(nsteps):
            sol.step()
        synchronize()
    finally:
        for n in INSTEP_ROWS:
            setattr(K, n, orig[n])
    out = {}
    for n, lst in pairs.items():
        us = []
        for e0, e1 in lst:
            ms = C.c_float(0.0)
            _hip.call("wf_event_elapsed_ms", C.byref(ms), e0, e1)
            us.append(1000.0 * ms.value)
            _hip.call("wf_event_destroy", e0)
            _hip.call("wf_event_destroy", e1)
        if us:
            out[INSTEP_ROWS[n]] = (sum(us) / len(us), len(us))
    return out


def kernel_table(a, s64, smx):
    T = Timer()
    flush = Flush()
    traffic = load_traffic()
    rows = []
    instep = instep_momentum_us(smx)
    for r in hot_kernels(a, s64, smx):
        fn = r.pop("fn")
        cold = T.cold_us(fn, a.reps, flush)
        warm = T.warm_us(fn, a.reps)
        # the time the ranking and roofline use: in-step where measured, else
        # the cold call (a cycle's level-0 smoothing started cold takes a few
        # per cent longer than inside a step, so the cold figure is the lower
        # bound on the rate)
        t = cold
        if r["name"] in instep:
            per_call, n = instep[r["name"]]  # one call = the row's launches
            t = per_call
            r["us_instep_per_launch"] = round(per_call / r["launches"], 2)
            r["instep_calls_timed"] = n
        r.update({"us_cold": round(cold, 2), "us_warm": round(warm, 2), "us": round(t, 2),
                  "us_per_launch": round(t / r["launches"], 2),
                  "GBps": round(r["bytes"] / t / 1e3, 1) if r["bytes"] else None,
                  "hbm_frac": round(r["bytes"] / t / 1e3 / HBM_PEAK_GBS, 3)
                  if r["bytes"] else None,
                  "us_per_step_mixed": round(t * r["per_step_mixed"], 1),
                  "us_per_step_fp64": round(t * r["per_step_fp64"], 1)})
        if r["name"] in traffic:
            r["traffic_per_launch"] = traffic[r["name"]]["bytes_per_launch"]
        r["bytes"] = int(r["bytes"])
        rows.append(r)
    return rows


def probe(a):
    """Short run for hardware-counter passes: build both solvers (no steps),
    then launch every hot kernel a.reps times, in order, on 256^3 arrays."""
    from warpflow import FlowSolver, build_topology
    from warpflow.device import synchronize
    s64 = FlowSolver(make_cfg(a, "fp64"), build_topology(make_cfg(a, "fp64")), "hip:0")
    smx = FlowSolver(make_cfg(a, "mixed"), build_topology(make_cfg(a, "mixed")), "hip:0")
    tune(a, s64)
    tune(a, smx)
    synchronize()
    for r in hot_kernels(a, s64, smx):
        print("# probe %s: %d calls, %d launches per call, %d algorithmic bytes per call"
              % (r["name"], a.reps, r["launches"], r["bytes"]), file=sys.stderr, flush=True)
        for _ in range(a.reps):
            r["fn"]()
        synchronize()
    return 0


# Trace/counter rows: the kernel each table row launches, as rocprofv3 names
# it (the level-0 launch is the one with the largest grid).
ROW_KERNELS = {
    "convdiff3 stage 1 (u, v, w)": ("convdiff3_save_kernel(",),
    "convdiff u edges stage 2": ("convdiff_kernel<0, false>",),
    "convdiff3 stage 2 (u, v, w)": ("convdiff3_kernel(",),
    "calvel3": ("calvel3_kernel(",),
    "calvel3+calmas": ("calvel3_calmas_kernel<true>",),
    "calmas of calvel3 (no store)": ("calvel3_calmas_kernel<false>",),
    "wale+checkdt": ("wale_checkdt_kernel(",),
    "add_ring": ("add_kernel(",),
    "copy (u -> uoo)": ("copy16_kernel(",),
    "bc_fill (u)": ("bc_axis_kernel(",),
    "rbgs 12 sweeps L0 (fp64)": ("sweep2_kernel<double", "sweep_kernel<double"),
    "restrict L0 (fp64)": ("restrict_kernel<double>",),
    "prolong L0 (fp64)": ("prolong8_kernel<double>", "prolong_kernel<double>"),
    "rbgs 12 sweeps L0 (fp32)": ("sweep2_kernel<float", "sweep_kernel<float"),
    "restrict L0 (fp32)": ("restrict_kernel<float>",),
    "prolong L0 (fp32)": ("prolong8_kernel<float>", "prolong_kernel<float>"),
}


def _row_dispatches(recs, row):
    """(kernel name, dispatches) of `row` among recs = [(name, grid, ...)]:
    the first listed kernel present, at its largest grid."""
    for pat in ROW_KERNELS[row]:
        hits = [r for r in recs if pat in r[0]]
        if hits:
            g = max(r[1] for r in hits)
            return hits[0][0], [r for r in hits if r[1] == g]
    return None, []


def summarize_pmc(rd_db, wr_db, out_json):
    """HBM traffic per launch from two rocprofv3 --pmc runs of --probe:
    reads = 32, 64 and 128-byte read requests (TCC_EA0_RDREQ_{32B,64B,128B}),
    writes = 64-byte write requests plus 32 bytes for each other write
    request (TCC_EA0_WRREQ, TCC_EA0_WRREQ_64B); no calibration factor (the
    copy kernel reads and writes exactly its array size this way)."""
    import collections
    import sqlite3

    def load(db):
        per = collections.defaultdict(dict)
        con = sqlite3.connect(db)
        for d, k, g, cn, v in con.execute(
                "select dispatch_id, kernel_name, grid_size, counter_name, value "
                "from counters_collection"):
            per[d].update({"name": k, "grid": g, cn: v})
        return per
    rd, wr = load(rd_db), load(wr_db)

    def recs(per, fn):
        return [(v["name"], v["grid"], fn(v)) for v in per.values()]
    rbytes = recs(rd, lambda v: 32 * v.get("TCC_EA0_RDREQ_32B_sum", 0) +
                  64 * v.get("TCC_EA0_RDREQ_64B_sum", 0) +
                  128 * v.get("TCC_EA0_RDREQ_128B_sum", 0))
    wbytes = recs(wr, lambda v: 64 * v.get("TCC_EA0_WRREQ_64B_sum", 0) +
                  32 * (v.get("TCC_EA0_WRREQ_sum", 0) - v.get("TCC_EA0_WRREQ_64B_sum", 0)))
    out = {}
    for row in ROW_KERNELS:
        kr, hr = _row_dispatches(rbytes, row)
        kw, hw = _row_dispatches(wbytes, row)
        if not hr or not hw:
            continue
        r = sum(x[2] for x in hr) / len(hr)
        w = sum(x[2] for x in hw) / len(hw)
        out[row] = {"kernel": kr.split("(")[0], "grid": hr[0][1], "dispatches": len(hr),
                    "read_bytes": int(r), "write_bytes": int(w), "bytes_per_launch": int(r + w)}
    doc = {"counters": ["TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum",
                        "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum"],
           "bytes": "read = 32*RDREQ_32B + 64*RDREQ_64B + 128*RDREQ_128B; write = "
                    "64*WRREQ_64B + 32*(WRREQ - WRREQ_64B); averaged over the dispatches of "
                    "the row's kernel at its largest grid", "kernels": out}
    with open(out_json, "w") as fh:
        json.dump(doc, fh, indent=1)
    return doc


def summarize_sq(dbs):
    """Per-wave SQ counters of the table rows from rocprofv3 --pmc runs of
    --probe (one database per pass): counter / SQ_WAVES averaged over the
    row's dispatches at its largest grid; "of life" = counter / SQ_WAVE_CYCLES
    (both in the same cycle units)."""
    import collections
    import sqlite3
    per = collections.defaultdict(dict)  # (db index, dispatch) -> counters
    for q, db in enumerate(dbs):
        con = sqlite3.connect(db)
        for d, k, g, cn, v in con.execute(
                "select dispatch_id, kernel_name, grid_size, counter_name, value "
                "from counters_collection"):
            per[(q, d)].update({"name": k, "grid": g, cn: v})
    recs = [(v["name"], v["grid"], v) for v in per.values()]
    cols = [("SQ_WAVES", "waves", False), ("SQ_WAVE_CYCLES", "cycles / wave", True),
            ("SQ_INSTS_VALU", "VALU insts", True), ("SQ_INSTS_SALU", "SALU insts", True),
            ("SQ_INSTS_LDS", "LDS insts", True), ("SQ_INSTS_VMEM_RD", "VMEM rd", True),
            ("SQ_INSTS_VMEM_WR", "VMEM wr", True), ("SQ_LDS_BANK_CONFLICT", "LDS conflict cyc",
                                                    True)]
    life = [("SQ_ACTIVE_INST_VALU", "VALU active"), ("SQ_ACTIVE_INST_LDS", "LDS active"),
            ("SQ_ACTIVE_INST_SCA", "SALU active"), ("SQ_WAIT_INST_ANY", "wait on deps"),
            ("SQ_WAIT_ANY", "any wait")]
    lines = ["| row | kernel | " + " | ".join(c[1] for c in cols) + " | " +
             " | ".join(c[1] + " (of life)" for c in life) + " |",
             "|---|---|" + "---:|" * (len(cols) + len(life))]
    for row in ROW_KERNELS:
        avg = {}
        name = None
        for cn, _, _ in cols + [(x[0], None, None) for x in life]:
            hits = [r for r in recs if cn in r[2]]
            k, h = _row_dispatches(hits, row)
            if h:
                name = k
                avg[cn] = sum(x[2][cn] for x in h) / len(h)
        if not name or not avg.get("SQ_WAVES"):
            continue
        w = avg["SQ_WAVES"]
        cells = []
        for cn, _, perwave in cols:
            v = avg.get(cn)
            cells.append("" if v is None else ("%.0f" % (v / w if perwave else v)))
        wc = avg.get("SQ_WAVE_CYCLES")
        for cn, _ in life:
            v = avg.get(cn)
            cells.append("" if v is None or not wc else "%.3f" % (v / wc))
        lines.append("| %s | `%s` | %s |" % (row, name.split("(")[0].replace("void ", ""),
                                               " | ".join(cells)))
    return "\n".join(lines)


def summarize_trace(db, nsteps):
    """Per-kernel table of a rocprofv3 --kernel-trace database: launches,
    total and average duration, share; per step = total / nsteps."""
    import collections
    import sqlite3
    con = sqlite3.connect(db)
    agg = collections.defaultdict(lambda: [0, 0.0])
    for name, t0, t1 in con.execute("select name, start, end from kernels"):
        a = agg[name]
        a[0] += 1
        a[1] += (t1 - t0) / 1e3
    tot = sum(v[1] for v in agg.values())
    # the bench's rows: their kernels at the largest grid (the 256^3 launches)
    recs = [(n, gx * gy * gz, (t1 - t0) / 1e3) for n, gx, gy, gz, t0, t1 in con.execute(
        "select name, grid_x, grid_y, grid_z, start, end from kernels")]
    lines = ["| bench row | kernel | launches | avg us | per step us |",
             "|---|---|---:|---:|---:|"]
    brows = []
    for row in ROW_KERNELS:
        k, hits = _row_dispatches(recs, row)
        if hits:
            t = sum(x[2] for x in hits)
            brows.append((t / nsteps, row, k, len(hits), t / len(hits)))
    for per, row, k, n, avg in sorted(brows, reverse=True):
        lines.append("| %s | `%s` | %d | %.2f | %.1f |"
                     % (row, k.split("(")[0].replace("void ", ""), n, avg, per))
    lines += ["", "| kernel | launches | total us | avg us | per step us | share |",
              "|---|---:|---:|---:|---:|---:|"]
    for name, (n, t) in sorted(agg.items(), key=lambda x: -x[1][1]):
        lines.append("| `%s` | %d | %.1f | %.2f | %.1f | %.1f%% |"
                     % (name.split("(")[0].replace("void ", ""), n, t, t / n, t / nsteps,
                        100.0 * t / tot))
    lines.append("")
    lines.append("kernel time %.1f us over %d launches; %.1f us per step over %d steps"
                 % (tot, sum(v[0] for v in agg.values()), tot / nsteps, nsteps))
    return "\n".join(lines)


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _time_steps(sol, warm, reps):
    for _ in range(warm):
        sol.step()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        sol.step()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def cpu_baselines(a):
    """The C oracle (the reference's FP64 schedule in C, OpenMP) and the NumPy
    oracle on the host: the headline workload on all cores (median of 3 steps
    after one) and on one thread (one step), and the 32^3 and 80^3 cases of
    the reference on one thread and on all cores."""
    from oracle import c_oracle as CO
    from oracle import np_oracle as O
    from warpflow import load_case
    model, ncpu = cpu_info()
    allc = CO.threads()
    legs = {}
    case = O.Case(n3=(a.n,) * 3, reynolds=a.reynolds, sgs_enabled=True, cw=0.46, noise=0.01,
                  mg_levels=a.levels, variable_dt=False, dt_init=a.dt)
    sol = CO.CSolver(case)
    sec = _time_steps(sol, 1, 3)
    cells = a.n ** 3
    base = {"value": round(cells / sec / 1e6, 3), "unit": "Mcell/s", "cores": allc,
            "kind": "port",
            "sample": "median of steps 2-4 of the same %d^3 Re-%d cavity (plain FP64 schedule, "
                      "%d MG levels, 4 V-cycles) in the C oracle (oracle/c/wf_oracle.c), %.3f s "
                      "per step on %d threads of %s" % (a.n, a.reynolds, a.levels, sec, allc,
                                                        model)}
    legs["cavity%d_c_all" % a.n] = {"threads": allc, "ms_per_step": round(1e3 * sec, 1),
                                    "Mcell_s": base["value"]}
    CO.set_threads(1)
    try:
        sec1 = _time_steps(sol, 0, 1)
        legs["cavity%d_c_1thread" % a.n] = {"threads": 1, "ms_per_step": round(1e3 * sec1, 1),
                                            "Mcell_s": round(cells / sec1 / 1e6, 3)}
        del sol
        for tag, path, warm, reps in (("cavity32_re100", "cavity_re100_32.yaml", 2, 5),
                                      ("cavity80_re400", "cavity_serial.yaml", 1, 3)):
            cfg = load_case(os.path.join(PKG, "cases", path))
            ncell = cfg.grid.nx * cfg.grid.ny * cfg.grid.nz
            for thr in (1, allc):
                CO.set_threads(thr)
                s = _time_steps(CO.CSolver(O.Case.from_config(cfg)), warm, reps)
                legs["%s_c_%s" % (tag, "1thread" if thr == 1 else "all")] = {
                    "threads": thr, "ms_per_step": round(1e3 * s, 2),
                    "Mcell_s": round(ncell / s / 1e6, 3)}
            if tag == "cavity32_re100":
                s = _time_steps(O.Solver(O.Case.from_config(cfg)), 1, 2)
                legs["cavity32_re100_numpy"] = {"threads": 1, "ms_per_step": round(1e3 * s, 1),
                                                "Mcell_s": round(ncell / s / 1e6, 4)}
    finally:
        CO.set_threads(allc)
    legs["host"] = {"cpu": model, "logical_cpus": ncpu}
    return base, legs


def cavity80(a):
    """cases/cavity_serial.yaml (80^3, Re 400, WALE, adaptive dt, 3 levels) in
    both forms: the only configuration the reference publishes a rate for."""
    from warpflow import FlowSolver, build_topology, load_case
    from warpflow.device import synchronize
    out = {}
    for form in ("fp64", "mixed"):
        cfg = load_case(os.path.join(PKG, "cases", "cavity_serial.yaml"))
        cfg.numerics.mg_precision = form
        sol = FlowSolver(cfg, build_topology(cfg), "hip:0")
        tune(a, sol)
        for _ in range(30):
            sol.step()
        synchronize()
        nsteps = 300
        t0 = time.perf_counter()
        for _ in range(nsteps):
            info = sol.step()
        synchronize()
        sec = (time.perf_counter() - t0) / nsteps
        out[form] = {"ms_per_step": round(1e3 * sec, 3),
                     "Mcell_s": round(80 ** 3 / sec / 1e6, 1), "dt": info["dt"],
                     "iters": info["iters"]}
        del sol
    out["reference_Mcell_s"] = REPORT_80_MCELLS
    out["reference_hardware"] = "RTX 5060 Laptop GPU, FP64 (REPORT.md:101-103)"
    out["speedup_fp64_vs_reference"] = round(out["fp64"]["Mcell_s"] / REPORT_80_MCELLS, 1)
    out["steps_timed"] = "steps 31-330 of the case (adaptive dt near its saturated 4.8e-3)"
    return out


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
                       "cells": cells, "graph": True,
                       "launches": "one launch per operator for all members (not one "
                                   "persistent launch per RK stage), graph-captured step"},
            "sequential_fp64_value": round(out["sequential_fp64"], 3),
            "sequential_mixed_value": round(out["sequential_mixed"], 3),
            "speedup_vs_sequential_fp64": round(out["batched"] / out["sequential_fp64"], 3),
            "speedup_vs_sequential_mixed": round(out["batched"] / out["sequential_mixed"], 3),
            "max_outer_iters": out["iters"]}
    print(json.dumps(line), flush=True)
    return 0


def run_poisson(a):
    """Standalone Poisson solve at n^3 (cell-centred, homogeneous Neumann walls,
    4 levels, the reference's standalone multigrid input: a seeded
    standard-normal rhs, ring zeroed, interior demeaned): plain FP64 V(6,6)
    cycles (300 coarsest sweeps) repeated until max|rhs - A phi| <= 1e-10
    max|rhs|, against FP32 V-cycles inside FP64 defect correction stopped by
    the same test; the per-kernel table of the level-0 and coarsest-level
    work of both hierarchies (cold HIP-event timings)."""
    from warpflow import _hip, build_topology
    from warpflow.config import CaseConfig, validate
    from warpflow.device import array, current_stream, synchronize, zeros
    from warpflow.mixed import MixedMultigrid
    from warpflow.multigrid import Multigrid
    n, levels, rtol = a.n, 4, 1e-10
    cfg = CaseConfig()
    cfg.grid.nx = cfg.grid.ny = cfg.grid.nz = n
    cfg.numerics.mg_levels = levels
    validate(cfg)
    topo = build_topology(cfg)
    h = 1.0 / n
    rhs = np.random.default_rng(0).standard_normal((n + 2,) * 3)
    for ax in range(3):
        sl = [slice(None)] * 3
        sl[ax] = [0, -1]
        rhs[tuple(sl)] = 0.0
    rhs[1:-1, 1:-1, 1:-1] -= rhs[1:-1, 1:-1, 1:-1].mean()
    bmax = float(np.abs(rhs[1:-1, 1:-1, 1:-1]).max())
    mg = Multigrid(cfg, topo, "hip:0", h, h, h)
    lev0 = mg.levels[0]
    out = {}
    for rep in range(2):  # an untimed solve first (first-touch, kernel loading)
        lev0.rhs.assign(rhs)
        lev0.phi.zero_()
        mg._refill(lev0)
        synchronize()
        t0 = time.perf_counter()
        hist = []
        for _ in range(40):
            mg._vcycle()
            hist.append(mg.residual_max() / bmax)
            if hist[-1] <= rtol:
                break
        out["fp64_ms"] = 1e3 * (time.perf_counter() - t0)
    out["fp64_cycles"], out["fp64_residual"] = len(hist), hist[-1]
    phi64 = lev0.phi.numpy()[1:-1, 1:-1, 1:-1]
    mm = MixedMultigrid(cfg, topo, "hip:0", h, h, h, rtol=rtol, max_cycles=40)
    su, pp = array(rhs, np.float64), zeros(rhs.shape, np.float64)
    for rep in range(2):
        synchronize()
        t0 = time.perf_counter()
        mm.solve(su, pp, 1)
        synchronize()
        out["mixed_ms"] = 1e3 * (time.perf_counter() - t0)
    out["mixed_cycles"], out["mixed_residual"] = mm.last_cycles, mm.last_residual
    pmx = pp.numpy()[1:-1, 1:-1, 1:-1]
    d64, dmx = phi64 - phi64.mean(), pmx - pmx.mean()
    err = float(np.sqrt(np.mean((d64 - dmx) ** 2)) / np.sqrt(np.mean(d64 ** 2)))
    # per-kernel table: level 0 and the coarsest level of each hierarchy
    T, flush, st = Timer(), Flush(), current_stream()
    rows = []
    for tag, m, eb in (("fp64", mg, 8), ("fp32", mm, 4)):
        L = [lv.desc() for lv in m.levels]
        cells0 = float(np.prod(m.levels[0].n3))
        cellsc = float(np.prod(m.levels[-1].n3))
        for name, fn, nbytes, launches in (
                ("rbgs 12 sweeps L0 (%s)" % tag,
                 lambda L=L: _hip.call("wf_mg_relax", C.byref(L[0]), 12, st),
                 12 * 4 * eb * cells0, None),
                ("restrict L0 (%s)" % tag,
                 lambda L=L: _hip.call("wf_mg_restrict", C.byref(L[0]), C.byref(L[1]), st),
                 17 * eb / 8 * cells0, 1),
                ("prolong L0 (%s)" % tag,
                 lambda L=L: _hip.call("wf_mg_prolong", C.byref(L[1]), C.byref(L[0]), st),
                 17 * eb / 8 * cells0, 1),
                ("rbgs 300 sweeps coarsest %d^3 (%s)" % (round(cellsc ** (1 / 3)), tag),
                 lambda L=L: _hip.call("wf_mg_relax", C.byref(L[-1]), 300, st),
                 300 * 4 * eb * cellsc, None)):
            us = T.cold_us(fn, max(3, a.reps // 2), flush)
            rows.append({"kernel": name, "us": round(us, 2), "bytes": int(nbytes),
                         "GBps": round(nbytes / us / 1e3, 1),
                         "hbm_frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 3)})
    print("# poisson%d: fp64 %d cycles %.1f ms, mixed %d cycles %.1f ms, rms(p_mix - p_64) / "
          "rms(p_64) = %.2e" % (n, out["fp64_cycles"], out["fp64_ms"], out["mixed_cycles"],
                                out["mixed_ms"], err), file=sys.stderr, flush=True)
    line = {"metric": "standalone Poisson %d^3 to 1e-10 relative residual: solve time" % n,
            "value": round(out["mixed_ms"], 2), "unit": "ms", "n_gpus": 1,
            "higher_is_better": False, "vs_baseline": None, "dtype": "f64 (V-cycle f32)",
            "data": "synthetic (seeded standard-normal rhs, demeaned)",
            "config": {"workload": "poisson%d: cell-centred Neumann, %d levels, V(6,6), 300 "
                                   "coarsest sweeps" % (n, levels)},
            "fp64_ms": round(out["fp64_ms"], 2), "fp64_cycles": out["fp64_cycles"],
            "mixed_cycles": out["mixed_cycles"],
            "speedup_mixed_vs_fp64": round(out["fp64_ms"] / out["mixed_ms"], 3),
            "residual": {"fp64": out["fp64_residual"], "mixed": out["mixed_residual"]},
            "rms_p_mixed_vs_fp64": err, "kernels": rows}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    a = parse(argv)
    if a.gpus != 1:
        print("bench.py: this build runs one cavity on one GPU; --gpus %d is not supported "
              "(multi-GPU domain decomposition is out of scope)" % a.gpus, file=sys.stderr)
        return 2
    if a.summarize_trace:
        print(summarize_trace(a.summarize_trace, a.steps))
        return 0
    if a.summarize_sq:
        print(summarize_sq(a.summarize_sq))
        return 0
    if a.summarize_pmc:
        print(json.dumps(summarize_pmc(*a.summarize_pmc)["kernels"], indent=1))
        return 0
    if a.probe:
        return probe(a)
    if a.workload == "ensemble64":
        return run_ensemble(a)
    if a.workload == "poisson256":
        return run_poisson(a)
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
                       "vcycles_per_solve": {"fp64": 4, "mixed": a.mixed_cycles, "mixed4": 4},
                       "sweeps": "V(6,6)+300 coarsest"}}
    line["config"]["launch_shapes"] = {f: res[f]["tuned"] for f in res}  # [knob, n3, value]
    for f in res:
        if f != head:
            line["%s_value" % f] = round(res[f]["mcell"], 3)
            line["%s_ms_per_step" % f] = round(res[f]["ms"], 4)
    if "fp64" in res and head == "mixed":
        line["speedup_mixed_vs_fp64"] = round(res["mixed"]["mcell"] / res["fp64"]["mcell"], 4)
        if "mixed4" in res:
            line["speedup_mixed_vs_fp64_equal_cycles"] = round(
                res["mixed4"]["mcell"] / res["fp64"]["mcell"], 4)
        line["deviation_mixed_vs_fp64"] = deviation(res["mixed"]["sol"], res["fp64"]["sol"])
        line["rmax_trajectory"] = {f: [float("%.4g" % i["rmax"]) for i in res[f]["infos"]]
                                   for f in res}
    for f in list(res):
        if f not in ("mixed", "fp64"):
            del res[f]["sol"]
    if not a.no_kernels and "fp64" in res and "mixed" in res:
        rows = kernel_table(a, res["fp64"]["sol"], res["mixed"]["sol"])
        line["kernels"] = rows
        key = "us_per_step_mixed" if head == "mixed" else "us_per_step_fp64"
        ranked = sorted(rows, key=lambda r: -r[key])
        dom = ranked[0]
        achieved = dom["bytes"] / dom["us"] / 1e3
        line["roofline"] = {"bound": "hbm", "kernel": dom["kernel"], "row": dom["name"],
                            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": dom.get("traffic_per_launch"),
                            "algorithmic_bytes_per_launch": int(dom["bytes"] / dom["launches"]),
                            "us_per_launch": dom["us_per_launch"],
                            "launches_per_step": dom["launches"] * dom[
                                "per_step_mixed" if head == "mixed" else "per_step_fp64"],
                            "us_per_step": dom[key]}
        nxt = ranked[1]
        line["roofline"]["runner_up"] = {
            "row": nxt["name"], "kernel": nxt["kernel"], "us_per_step": nxt[key],
            "frac": nxt["hbm_frac"], "traffic": nxt.get("traffic_per_launch")}
    for f in list(res):
        res[f].pop("sol", None)
    if not a.no_80:
        line["cavity80"] = cavity80(a)
    if not a.no_cpu:
        line["cpu_baseline"], line["cpu_legs"] = cpu_baselines(a)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
