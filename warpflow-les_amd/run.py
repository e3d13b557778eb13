"""Run a YAML cavity case on one MI355X.

    python run.py --case CASE.yaml [--steps N] [--outdir DIR] [--device hip:0]
                  [--precision fp64|mixed] [--restart SNAPSHOT]

The console output is the reference's (run.py:40-42 banner, :52-54 log line,
:62-64 closing line), so logs of the two programs can be compared line by
line.  Snapshots ``snap_%06d_r000.npz`` are written at every ``output.n_out``
steps and after the last one.  Besides the reference's keys they carry the
viscosity field and the step number, which make ``--restart`` exact: the run
continues from the snapshot's step up to the final step number as if it had
never stopped, adaptive dt included.  One domain on one GPU; a case with
``mpi.domains`` other than [1, 1, 1] is refused by ``build_topology``.
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from dataclasses import dataclass

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from warpflow import FlowSolver, __version__, build_topology, load_case  # noqa: E402
from warpflow.io import load_into, write_snapshot  # noqa: E402

# console formats kept from the reference (run.py:40-42, :52-54, :62-64)
BANNER = "warpflow-les %s: case=%s grid=%dx%dx%d domains=%s device=%s"
LOG_LINE = "it %6d  ctime %.6f  dt %.3e  rmax %.3e  outer %2d  cfl %.2f"
DONE_LINE = "done: itime=%d ctime=%.6f wall=%.1fs"


@dataclass(frozen=True)
class Schedule:
    """Which steps log and which write a snapshot: steps first+1 .. last."""
    first: int
    last: int
    log_every: int
    snap_every: int

    def steps(self):
        return range(self.first + 1, self.last + 1)

    def logs(self, it: int) -> bool:
        return it % self.log_every == 0

    def snaps(self, it: int) -> bool:
        # every n_out steps, and the last step when the interval skips it
        return it % self.snap_every == 0 or it == self.last


def _cli() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(
        prog="run.py", description="Advance a lid-driven-cavity case on one MI355X GPU.")
    ap.add_argument("--case", required=True, metavar="YAML", help="the case file to run")
    ap.add_argument("--steps", type=int, metavar="N",
                    help="final step number (default: time.n_steps of the case)")
    ap.add_argument("--outdir", metavar="DIR",
                    help="snapshot directory (default: output.outdir of the case)")
    ap.add_argument("--device", default="auto", metavar="DEV",
                    help="hip:N or cuda:N (the same GPU); auto = hip:0")
    ap.add_argument("--precision", choices=("fp64", "mixed"),
                    help="pressure solve: plain FP64 V-cycles, or FP32 V-cycles inside FP64 "
                         "defect correction (default: numerics.mg_precision of the case)")
    ap.add_argument("--restart", metavar="SNAPSHOT",
                    help="continue from a snapshot written by this program")
    return ap


def configure(opts):
    """The case with the command line's overrides applied."""
    cfg = load_case(opts.case)
    if opts.outdir is not None:
        cfg.output.outdir = opts.outdir
    if opts.precision is not None:
        cfg.numerics.mg_precision = opts.precision
    if opts.steps is not None:
        cfg.time.n_steps = opts.steps
    return cfg


def main(argv=None) -> int:
    opts = _cli().parse_args(argv)
    cfg = configure(opts)
    device = "hip:0" if opts.device == "auto" else opts.device
    topo = build_topology(cfg)
    g = cfg.grid
    print(BANNER % (__version__, cfg.name, g.nx, g.ny, g.nz, list(cfg.mpi.domains), device),
          flush=True)
    solver = FlowSolver(cfg, topo, device=device)
    if opts.restart:
        load_into(solver, opts.restart)
    plan = Schedule(solver.itime, cfg.time.n_steps, cfg.output.log_interval, cfg.output.n_out)
    started = time.time()
    for it in plan.steps():
        info = solver.step()
        if plan.logs(it):
            print(LOG_LINE % (it, info["ctime"], info["dt"], info["rmax"], info["iters"],
                              info["cfl"]), flush=True)
        if plan.snaps(it):
            write_snapshot(solver, it, restart=True)
    print(DONE_LINE % (plan.last, solver.ctime, time.time() - started), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
