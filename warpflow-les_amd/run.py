"""CLI entry point (the reference's run.py), on one MI355X.

Usage:
    python run.py --case cases/cavity_serial.yaml [--device hip:0] [--steps N]
                  [--outdir DIR]

Prints the reference's banner, one log line every output.log_interval steps
and a closing line, and writes snap_%06d_r000.npz every output.n_out steps
and after the last step.  A decomposed case (mpi.domains != [1, 1, 1])
raises: this build runs one domain on one GPU.
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from warpflow import FlowSolver, build_topology, load_case  # noqa: E402
from warpflow.io import write_snapshot  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser(description="warpflow-les solver (MI355X)")
    ap.add_argument("--case", required=True, help="case YAML file")
    ap.add_argument("--device", default="auto", help="auto | hip:0 | cuda:0")
    ap.add_argument("--steps", type=int, default=None,
                    help="override time.n_steps from the case file")
    ap.add_argument("--outdir", default=None, help="override output.outdir from the case file")
    args = ap.parse_args(argv)

    cfg = load_case(args.case)
    if args.outdir:
        cfg.output.outdir = args.outdir
    nsteps = args.steps if args.steps is not None else cfg.time.n_steps
    topo = build_topology(cfg)
    device = "hip:0" if args.device == "auto" else args.device
    print("warpflow-les %s: case=%s grid=%dx%dx%d domains=%s device=%s"
          % ("0.1.0", cfg.name, cfg.grid.nx, cfg.grid.ny, cfg.grid.nz, list(cfg.mpi.domains),
             device), flush=True)
    sol = FlowSolver(cfg, topo, device=device)
    t0 = time.time()
    for itime in range(1, nsteps + 1):
        info = sol.step()
        if itime % cfg.output.log_interval == 0:
            print("it %6d  ctime %.6f  dt %.3e  rmax %.3e  outer %2d  cfl %.2f"
                  % (itime, info["ctime"], info["dt"], info["rmax"], info["iters"], info["cfl"]),
                  flush=True)
        if itime % cfg.output.n_out == 0:
            write_snapshot(sol, itime)
    if nsteps % cfg.output.n_out != 0:
        write_snapshot(sol, nsteps)
    print("done: itime=%d ctime=%.6f wall=%.1fs" % (nsteps, sol.ctime, time.time() - t0),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
