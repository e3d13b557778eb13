"""Snapshot output (the reference's warpflow/io.py:9-27): one compressed
.npz per rank, ``snap_%06d_r%03d.npz``, holding the local block of u, v, w
and p including the ghost layers (float64), plus ctime, dt, the block's
global offsets, its cell counts and the global grid shape.  Files written
here load with the reference's post-processing and vice versa.

``write_snapshot(..., restart=True)`` adds two keys the reference does not
write: ``vis`` (the eddy viscosity of the last step, which the next step's
dt limit reads) and ``itime``.  With them ``load_into`` restarts a run
exactly, adaptive dt included; readers that look up the reference's keys are
unaffected."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

SNAP_FIELDS = ("u", "v", "w", "p")


def snapshot_path(outdir, itime, rank=0) -> Path:
    return Path(outdir) / ("snap_%06d_r%03d.npz" % (itime, rank))


def write_snapshot(sol, itime, restart=False):
    """io.py:9-27; restart=True also stores vis and itime."""
    outdir = Path(sol.cfg.output.outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    data = {name: sol.f[name].numpy() for name in SNAP_FIELDS}
    if restart:
        data["vis"] = sol.f["vis"].numpy()
        data["itime"] = np.int64(itime)
    path = snapshot_path(outdir, itime, sol.topo.rank)
    np.savez_compressed(
        path, **data,
        ctime=np.float64(sol.ctime), dt=np.float64(sol.dt),
        offsets=np.array(sol.topo.global_offsets, dtype=np.int64),
        ncells=np.array(sol.ncells, dtype=np.int64),
        gshape=np.array((sol.cfg.grid.nx, sol.cfg.grid.ny, sol.cfg.grid.nz), dtype=np.int64))
    return path


def read_snapshot(path) -> dict:
    """The arrays and scalars of one snapshot (no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_into(sol, path):
    """Restart: put a snapshot's u, v, w, p (ghosts included), ctime, dt and
    step number into a solver of the same grid.

    vis comes from the snapshot's ``vis`` key.  A file without it (one the
    reference wrote) gets the viscosity a step would form from the loaded
    velocities (WALE and its wall fill, or 1/Re without SGS); the step after
    the restart then starts from the same fields, but with an adaptive dt its
    dt limit reads that viscosity instead of the one the interrupted run had.
    The step number is ``itime`` if stored, else the one in the file name."""
    snap = read_snapshot(path)
    if tuple(snap["ncells"]) != tuple(sol.ncells):
        raise ValueError("snapshot grid %s != solver grid %s"
                         % (tuple(snap["ncells"]), tuple(sol.ncells)))
    for name in SNAP_FIELDS:
        sol.f[name].assign(snap[name])
    sol.ctime = float(snap["ctime"])
    sol.dt = float(snap["dt"])
    if "itime" in snap:
        sol.itime = int(snap["itime"])
    else:
        m = re.match(r"snap_(\d+)_r\d+\.npz$", Path(path).name)
        sol.itime = int(m.group(1)) if m else 0
    if "vis" in snap:
        sol.f["vis"].assign(snap["vis"])
    else:
        sol._sgs()
    return snap
