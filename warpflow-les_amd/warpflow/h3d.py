"""Hydro3D restart bridge (the reference's tests/h3d_bridge.py:59-172).

Hydro3D is the Fortran code the cavity schedule follows.  To compare the two
codes step by step, its run is started from the very state a FlowSolver
starts from:

  export   write ``final_ctime.dat`` and ``tecbin0000.bin`` (a Hydro3D
           restart) from a solver's fields: the real cells of u, v, w and p
           with Hydro3D's 4 ghost layers (LMR = 2, pl_ex = 2) rebuilt by its
           face rules, vis = 1/Re everywhere;
  compare  read Hydro3D's tecbin dumps (one every 25 steps) and this build's
           snapshots of the same steps, and report per-field max and rms
           differences.

tecbin layout (Fortran unformatted sequential records, little endian; every
record is framed by its byte length as an int32 before and after it):
  (tti, ttj, ttk) int32 x 3;  (pl) int32;  (-1, -1, -1) int32 x 3;  then one
  record of 23 float64 per cell, i fastest, then j, then k: x, y, z at the
  cell centre (i - pl + 1/2) * dx, p at column 3, vis at 6, u at 8, v at 11,
  w at 14, zeros elsewhere (h3d_bridge.py:87-115).

Everything here is host-side file handling with NumPy (no per-cell Python
loops); the fields come from the solver through ``DeviceArray.numpy()``.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np

from .topology import NPL

PL_H3D = 4  # Hydro3D ghost layers with LMR = 2, pl_ex = 2
NCOL = 23
COLUMNS = {"p": 3, "vis": 6, "u": 8, "v": 11, "w": 14}
_SIDES = (("west", 0, False), ("east", 0, True), ("south", 1, False), ("north", 1, True),
          ("bottom", 2, False), ("top", 2, True))
_NORMAL = {"west": "u", "east": "u", "south": "v", "north": "v", "bottom": "w", "top": "w"}
# one cell record: 4-byte length, 23 doubles, 4-byte length
_REC = np.dtype([("head", "<i4"), ("val", "<f8", (NCOL,)), ("tail", "<i4")])


def h3d_ghosts(real, name, bnd, npl, ulid):
    """Field `name` with `npl` ghost layers rebuilt around its real cells by
    Hydro3D's bounds.for rules for the cavity set (h3d_bridge.py:27-56): p is
    mirrored (Neumann); the lid side of u is the lid velocity; the component
    normal to a side is zero in its ghosts; the other components mirror with
    +1 at slip and lid sides and -1 at walls.  Sides are applied in the order
    west, east, south, north, bottom, top, each over the whole ghost plane,
    so a later side overwrites the edges an earlier one wrote."""
    real = np.asarray(real, dtype=np.float64)
    n3 = real.shape
    g = np.zeros(tuple(n + 2 * npl for n in n3))
    g[npl:npl + n3[0], npl:npl + n3[1], npl:npl + n3[2]] = real
    for side, ax, high in _SIDES:
        kind = bnd[side]
        n = n3[ax]
        for m in range(npl):
            gi, ri = (npl + n + m, npl + n - 1 - m) if high else (npl - 1 - m, npl + m)
            dst = [slice(None)] * 3
            src = [slice(None)] * 3
            dst[ax], src[ax] = gi, ri
            if name == "p":
                g[tuple(dst)] = g[tuple(src)]
            elif kind == "moving_lid" and name == "u":
                g[tuple(dst)] = ulid
            elif name == _NORMAL[side]:
                g[tuple(dst)] = 0.0
            else:
                sign = 1.0 if kind in ("slip", "moving_lid") else -1.0
                g[tuple(dst)] = sign * g[tuple(src)]
    return g


def write_tecbin(path, fields, dx, pl=PL_H3D):
    """Write fields {"u","v","w","p","vis"} (arrays of (n + 2 pl)^3, indexed
    [i, j, k]) as a tecbin file; returns its path."""
    shape = fields["u"].shape
    tti, ttj, ttk = shape
    with open(path, "wb") as fh:
        for rec in (np.array([tti, ttj, ttk], "<i4"), np.array([pl], "<i4"),
                    np.array([-1, -1, -1], "<i4")):
            mark = np.array([rec.nbytes], "<i4")
            fh.write(mark.tobytes() + rec.tobytes() + mark.tobytes())
        cells = np.zeros(tti * ttj * ttk, dtype=_REC)
        cells["head"] = cells["tail"] = NCOL * 8
        val = cells["val"]
        # record order: i fastest, then j, then k  ->  C order of [k, j, i]
        ci, cj, ck = np.meshgrid(*[(np.arange(t) - pl + 0.5) * dx for t in shape],
                                 indexing="ij")
        for col, a in ((0, ci), (1, cj), (2, ck)):
            val[:, col] = a.transpose(2, 1, 0).ravel()
        for name, col in COLUMNS.items():
            val[:, col] = np.asarray(fields[name], np.float64).transpose(2, 1, 0).ravel()
        fh.write(cells.tobytes())
    return Path(path)


def read_tecbin(path, n=None, raw=False):
    """Fields {"u","v","w","p","vis"} of a tecbin file, indexed [i, j, k]:
    the whole (n + 2 pl)^3 block with raw=True, else its real cells
    [pl, pl + n) (n defaults to the extent minus 2 pl)."""
    with open(path, "rb") as fh:
        data = fh.read()
    hdr = np.frombuffer(data, "<i4", count=13)
    if not (hdr[0] == 12 and hdr[4] == 12 and hdr[5] == 4 and hdr[7] == 4 and hdr[8] == 12
            and hdr[12] == 12):
        raise ValueError("%s: not a tecbin file (record markers %s)" % (path, hdr.tolist()))
    tti, ttj, ttk = (int(x) for x in hdr[1:4])
    pl = int(hdr[6])
    cells = np.frombuffer(data, _REC, count=tti * ttj * ttk, offset=13 * 4)
    if np.any(cells["head"] != NCOL * 8) or np.any(cells["tail"] != NCOL * 8):
        raise ValueError("%s: bad cell record markers" % path)
    out = {}
    for name, col in COLUMNS.items():
        a = cells["val"][:, col].reshape(ttk, ttj, tti).transpose(2, 1, 0).copy()
        if not raw:
            m = (tti - 2 * pl) if n is None else n
            a = a[pl:pl + m, pl:pl + m, pl:pl + m].copy()
        out[name] = a
    out["pl"] = pl
    return out


def export_restart(sol, outdir):
    """final_ctime.dat and tecbin0000.bin of `sol`'s current fields
    (h3d_bridge.py:59-115); the reference exports a freshly initialised
    solver.  Returns the tecbin path."""
    n = sol.ncells
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    with open(outdir / "final_ctime.dat", "w") as fh:
        fh.write("       0%18.12f%18.12f%18.12f\n" % (0.0, 0.0, 1.0))
    fields = {}
    for name in ("u", "v", "w", "p"):
        real = sol.f[name].numpy()[NPL:NPL + n[0], NPL:NPL + n[1], NPL:NPL + n[2]]
        fields[name] = h3d_ghosts(real, name, sol.bnd, PL_H3D, sol.ulid)
    fields["vis"] = np.full_like(fields["u"], sol.rey)
    return write_tecbin(outdir / "tecbin0000.bin", fields, sol.dx)


def compare_dumps(h3ddir, snapdir, n=80, step=25, upto=100):
    """Per-field differences between Hydro3D dumps h3d_it%03d.bin and
    snapshots snap_%06d_r000.npz of the same steps (h3d_bridge.py:137-172):
    Hydro3D file plane s pairs with this build's array plane s - 1 (its real
    velocity planes are file planes 4 .. n + 2); p is compared demeaned.
    Returns {step: {field: (max |d|, rms d)}}; missing pairs are skipped."""
    lo, hi = PL_H3D, n + 3
    res = {}
    for it in range(step, upto + 1, step):
        hp = os.path.join(h3ddir, "h3d_it%03d.bin" % it)
        sp = os.path.join(snapdir, "snap_%06d_r000.npz" % it)
        if not (os.path.exists(hp) and os.path.exists(sp)):
            continue
        h3d = read_tecbin(hp, n, raw=True)
        with np.load(sp, allow_pickle=False) as z:
            mine_all = {k: z[k] for k in ("u", "v", "w", "p")}
        res[it] = {}
        for name in ("u", "v", "w", "p"):
            mine = mine_all[name][lo - 1:hi - 1, lo - 1:hi - 1, lo - 1:hi - 1].copy()
            href = h3d[name][lo:hi, lo:hi, lo:hi].copy()
            if name == "p":
                mine -= mine.mean()
                href -= href.mean()
            d = mine - href
            res[it][name] = (float(np.abs(d).max()), float(np.sqrt(np.mean(d ** 2))))
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Hydro3D restart bridge")
    sub = ap.add_subparsers(dest="cmd", required=True)
    ex = sub.add_parser("export", help="write a Hydro3D restart of a fresh cavity case")
    ex.add_argument("outdir")
    ex.add_argument("--case", default=str(Path(__file__).resolve().parent.parent / "cases" /
                                          "cavity_serial.yaml"))
    cp = sub.add_parser("compare", help="compare Hydro3D dumps with snapshots")
    cp.add_argument("h3ddir")
    cp.add_argument("snapdir")
    cp.add_argument("--n", type=int, default=80)
    a = ap.parse_args(argv)
    if a.cmd == "export":
        from . import FlowSolver, build_topology, load_case
        cfg = load_case(a.case)
        sol = FlowSolver(cfg, build_topology(cfg), "hip:0")
        print("wrote", export_restart(sol, a.outdir))
        return 0
    for it, per in compare_dumps(a.h3ddir, a.snapdir, a.n).items():
        print("it %3d  " % it + "  ".join("%s max|d| %.3e rms %.3e" % (k, *v)
                                          for k, v in per.items()))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
