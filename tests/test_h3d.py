"""The Hydro3D restart bridge (warpflow/h3d.py; the reference's
tests/h3d_bridge.py:27-172): tecbin byte layout, round trip, the ghost rules
and the step-by-step comparison, on the CPU.  The GPU export from a live
solver is in test_gpu_run.py."""
import struct

import numpy as np

from warpflow import h3d

BND = {"west": "wall", "east": "wall", "south": "slip", "north": "slip", "bottom": "wall",
       "top": "moving_lid"}


def _fields(n, pl, seed=0):
    rng = np.random.default_rng(seed)
    shape = (n + 2 * pl,) * 3
    return {k: rng.standard_normal(shape) for k in ("u", "v", "w", "p", "vis")}


def _record_writer(path, f, dx, pl):
    """A record-by-record writer (one struct.pack per record, loops over k,
    j, i), independent of the vectorised one."""
    t = f["u"].shape[0]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<i3ii", 12, t, t, t, 12))
        fh.write(struct.pack("<iii", 4, pl, 4))
        fh.write(struct.pack("<i3ii", 12, -1, -1, -1, 12))
        for k in range(t):
            for j in range(t):
                for i in range(t):
                    rec = [0.0] * 23
                    rec[0], rec[1], rec[2] = ((i - pl + 0.5) * dx, (j - pl + 0.5) * dx,
                                              (k - pl + 0.5) * dx)
                    rec[3], rec[6] = f["p"][i, j, k], f["vis"][i, j, k]
                    rec[8], rec[11], rec[14] = f["u"][i, j, k], f["v"][i, j, k], f["w"][i, j, k]
                    fh.write(struct.pack("<i23di", 184, *rec, 184))


def test_tecbin_bytes_match_a_record_writer(tmp_path):
    n, pl, dx = 2, 4, 1.0 / 2
    f = _fields(n, pl)
    a = h3d.write_tecbin(tmp_path / "a.bin", f, dx, pl)
    _record_writer(tmp_path / "b.bin", f, dx, pl)
    ba, bb = a.read_bytes(), (tmp_path / "b.bin").read_bytes()
    t = n + 2 * pl
    assert len(ba) == 20 + 12 + 20 + t ** 3 * 192
    assert ba == bb


def test_tecbin_round_trip(tmp_path):
    n, pl = 3, 4
    f = _fields(n, pl, seed=1)
    p = h3d.write_tecbin(tmp_path / "t.bin", f, 0.25, pl)
    raw = h3d.read_tecbin(p, raw=True)
    assert raw["pl"] == pl
    for k in ("u", "v", "w", "p", "vis"):
        assert np.array_equal(raw[k], f[k]), k
    real = h3d.read_tecbin(p, n)
    s = (slice(pl, pl + n),) * 3
    assert np.array_equal(real["u"], f["u"][s])


def test_h3d_ghost_rules():
    n, pl, ulid = 5, 4, 0.75
    rng = np.random.default_rng(2)
    real = {k: rng.standard_normal((n, n, n)) for k in ("u", "v", "w", "p")}
    g = {k: h3d.h3d_ghosts(real[k], k, BND, pl, ulid) for k in real}
    r = slice(pl, pl + n)
    for k in real:
        assert np.array_equal(g[k][r, r, r], real[k])
    for m in range(pl):
        lo, hi, slo, shi = pl - 1 - m, pl + n + m, pl + m, pl + n - 1 - m
        # faces away from edges: the single side that writes them
        assert np.array_equal(g["p"][lo, r, r], g["p"][slo, r, r])       # Neumann
        assert np.array_equal(g["p"][r, r, hi], g["p"][r, r, shi])
        assert np.all(g["u"][lo, r, r] == 0.0) and np.all(g["u"][hi, r, r] == 0.0)  # normal
        assert np.all(g["v"][r, lo, r] == 0.0) and np.all(g["w"][r, r, lo] == 0.0)
        assert np.array_equal(g["v"][lo, r, r], -g["v"][slo, r, r])      # wall mirror
        assert np.array_equal(g["u"][r, hi, r], g["u"][r, shi, r])       # slip mirror
        assert np.array_equal(g["w"][r, lo, r], g["w"][r, slo, r])
        assert np.all(g["u"][r, r, hi] == ulid)                          # lid
        assert np.array_equal(g["v"][r, r, hi], g["v"][r, r, shi])       # lid side: +1
        assert np.array_equal(g["u"][r, r, lo], -g["u"][r, r, slo])      # bottom wall
    # an edge ghost takes the last side's rule: the top lid over the east wall
    assert np.all(g["u"][pl + n, :, pl + n] == ulid)


def test_compare_dumps_pairs_the_planes(tmp_path):
    """A Hydro3D dump whose file plane s holds this build's array plane s - 1
    compares with zero difference (p up to a constant)."""
    n, pl, npl = 6, 4, 3
    rng = np.random.default_rng(3)
    mine = {k: rng.standard_normal((n + 2 * npl,) * 3) for k in ("u", "v", "w", "p")}
    h = {k: np.zeros((n + 2 * pl,) * 3) for k in ("u", "v", "w", "p", "vis")}
    for k in mine:
        h[k][1:n + 2 * npl + 1, 1:n + 2 * npl + 1, 1:n + 2 * npl + 1] = mine[k]
    h["p"] += 5.0
    h3d.write_tecbin(tmp_path / "h3d_it025.bin", h, 1.0 / n, pl)
    np.savez(tmp_path / "snap_000025_r000.npz", **mine)
    res = h3d.compare_dumps(tmp_path, tmp_path, n=n, step=25, upto=25)
    assert list(res) == [25]
    for k, (dmax, drms) in res[25].items():
        assert dmax < 1e-12 and drms < 1e-12, k
