"""Snapshot files (the reference's warpflow/io.py:9-27): name pattern, the
arrays and scalars they hold, and that they load without pickle."""
import types

import numpy as np

from warpflow.io import read_snapshot, snapshot_path, write_snapshot


class _Arr:
    def __init__(self, a):
        self.a = a

    def numpy(self):
        return self.a.copy()


def fake_solver(tmp_path, n3=(4, 5, 6)):
    rng = np.random.default_rng(3)
    shape = tuple(n + 6 for n in n3)
    f = {k: _Arr(rng.standard_normal(shape)) for k in ("u", "v", "w", "p", "vis")}
    cfg = types.SimpleNamespace(output=types.SimpleNamespace(outdir=str(tmp_path / "out")),
                                grid=types.SimpleNamespace(nx=n3[0], ny=n3[1], nz=n3[2]))
    topo = types.SimpleNamespace(rank=0, global_offsets=(0, 0, 0))
    return types.SimpleNamespace(f=f, cfg=cfg, topo=topo, ncells=n3, ctime=0.125, dt=1e-3)


def test_snapshot_layout(tmp_path):
    sol = fake_solver(tmp_path)
    path = write_snapshot(sol, 42)
    assert path == snapshot_path(tmp_path / "out", 42, 0)
    assert path.name == "snap_000042_r000.npz"
    z = read_snapshot(path)
    assert set(z) == {"u", "v", "w", "p", "ctime", "dt", "offsets", "ncells", "gshape"}
    for k in ("u", "v", "w", "p"):
        assert z[k].dtype == np.float64 and np.array_equal(z[k], sol.f[k].a)
    assert z["ctime"] == 0.125 and z["dt"] == 1e-3
    assert list(z["ncells"]) == [4, 5, 6] and list(z["gshape"]) == [4, 5, 6]
    assert list(z["offsets"]) == [0, 0, 0]


def test_restart_snapshot_adds_vis_and_itime(tmp_path):
    sol = fake_solver(tmp_path)
    z = read_snapshot(write_snapshot(sol, 7, restart=True))
    assert set(z) == {"u", "v", "w", "p", "vis", "itime", "ctime", "dt", "offsets", "ncells",
                      "gshape"}
    assert np.array_equal(z["vis"], sol.f["vis"].a) and int(z["itime"]) == 7
