"""The CLI (run.py) on a small case, and restart from a snapshot: with a
fixed time step and WALE on, a run continued from the snapshot of step k
equals the uninterrupted run bit for bit (every other field is recomputed
from u, v, w and p within a step)."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "warpflow-les_amd")


def test_cli_log_and_snapshots(gpu, tmp_path, capsys):
    sys.path.insert(0, ROOT)
    import run
    case = os.path.join(ROOT, "cases", "smoke24.yaml")
    assert run.main(["--case", case, "--steps", "4", "--outdir", str(tmp_path)]) == 0
    out = capsys.readouterr().out.splitlines()
    assert out[0].startswith("warpflow-les 0.1.0: case=")
    assert out[-1].startswith("done: itime=4 ctime=")
    snaps = sorted(p.name for p in tmp_path.iterdir())
    assert "snap_000004_r000.npz" in snaps


def test_restart_from_snapshot_is_exact(gpu, tmp_path):
    from warpflow import FlowSolver, build_topology
    from warpflow.config import CaseConfig, validate
    from warpflow.io import load_into, write_snapshot
    c = CaseConfig()
    c.grid.nx = c.grid.ny = c.grid.nz = 24
    c.time.variable_dt = False
    c.time.dt_init = 1e-3
    c.output.outdir = str(tmp_path)
    validate(c)
    a = FlowSolver(c, build_topology(c), "hip:0")
    for _ in range(2):
        a.step()
    path = write_snapshot(a, 2)
    for _ in range(2):
        a.step()
    b = FlowSolver(c, build_topology(c), "hip:0")
    load_into(b, path)
    for _ in range(2):
        b.step()
    for nm in ("u", "v", "w", "p"):
        assert np.array_equal(a.f[nm].numpy(), b.f[nm].numpy()), nm
    assert a.ctime == b.ctime


def test_cli_restart_is_exact_with_adaptive_dt(gpu, tmp_path, capsys):
    """run.py to step 8 in one go, and to step 4 then --restart to step 8
    (24^3, adaptive dt, WALE): the final snapshots agree bit for bit, because
    the restart snapshot carries vis (read by the dt limit) and itime."""
    sys.path.insert(0, ROOT)
    import run
    from warpflow.io import read_snapshot
    case = os.path.join(ROOT, "cases", "smoke24.yaml")
    a, b, c = (tmp_path / x for x in "abc")
    assert run.main(["--case", case, "--steps", "8", "--outdir", str(a)]) == 0
    assert run.main(["--case", case, "--steps", "4", "--outdir", str(b)]) == 0
    assert run.main(["--case", case, "--steps", "8", "--outdir", str(c),
                     "--restart", str(b / "snap_000004_r000.npz")]) == 0
    logs = capsys.readouterr().out
    assert logs.count("done: itime=8 ") == 2
    x = read_snapshot(a / "snap_000008_r000.npz")
    y = read_snapshot(c / "snap_000008_r000.npz")
    assert int(x["itime"]) == int(y["itime"]) == 8
    for k in ("u", "v", "w", "p", "vis", "ctime", "dt"):
        assert np.array_equal(x[k], y[k]), k


def test_h3d_restart_export_of_a_fresh_solver(gpu, tmp_path):
    """export (tests/h3d_bridge.py:59-115 in the reference): the restart of
    a freshly initialised cavity holds the solver's real cells of u, v, w, p,
    Hydro3D's 4 ghost layers and vis = 1/Re."""
    from warpflow import FlowSolver, build_topology, load_case
    from warpflow import h3d
    cfg = load_case(os.path.join(ROOT, "cases", "cavity_serial.yaml"))
    cfg.grid.nx = cfg.grid.ny = cfg.grid.nz = 16
    sol = FlowSolver(cfg, build_topology(cfg), "hip:0")
    path = h3d.export_restart(sol, tmp_path)
    assert (tmp_path / "final_ctime.dat").read_text().startswith("       0")
    got = h3d.read_tecbin(path, raw=True)
    n, pl = 16, h3d.PL_H3D
    r = slice(pl, pl + n)
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(got[k][r, r, r], sol.f[k].numpy()[3:3 + n, 3:3 + n, 3:3 + n]), k
    assert np.all(got["vis"] == sol.rey)
    assert np.all(got["u"][r, r, pl + n:] == sol.ulid)  # lid ghosts
