"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/warpflow_hip.h declares, the ctypes binding covers them all,
case files load like the reference's, and host-side argument checks reject
bad shapes (no GPU calls are made here)."""
import os
import re

import numpy as np
import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "warpflow_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(wf_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from warpflow import _hip
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(_hip.lib, s)]
    assert not missing, missing
    # the ctypes binding declares exactly the header's entry points
    assert sorted(_hip.exported_symbols()) == syms


def test_library_path_is_in_tree():
    from warpflow import _hip
    assert str(_hip.LIB_PATH).startswith(PKG)


def test_error_strings():
    from warpflow import _hip
    assert b"invalid argument" in _hip.lib.wf_error_string(_hip.WF_EINVAL)


CASE = """
case: t
grid: {nx: 32, ny: 32, nz: 32}
physics: {reynolds: 100, sgs: {enabled: false, model: wale, cw: 0.46}}
initial: {type: rest, noise: 0.01}
numerics: {mg_levels: 3, max_outer_iters: 30, div_tolerance: 1.0e-6}
time: {n_steps: 200, variable_dt: false, dt_init: 1.0e-3}
mpi: {domains: [1, 1, 1]}
"""


def test_case_loading_and_defaults(tmp_path):
    from warpflow.config import load_case
    p = tmp_path / "c.yaml"
    p.write_text(CASE)
    cfg = load_case(p)
    assert cfg.grid.nx == 32 and cfg.physics.reynolds == 100
    assert cfg.physics.sgs_enabled is False and cfg.physics.sgs_cw == 0.46
    assert cfg.time.variable_dt is False and cfg.time.dt_init == 1e-3
    # reference defaults (config.py:15-98)
    assert cfg.boundaries.top == "moving_lid" and cfg.boundaries.lid_velocity == 1.0
    assert cfg.numerics.ghost_layers == 3 and cfg.output.log_interval == 20
    assert cfg.numerics.mg_precision == "fp64" and cfg.initial.seed == 1


@pytest.mark.parametrize("bad,err", [
    ("grid: {nx: 30, ny: 32, nz: 32}\nnumerics: {mg_levels: 3}\n", "divisible by 2"),
    ("grid: {nx: 32, ny: 32, nz: 32}\nmpi: {domains: [3, 1, 1]}\n", "not divisible by"),
    ("numerics: {mg_precision: half}\n", "mg_precision"),
])
def test_case_validation(tmp_path, bad, err):
    from warpflow.config import load_case
    p = tmp_path / "bad.yaml"
    p.write_text(bad)
    with pytest.raises(ValueError, match=err):
        load_case(p)


def test_shipped_cases_load():
    from warpflow.config import load_case
    cases = os.path.join(PKG, "cases")
    files = [f for f in os.listdir(cases) if f.endswith(".yaml")]
    assert files
    for f in files:
        load_case(os.path.join(cases, f))


def test_topology_is_single_domain():
    from warpflow.config import CaseConfig
    from warpflow.topology import build_topology
    cfg = CaseConfig()
    topo = build_topology(cfg)
    assert topo.local_ncells == (80, 80, 80)
    assert all(v == (-1, -1) for v in topo.neighbors.values())
    assert topo.comm.allreduce(3.5) == 3.5
    cfg.mpi.domains = (2, 1, 1)
    with pytest.raises(ValueError):
        build_topology(cfg)


def test_boundary_set_is_validated():
    from warpflow.bcs import BoundaryKernels
    good = {"west": "wall", "east": "wall", "south": "slip", "north": "slip",
            "bottom": "wall", "top": "moving_lid"}
    BoundaryKernels(good)
    with pytest.raises(ValueError):
        BoundaryKernels(dict(good, south="wall"))


def test_grid_descriptor():
    from warpflow import _hip
    g = _hip.WfGrid(10, 12, 14, 3, 20, 18 * 20)
    assert (g.ni, g.nj, g.nk, g.npl, g.sj, g.si) == (10, 12, 14, 3, 20, 360)
    assert np.dtype(np.float64).itemsize == 8
