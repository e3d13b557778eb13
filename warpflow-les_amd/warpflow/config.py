"""YAML case files -> ``CaseConfig``.

Same schema and defaults as the reference's ``warpflow/config.py:15-132``
(sections ``case, domain, grid, physics{reynolds, sgs{enabled, model, cw}},
initial, boundaries, numerics, time, output, mpi``) so existing case files
load unchanged, plus optional keys this build reads:
``numerics.mg_precision`` (``fp64`` = the reference's V-cycle, ``mixed`` =
FP32 V-cycles inside FP64 defect correction), ``numerics.mg_mixed_cycles``
(FP64-refined FP32 V-cycles per solve in the mixed form, default 2) and
``initial.seed`` (the init_noise seed, 1 in the reference, solver.py:70-73).
"""
from __future__ import annotations

from dataclasses import dataclass, field, fields
from pathlib import Path

import yaml


@dataclass
class DomainCfg:
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0


@dataclass
class GridCfg:
    nx: int = 80
    ny: int = 80
    nz: int = 80


@dataclass
class PhysicsCfg:
    reynolds: float = 400.0
    sgs_enabled: bool = True
    sgs_model: str = "wale"
    sgs_cw: float = 0.46


@dataclass
class InitialCfg:
    type: str = "rest"
    noise: float = 0.01
    seed: int = 1


@dataclass
class BoundaryCfg:
    west: str = "wall"
    east: str = "wall"
    south: str = "slip"
    north: str = "slip"
    bottom: str = "wall"
    top: str = "moving_lid"
    lid_velocity: float = 1.0


@dataclass
class NumericsCfg:
    time_scheme: str = "rk2"
    convection: str = "cds4"
    diffusion: str = "explicit4"
    pressure_solver: str = "multigrid"
    mg_levels: int = 3
    max_outer_iters: int = 30
    div_tolerance: float = 1.0e-6
    ghost_layers: int = 3
    mg_precision: str = "fp64"
    mg_mixed_cycles: int = 2


@dataclass
class TimeCfg:
    n_steps: int = 1000
    variable_dt: bool = True
    dt_init: float = 3.0e-4
    safety_factor: float = 0.8
    dt_growth: float = 1.1


@dataclass
class OutputCfg:
    n_out: int = 1500
    log_interval: int = 20
    outdir: str = "output"


@dataclass
class MpiCfg:
    domains: tuple = (1, 1, 1)


@dataclass
class CaseConfig:
    name: str = "cavity"
    domain: DomainCfg = field(default_factory=DomainCfg)
    grid: GridCfg = field(default_factory=GridCfg)
    physics: PhysicsCfg = field(default_factory=PhysicsCfg)
    initial: InitialCfg = field(default_factory=InitialCfg)
    boundaries: BoundaryCfg = field(default_factory=BoundaryCfg)
    numerics: NumericsCfg = field(default_factory=NumericsCfg)
    time: TimeCfg = field(default_factory=TimeCfg)
    output: OutputCfg = field(default_factory=OutputCfg)
    mpi: MpiCfg = field(default_factory=MpiCfg)


def _section(cls, raw: dict | None):
    raw = dict(raw or {})
    known = {f.name for f in fields(cls)}
    unknown = set(raw) - known
    if unknown:
        raise TypeError("%s: unknown key(s) %s" % (cls.__name__, sorted(unknown)))
    return cls(**raw)


def case_from_dict(raw: dict) -> CaseConfig:
    """Build and validate a CaseConfig from the parsed YAML mapping."""
    cfg = CaseConfig(name=raw.get("case", "cavity"))
    cfg.domain = _section(DomainCfg, raw.get("domain"))
    cfg.grid = _section(GridCfg, raw.get("grid"))
    phys = dict(raw.get("physics") or {})
    sgs = dict(phys.pop("sgs", None) or {})
    cfg.physics = PhysicsCfg(reynolds=phys.get("reynolds", 400.0),
                             sgs_enabled=sgs.get("enabled", True),
                             sgs_model=sgs.get("model", "wale"),
                             sgs_cw=sgs.get("cw", 0.46))
    cfg.initial = _section(InitialCfg, raw.get("initial"))
    cfg.boundaries = _section(BoundaryCfg, raw.get("boundaries"))
    cfg.numerics = _section(NumericsCfg, raw.get("numerics"))
    cfg.time = _section(TimeCfg, raw.get("time"))
    cfg.output = _section(OutputCfg, raw.get("output"))
    mpi = dict(raw.get("mpi") or {})
    if "domains" in mpi:
        mpi["domains"] = tuple(int(m) for m in mpi["domains"])
    cfg.mpi = _section(MpiCfg, mpi)
    validate(cfg)
    return cfg


def validate(cfg: CaseConfig) -> None:
    """The reference's divisibility checks (config.py:121-131)."""
    lev = cfg.numerics.mg_levels
    if lev < 1:
        raise ValueError("numerics.mg_levels must be >= 1")
    for axis, (n, m) in enumerate(zip((cfg.grid.nx, cfg.grid.ny, cfg.grid.nz),
                                      cfg.mpi.domains)):
        if n % m != 0:
            raise ValueError(f"grid axis {axis} ({n} cells) not divisible by "
                             f"mpi domains ({m})")
        if (n // m) % (2 ** (lev - 1)) != 0:
            raise ValueError(f"sub-domain cells on axis {axis} ({n // m}) must be "
                             f"divisible by 2**(mg_levels-1)={2 ** (lev - 1)}")
    if cfg.numerics.mg_precision not in ("fp64", "mixed"):
        raise ValueError("numerics.mg_precision must be 'fp64' or 'mixed'")
    if cfg.numerics.mg_mixed_cycles < 1:
        raise ValueError("numerics.mg_mixed_cycles must be >= 1")


def load_case(path) -> CaseConfig:
    """Read a YAML case file (config.py:101-132)."""
    return case_from_dict(yaml.safe_load(Path(path).read_text()) or {})
