"""warpflow (MI355X build): the per-step hot path of the warpflow-les
lid-driven-cavity LES -- RK2 fractional step with 4th-order staggered
convection/diffusion, WALE SGS viscosity, checkdt reductions, 3-layer ghost
fills and an RBGS multigrid projection -- on hand-written gfx950 HIP kernels.

Public surface as in the reference package (warpflow/__init__.py:9-14):
``load_case``, ``build_topology``, ``FlowSolver``, ``NPL``.
"""
from .config import load_case
from .topology import NPL, build_topology
from .solver import FlowSolver

__version__ = "0.1.0"
__all__ = ["load_case", "build_topology", "FlowSolver", "NPL", "__version__"]
