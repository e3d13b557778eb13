"""Single-domain topology (the reference's warpflow/topology.py:18-53 for
``mpi.domains = [1, 1, 1]``).

This build runs one cavity on one GPU: every neighbour is -1 (a physical
boundary), the local block is the whole grid and the communicator is a
one-rank stand-in whose reductions return their argument.  A decomposed case
raises: multi-GPU domain decomposition is not part of this build.
"""
from __future__ import annotations

from dataclasses import dataclass

NPL = 3  # ghost layers on each side of every field


class SerialComm:
    """One-rank communicator: allreduce(x) == x (solver.py:120, :134, :150)."""

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def allreduce(self, value, op=None):
        return value

    def Barrier(self):
        return None


@dataclass
class Topology:
    rank: int
    nranks: int
    comm: SerialComm
    sd: tuple
    coords: tuple
    neighbors: dict
    local_ncells: tuple
    global_offsets: tuple


def build_topology(cfg) -> Topology:
    sd = tuple(int(m) for m in cfg.mpi.domains)
    if sd != (1, 1, 1):
        raise ValueError("mpi.domains %s: this build runs a single domain on one GPU "
                         "(use [1, 1, 1])" % (sd,))
    n = (cfg.grid.nx, cfg.grid.ny, cfg.grid.nz)
    return Topology(rank=0, nranks=1, comm=SerialComm(), sd=sd, coords=(0, 0, 0),
                    neighbors={"x": (-1, -1), "y": (-1, -1), "z": (-1, -1)},
                    local_ncells=n, global_offsets=(0, 0, 0))
