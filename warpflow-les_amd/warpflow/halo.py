"""Halo exchange (the reference's warpflow/halo.py:59-130).

In one domain every neighbour is a physical boundary, so the reference's
``exchange`` returns without launching anything (halo.py:86-87 skips an axis
whose two neighbours are -1).  The same holds here; a topology with a real
neighbour raises instead of silently leaving stale ghosts.
"""
from __future__ import annotations


class HaloExchanger:
    def __init__(self, topo, device, npl):
        self.topo = topo
        self.device = device
        self.npl = npl

    def exchange(self, arr, nl=None):
        for axis in ("x", "y", "z"):
            lo, hi = self.topo.neighbors[axis]
            if lo >= 0 or hi >= 0:
                raise NotImplementedError(
                    "halo exchange across sub-domains is not part of this single-GPU build")
