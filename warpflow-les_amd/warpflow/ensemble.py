"""Ensembles: many independent lid-driven cavities of different sizes and
Reynolds numbers advanced together on one GPU.

The reference runs one cavity per process (``FlowSolver``, solver.py:29-271).
``Ensemble`` advances a list of such cavities with one launch per operator for
all members: every field is one pool holding all members' arrays, and each
batched C-ABI entry applies the single-domain kernel body to every member.
Each member follows the schedule of ``FlowSolver.step`` with the
mixed-precision pressure solve (FP32 V-cycles inside FP64 defect correction)
and a fixed time step; the outer projection loop runs per member on the
device (``active`` masks), so a whole step is captured into HIP graphs and
replayed.  A member's result equals that of a ``FlowSolver`` built from
``member_config`` for the same member.

Step as graphs: graph A = time-step checks, RK2 stages and the projection
up to the outer-iteration test of iteration ``graph_iters + 1``; the host
then reads the members' flags (one synchronisation per step), runs further
outer iterations eagerly for the members that need them, and replays graph
B (the final ghost fills).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _hip
from .config import CaseConfig, validate
from .device import zeros
from .kernels import wale_delta
from .multigrid import COARSE_OPERATOR_SERIAL_MAX_CELLS
from .solver import ALFARK, FIELDS
from .topology import NPL

_BC = {"u": 0, "v": 1, "w": 2, "vis": 3, "pbound": 4, "p_mirror": 5}


@dataclass(frozen=True)
class MemberSpec:
    """One cavity: cells per axis, Reynolds number, noise seed and amplitude."""
    n3: tuple
    reynolds: float
    seed: int
    noise: float = 0.01


def draw_members(nmem=64, sizes=(32, 48, 64, 96), re_range=(100.0, 1000.0), seed=2024,
                 noise=0.01):
    """Seeded ensemble: nx, ny, nz drawn independently from `sizes`, Re
    uniform in `re_range`, noise seed = member index + 1."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(sizes), size=(nmem, 3))
    re = rng.uniform(re_range[0], re_range[1], size=nmem)
    return [MemberSpec(tuple(int(x) for x in n[m]), float(re[m]), m + 1, noise)
            for m in range(nmem)]


def member_config(spec: MemberSpec, dt, mg_levels=4, precision="mixed", mixed_cycles=2,
                  sgs=True, cw=0.46, eps=1e-6, maxcy=30) -> CaseConfig:
    """The single-cavity case of one member (unit cube, fixed dt)."""
    c = CaseConfig()
    c.grid.nx, c.grid.ny, c.grid.nz = spec.n3
    c.physics.reynolds = spec.reynolds
    c.physics.sgs_enabled = sgs
    c.physics.sgs_cw = cw
    c.initial.noise = spec.noise
    c.initial.seed = spec.seed
    c.numerics.mg_levels = mg_levels
    c.numerics.mg_precision = precision
    c.numerics.mg_mixed_cycles = mixed_cycles
    c.numerics.div_tolerance = eps
    c.numerics.max_outer_iters = maxcy
    c.time.variable_dt = False
    c.time.dt_init = dt
    validate(c)
    return c


class Ensemble:
    """Batched cavities (see the module docstring)."""

    N_PRE, N_POST, N_COARSE = 6, 6, 300  # multigrid.py:231-234

    def __init__(self, specs, dt, mg_levels=4, mixed_cycles=2, sgs=True, cw=0.46, eps=1e-6,
                 maxcy=30, ulid=1.0, device="hip:0", graph=True, graph_iters=1):
        _hip.ensure_device(device)
        self.specs = list(specs)
        self.nmem = len(self.specs)
        self.dt = float(dt)
        self.nlev = int(mg_levels)
        self.mixed_cycles = int(mixed_cycles)
        self.sgs = bool(sgs)
        self.cw = float(cw)
        self.eps = float(eps)
        self.maxcy = int(maxcy)
        self.device = device
        self.use_graph = bool(graph)
        self.graph_iters = int(graph_iters)
        self.itime = 0
        self.ctime = 0.0
        mem = (_hip.WfMember * self.nmem)()
        for m, sp in enumerate(self.specs):
            nx, ny, nz = sp.n3
            dx, dy, dz = 1.0 / nx, 1.0 / ny, 1.0 / nz
            mem[m] = _hip.WfMember(nx, ny, nz, sp.seed, dx, dy, dz, 1.0 / sp.reynolds, ulid,
                                   self.cw * wale_delta(dx, dy, dz), sp.noise, self.dt)
        self._mem = mem
        h = C.c_void_p()
        _hip.call("wf_batch_create", mem, self.nmem, self.nlev, C.byref(h))
        self.handle = h.value
        # pool layout
        I64 = C.c_int64
        foff = (I64 * self.nmem)()
        loff = (I64 * (self.nlev * self.nmem))()
        coff = (I64 * self.nmem)()
        fel, cel = I64(0), I64(0)
        lel = (I64 * self.nlev)()
        _hip.call("wf_batch_layout", self.handle, foff, C.byref(fel), loff, lel, coff,
                  C.byref(cel))
        self.field_off = list(foff)
        self.pool = {nm: zeros(fel.value, np.float64, device) for nm in FIELDS + ("w_pre",)}
        self.level_n3 = [[tuple(x >> l for x in sp.n3) for sp in self.specs]
                         for l in range(self.nlev)]
        self.level_off = [[loff[l * self.nmem + m] for m in range(self.nmem)]
                          for l in range(self.nlev)]
        self.lpool = {k: [zeros(max(lel[l], 1), np.float32, device) for l in range(self.nlev)]
                      for k in ("phi", "rhs", "tmp")}
        self._lptr = {k: (C.c_void_p * self.nlev)(*[a.ptr for a in v])
                      for k, v in self.lpool.items()}
        # coarsest-level operators, built per member exactly as Multigrid does
        coarse_cells = [int(np.prod(n)) for n in self.level_n3[-1]]
        # the batched apply is the serial-row form, so a member whose coarsest
        # level is larger relaxes it instead (its single-cavity solver would
        # apply the split form and differ from it by rounding)
        self.use_cop = self.nlev > 1 and max(coarse_cells) <= COARSE_OPERATOR_SERIAL_MAX_CELLS
        self.cop = zeros(max(cel.value, 1), np.float32, device)
        if self.use_cop:
            for m in range(self.nmem):
                d = self._level_desc(self.nlev - 1, m)
                _hip.call("wf_mg_coarse_op_build", C.byref(d), self.N_COARSE,
                          self.cop.ptr + 4 * coff[m], None)
        # device scalars
        self.out4 = zeros(4 * self.nmem, np.float64, device)
        self.rmax = zeros(self.nmem, np.float64, device)
        self.active = zeros(self.nmem, np.int32, device)
        self.iters = zeros(self.nmem, np.int32, device)
        self.status = zeros(self.nmem, np.int32, device)
        s = C.c_void_p()
        _hip.call("wf_stream_create", C.byref(s))
        self.stream = s.value
        self._graphs = None
        self._init_fields()

    # ------------------------------------------------------------------ #
    def _level_desc(self, l, m):
        n3 = self.level_n3[l][m]
        sp = self.specs[m]
        h3 = tuple((1.0 / n) * 2.0 ** l for n in sp.n3)
        off = 4 * self.level_off[l][m]
        return _hip.WfMgLevel(self.lpool["phi"][l].ptr + off, self.lpool["rhs"][l].ptr + off,
                              *n3, 4, *h3, self.lpool["tmp"][l].ptr + off, None, 0, 0)

    def member_shape(self, m):
        return tuple(n + 2 * NPL for n in self.specs[m].n3)

    def field(self, name, m):
        """Member m's array of field `name` (a view into the pool)."""
        return self.pool[name].view(self.field_off[m], self.member_shape(m))

    def synchronize(self):
        _hip.call("wf_stream_synchronize", self.stream)

    def _call(self, fn, *args):
        _hip.call(fn, self.handle, *args)

    def _init_fields(self):
        """FlowSolver.__init__ (solver.py:67-78) for every member."""
        s, P = self.stream, self.pool
        self.synchronize()
        self._call("wf_batch_init_noise", P["u"].ptr, P["v"].ptr, P["w"].ptr, s)
        self.synchronize()
        for m, sp in enumerate(self.specs):  # vis = 1/Re on the whole array
            self.field("vis", m).fill_(1.0 / sp.reynolds)
        _hip.call("wf_synchronize")
        for nm in ("u", "v", "w"):
            self._call("wf_batch_bc_fill", P[nm].ptr, _BC[nm], None, s)
        self._call("wf_batch_bc_fill", P["p"].ptr, _BC["p_mirror"], None, s)
        self.synchronize()

    # ------------------------------------------------------------------ #
    # pieces of one step (FlowSolver.step)                                #
    # ------------------------------------------------------------------ #
    def _fill(self, field, which=None, active=None):
        """Ghost fill `which` (default: the field's own, bcs.py:201-311) of
        pool `field` for every (active) member."""
        self._call("wf_batch_bc_fill", self.pool[field].ptr, _BC[which or field],
                   active.ptr if active is not None else None, self.stream)

    def _fill_uvw(self, with_u=True):
        """The fills of u (unless with_u is false), v and w for every member,
        one launch per axis for the fields."""
        P = self.pool
        self._call("wf_batch_bc_fill_uvw", P["u"].ptr if with_u else None, P["v"].ptr,
                   P["w"].ptr, None, self.stream)

    def _calvel(self, alpha, active=None):
        P = self.pool
        self._call("wf_batch_calvel3", P["ustar"].ptr, P["vstar"].ptr, P["wstar"].ptr,
                   P["p"].ptr, P["u"].ptr, P["v"].ptr, P["w"].ptr, float(alpha),
                   active.ptr if active is not None else None, self.stream)

    def _part_a_head(self):
        """compute_dt, SGS and both RK2 stages (the first storing uoo, voo, woo)."""
        P, s = self.pool, self.stream
        self._call("wf_batch_checkdt", P["u"].ptr, P["v"].ptr, P["w"].ptr, P["vis"].ptr,
                   self.out4.ptr, s)
        # no copies u -> uoo: the first stage's conv+diff stores them on the
        # operator boxes (the cells the second stage reads)
        if self.sgs:
            self._call("wf_batch_wale", P["u"].ptr, P["v"].ptr, P["w"].ptr, P["vis"].ptr, s)
            self._fill("vis")
        else:
            self._call("wf_batch_fill_vis", P["vis"].ptr, s)
        for stage, alpha in enumerate(ALFARK):
            if stage == 0:
                # u, v, w as the initial or end-of-step fills left them: the
                # three conv+diff pairs in one pass (FlowSolver._rk_stage)
                self._call("wf_batch_convdiff3_save", P["u"].ptr, P["v"].ptr, P["w"].ptr,
                           P["vis"].ptr, P["ustar"].ptr, P["vstar"].ptr, P["wstar"].ptr,
                           P["uoo"].ptr, P["voo"].ptr, P["woo"].ptr, float(alpha), s)
            else:
                # u's fill, u's planes 0 and ni-2, w's cells before its fill,
                # the fills of v and w, then the three pairs in one pass
                self._fill("u")
                self._call("wf_batch_convdiff4th_u_edges", P["u"].ptr, P["v"].ptr, P["w"].ptr,
                           P["uoo"].ptr, P["vis"].ptr, P["ustar"].ptr, float(alpha), s)
                self._call("wf_batch_save_cross_ghosts", P["w"].ptr, P["w_pre"].ptr, s)
                self._fill_uvw(with_u=False)
                self._call("wf_batch_convdiff3", P["u"].ptr, P["v"].ptr, P["w"].ptr,
                           P["uoo"].ptr, P["voo"].ptr, P["woo"].ptr, P["vis"].ptr,
                           P["w_pre"].ptr, P["ustar"].ptr, P["vstar"].ptr, P["wstar"].ptr,
                           float(alpha), s)
            self._fill_uvw()
            if stage + 1 < len(ALFARK):
                self._calvel(alpha)
        # projection: reset the per-member loop state; every outer test opens
        # with the calvel/calmas pair in one pass, so a calvel of its own runs
        # here only when no test does (solver.py:178-212)
        if self.maxcy < 1:
            self._calvel(1.0)
        _hip.call("wf_batch_set_int", self.active.ptr, self.nmem, 1, s)
        _hip.call("wf_batch_set_int", self.iters.ptr, self.nmem, 0, s)

    def _outer_test(self, it):
        """calvel + calmas in one pass and the break test of outer iteration
        `it` (solver.py:185-197) for the active members."""
        P = self.pool
        self._call("wf_batch_calvel3_calmas", P["ustar"].ptr, P["vstar"].ptr, P["wstar"].ptr,
                   P["p"].ptr, P["u"].ptr, P["v"].ptr, P["w"].ptr, P["su"].ptr, self.rmax.ptr,
                   1.0, self.active.ptr, self.stream)
        self._call("wf_batch_project_mask", self.rmax.ptr, self.eps, int(it), self.active.ptr,
                   self.iters.ptr, self.status.ptr, self.stream)

    def _outer_rest(self, it):
        """MG solve, p += pp, p ghosts for the active members; after the last
        allowed iteration `it` also the calvel that closes it (the next test
        would otherwise run it)."""
        P, s, a = self.pool, self.stream, self.active.ptr
        self._call("wf_batch_mg_solve_mixed", self._lptr["phi"], self._lptr["rhs"],
                   self._lptr["tmp"], self.cop.ptr if self.use_cop else None, P["su"].ptr,
                   P["pp"].ptr, self.mixed_cycles, self.N_PRE, self.N_POST, self.N_COARSE, 2,
                   a, s)
        self._call("wf_batch_add_ring", P["p"].ptr, P["pp"].ptr, a, s)
        self._fill("p", "pbound", self.active)
        if it == self.maxcy:
            self._calvel(1.0, self.active)

    def _part_a(self):
        self._part_a_head()
        for it in range(1, min(self.graph_iters, self.maxcy) + 1):
            self._outer_test(it)
            self._outer_rest(it)
        if self.graph_iters + 1 <= self.maxcy:
            self._outer_test(self.graph_iters + 1)

    def _part_b(self):
        self._fill_uvw()

    def _capture(self, fn):
        _hip.call("wf_graph_begin", self.stream)
        try:
            fn()
        finally:
            g = C.c_void_p()
            _hip.call("wf_graph_end", self.stream, C.byref(g))
        return g.value

    # ------------------------------------------------------------------ #
    def step(self):
        """One RK2 step of every member; returns per-member iters and rmax."""
        self.itime += 1
        self.ctime += self.dt
        if self.use_graph:
            if self._graphs is None:
                self._graphs = (self._capture(self._part_a), self._capture(self._part_b))
            _hip.call("wf_graph_launch", self._graphs[0], self.stream)
        else:
            self._part_a()
        self.synchronize()
        self._check_dt()
        it = min(self.graph_iters, self.maxcy) + 1
        settled = False  # every member left the loop through its break test
        while it <= self.maxcy:
            if self.status.numpy().any():
                break
            if not self.active.numpy().any():
                settled = True
                break
            self._outer_rest(it)
            if it + 1 > self.maxcy:
                break
            it += 1
            self._outer_test(it)
            self.synchronize()
        status = self.status.numpy()
        if status.any():
            bad = [m for m, s in enumerate(status) if s]
            raise RuntimeError("projection diverged (rmax > 10 or NaN) in members %s at "
                               "itime %d" % (bad, self.itime))
        if not settled:
            self._warn_unconverged()
        if self.use_graph:
            _hip.call("wf_graph_launch", self._graphs[1], self.stream)
        else:
            self._part_b()
        self.synchronize()
        return {"iters": self.iters.numpy().copy(), "rmax": self.rmax.numpy().copy()}

    def cells(self):
        return int(sum(np.prod(sp.n3) for sp in self.specs))

    def _dtmax(self, r):
        """dtmax of each member from its checkdt reductions r[m] = (max|u|,
        max|v|, max|w|, dtvisc), as compute_dt forms it (solver.py:135-141)."""
        out = []
        for m, sp in enumerate(self.specs):
            dtcfl = 1.0e10
            for n, um in zip(sp.n3, r[m, :3]):
                if um > 0.0:
                    dtcfl = min(dtcfl, (1.0 / n) / um)
            out.append(min(r[m, 3], dtcfl))
        return out

    def _check_dt(self):
        """compute_dt's collapse test per member (solver.py:142-144) on this
        step's checkdt reductions; the batch runs a fixed dt, so nothing else
        of compute_dt applies.  A NaN field gives NaN maxima and dtvisc."""
        dtmax = self._dtmax(self.out4.numpy().reshape(self.nmem, 4))
        bad = [m for m, d in enumerate(dtmax) if d <= 0.0 or not (d < 1.0e10)]
        if bad:
            raise RuntimeError("time step collapsed in members %s at itime %d (dtmax %s)"
                               % (bad, self.itime, ["%.3e" % dtmax[m] for m in bad]))

    def _warn_unconverged(self):
        """solver.py:205-209 per member: the members still iterating after the
        last allowed outer iteration."""
        act = self.active.numpy()
        left = [m for m in range(self.nmem) if act[m]]
        if left:
            rmax = self.rmax.numpy()
            print("warning: projection did not converge in members %s after %d iterations, "
                  "rmax = %s" % (left, self.maxcy, ["%.6e" % rmax[m] for m in left]),
                  flush=True)

    def cfl(self):
        """dt / dtmax per member from the last step's checkdt reductions
        (solver.py:122-150)."""
        r = self.out4.numpy().reshape(self.nmem, 4)
        return np.array([self.dt / d for d in self._dtmax(r)])

    def close(self):
        if getattr(self, "_graphs", None):
            for g in self._graphs:
                _hip.call("wf_graph_destroy", g)
            self._graphs = None
        if getattr(self, "handle", None):
            _hip.call("wf_batch_destroy", self.handle)
            self.handle = None
        if getattr(self, "stream", None):
            _hip.call("wf_stream_destroy", self.stream)
            self.stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
