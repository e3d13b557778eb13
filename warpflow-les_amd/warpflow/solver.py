"""Time-marching flow solver (the reference's warpflow/solver.py).

One RK2 (Heun) fractional step of the lid-driven-cavity LES, in the
reference's order:

  1. checkdt: max|u|, max|v|, max|w| and the viscous/advective dt limit
  2. uoo, voo, woo <- u, v, w
  3. WALE eddy viscosity + its ghost fill (or vis = 1/Re)
  4. RK2 stages alpha = 1/2, 1: per component fill_bc + conv4th, then per
     component fill_bc + diff4th; calvel(1/2) after the first stage
  5. projection: calvel(1); repeat {calmas; MG Poisson solve; p += pp;
     pbound; calvel(1)} until rmax < eps (at least one solve); final fills

Every operator runs on the GPU through the HIP library; the host keeps only
the scalars the reference keeps (dt, ctime, rmax, cfl) and reads back the
reductions it branches on, as the reference does.  ``fused=True`` (default)
runs each component's conv4th+diff4th pair and the three calvel launches as
single passes, opens each outer iteration with calvel and calmas in one pass,
and has the first stage's conv+diff store uoo, voo, woo on the operator boxes
instead of copying the whole fields; ``fused=False`` replays the reference's
launches one for one.  Both give the same fields (uoo, voo, woo agree on the
operator boxes, the only cells any operator reads).
"""
from __future__ import annotations

import numpy as np

from . import kernels as K
from .bcs import BoundaryKernels
from .device import copy, zeros
from .halo import HaloExchanger
from .multigrid import Multigrid
from .topology import NPL

ALFARK = (0.5, 1.0)  # Heun RK2 stage coefficients (solver.py:26)
FIELDS = ("u", "v", "w", "p", "ustar", "vstar", "wstar",
          "uoo", "voo", "woo", "vis", "su", "pp")


class FlowSolver:
    def __init__(self, cfg, topo, device=None, fused=True):
        self.cfg = cfg
        self.topo = topo
        self.device = device
        self.fused = fused

        ni, nj, nk = topo.local_ncells
        self.ncells = (ni, nj, nk)
        self.dx = cfg.domain.lx / cfg.grid.nx
        self.dy = cfg.domain.ly / cfg.grid.ny
        self.dz = cfg.domain.lz / cfg.grid.nz
        self.rey = 1.0 / cfg.physics.reynolds

        shape = (ni + 2 * NPL, nj + 2 * NPL, nk + 2 * NPL)
        # rows padded so that real cell k = NPL of every row starts a 128-byte line
        self.f = {name: zeros(shape, np.float64, device, aligned=True) for name in FIELDS}
        if fused:  # w's pre-fill cells for the second stage's v/w pass
            self.f["w_pre"] = zeros(shape, np.float64, device, aligned=True)

        self.halo = HaloExchanger(topo, device, NPL)
        b = cfg.boundaries
        self.bnd = {"west": b.west, "east": b.east, "south": b.south,
                    "north": b.north, "bottom": b.bottom, "top": b.top}
        self.bcs = BoundaryKernels(self.bnd, device)
        self.ulid = b.lid_velocity

        self.mg = self._make_mg()
        # reduction scratch: [0:4] maxabs / calmas, [4:8] checkdt; one read-back
        # returns both
        self._buf = zeros(8, np.float64, device)
        self._rbuf = self._buf.view(0, (4,))
        self._dbuf = self._buf.view(4, (4,))
        self._dt_pending = False

        self.dt = cfg.time.dt_init
        self.ctime = 0.0
        self.itime = 0
        self.rmax = 0.0
        self.cfl = 0.0
        self.outer_iters = 0
        self.rmax_history = []

        # initial condition: rest + deterministic hash noise (solver.py:67-78)
        amp = cfg.initial.noise if cfg.initial.type == "rest" else 0.0
        if amp > 0.0:
            K.init_noise(self.ncells, self.f["u"], self.f["v"], self.f["w"], amp,
                         getattr(cfg.initial, "seed", 1), NPL)
        self.f["vis"].fill_(self.rey)
        self.bcs.apply(self.f, self.ncells, self.ulid, NPL)
        # u, v, w are as a fill left them: see _rk_stage
        self._fill_stable = True
        # the second stage's u/v/w pass (off: the v/w pass, with fused_vw)
        self.fused_uvw = fused
        # the second stage's v/w pass (off: one pass per component)
        self.fused_vw = fused

    def _make_mg(self):
        if getattr(self.cfg.numerics, "mg_precision", "fp64") == "mixed":
            from .mixed import MixedMultigrid
            return MixedMultigrid(self.cfg, self.topo, self.device, self.dx, self.dy, self.dz)
        return Multigrid(self.cfg, self.topo, self.device, self.dx, self.dy, self.dz)

    # ------------------------------------------------------------------ #
    # ghost fills                                                        #
    # ------------------------------------------------------------------ #
    def fill_bc_uvw(self, alfabc=1):
        """fill_bc of u, v, w back to back (solver.py:88-97 three times): the
        three halo exchanges, then one fill launch per axis for the three
        fields (independent of each other, so the order within each field is
        the reference's)."""
        for name in ("u", "v", "w"):
            self.halo.exchange(self.f[name])
        self.bcs.apply_uvw(self.f["u"], self.f["v"], self.f["w"], self.ncells[0], self.ulid)

    def refill_bc_uvw(self, alfabc=1):
        """fill_bc_uvw where u, v, w were filled since their last change (the
        second fills of an RK stage): the same ghosts, computed on the edge
        lines only."""
        for name in ("u", "v", "w"):
            self.halo.exchange(self.f[name])
        self.bcs.refill_uvw(self.f["u"], self.f["v"], self.f["w"], self.ncells[0], self.ulid)

    def fill_bc_vw(self, alfabc=1):
        """fill_bc of v, then w (solver.py:88-97 twice), as fill_bc_uvw does
        it for three fields."""
        for name in ("v", "w"):
            self.halo.exchange(self.f[name])
        self.bcs.apply_vw(self.f["v"], self.f["w"], self.ncells[0])

    def fill_bc(self, name, alfabc=1):
        """solver.py:88-97: one boundu/boundv/boundw call for one field."""
        self.halo.exchange(self.f[name])
        n = self.ncells[0]
        if name == "u":
            self.bcs.apply_u(self.f["u"], n, self.ulid, alfabc)
        elif name == "v":
            self.bcs.apply_v(self.f["v"], n, alfabc)
        elif name == "w":
            self.bcs.apply_w(self.f["w"], n, alfabc)

    def _op_dim(self, comp):
        """solver.py:99-110: u (ni-1,nj,nk), v (ni,nj-1,nk), w (ni,nj,nk-1)."""
        ni, nj, nk = self.ncells
        if comp == "u":
            return (ni - 1, nj, nk)
        if comp == "v":
            return (ni, nj - 1, nk)
        if comp == "w":
            return (ni, nj, nk - 1)
        return (ni, nj, nk)

    # ------------------------------------------------------------------ #
    # time step control (checkdt)                                        #
    # ------------------------------------------------------------------ #
    def _maxabs(self, name):
        """solver.py:115-120"""
        self._rbuf.zero_()
        K.maxabs_reduce(self.ncells, self.f[name], self._rbuf, NPL)
        loc = float(self._rbuf.numpy()[0])
        return self.topo.comm.allreduce(loc)

    def _reductions(self):
        if not self.fused:
            umax, vmax, wmax = (self._maxabs(c) for c in ("u", "v", "w"))
            self._rbuf.fill_(1.0e10)
            K.dtvisc_reduce(self.ncells, self.f["u"], self.f["v"], self.f["w"],
                            self.f["vis"], self._rbuf, self.dx, self.dy, self.dz, NPL)
            return umax, vmax, wmax, float(self._rbuf.numpy()[0])
        self._launch_checkdt()
        r = self._dbuf.numpy()
        return float(r[0]), float(r[1]), float(r[2]), float(r[3])

    def _launch_checkdt(self):
        K.checkdt(self.ncells, self.f["u"], self.f["v"], self.f["w"], self.f["vis"],
                  self._dbuf, self.dx, self.dy, self.dz, NPL)

    def compute_dt(self, defer=False):
        """solver.py:122-150.  With a fixed dt nothing in a step needs the
        reductions before the projection, so step() (defer=True, fused path)
        only launches checkdt here and finishes (cfl, the collapse check)
        when the first calmas result is read back in the same step -- one
        host round trip fewer per step, same values."""
        if defer and self.fused and not self.cfg.time.variable_dt:
            self._launch_checkdt()
            self._dt_pending = True
            return
        self._finish_dt(*self._reductions())

    def _finish_dt(self, umax, vmax, wmax, dtvisc):
        dtvisc = self.topo.comm.allreduce(dtvisc)
        big = 1.0e10
        dtcfl = big
        for d, um in zip((self.dx, self.dy, self.dz), (umax, vmax, wmax)):
            if um > 0.0:
                dtcfl = min(dtcfl, d / um)
        dtmax = min(dtvisc, dtcfl)
        if dtmax <= 0.0 or not (dtmax < big):
            raise RuntimeError("time step collapsed (dtmax = %.3e)" % dtmax)
        self.cfl = self.dt / dtmax
        if self.cfg.time.variable_dt:
            dt_new = min(self.dt * self.cfg.time.dt_growth,
                         self.cfg.time.safety_factor * dtmax)
            self.dt = self.topo.comm.allreduce(dt_new)

    # ------------------------------------------------------------------ #
    # projection pieces                                                  #
    # ------------------------------------------------------------------ #
    def _calvel(self, alpha):
        """solver.py:155-166: calvel for u, v, w over their operator ranges,
        then halo exchange only (no BC fill)."""
        f = self.f
        if self.fused:
            K.calvel3(self.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"],
                      f["u"], f["v"], f["w"], self.dx, self.dy, self.dz, self.dt,
                      alpha, NPL)
        else:
            for ker, star, comp, d in ((K.calvel_u, "ustar", "u", self.dx),
                                       (K.calvel_v, "vstar", "v", self.dy),
                                       (K.calvel_w, "wstar", "w", self.dz)):
                ker(self._op_dim(comp), f[star], f["p"], f[comp], d, self.dt, alpha, NPL)
        for name in ("u", "v", "w"):
            self.halo.exchange(f[name])

    def _calmas(self):
        """solver.py:168-176"""
        self._rbuf.zero_()
        K.calmas(self.ncells, self.f["u"], self.f["v"], self.f["w"], self.f["su"],
                 self._rbuf, self.dx, self.dy, self.dz, self.dt, NPL)
        self._read_calmas()

    def _read_calmas(self):
        """Read back rmax (and the deferred checkdt values with it)."""
        r = self._buf.numpy()
        if self._dt_pending:
            self._dt_pending = False
            self._finish_dt(float(r[4]), float(r[5]), float(r[6]), float(r[7]))
        loc = float(r[0])
        self.rmax = self.topo.comm.allreduce(loc)

    def _calvel_calmas(self, store_uvw=True):
        """``_calvel(1.0)`` then ``_calmas()`` in one pass (solver.py:155-176 as
        _project runs them back to back): calvel3_calmas writes u, v, w, su
        and rmax as the two kernels do; the halo exchange between them does
        nothing on one domain and is issued afterwards.  store_uvw=False
        leaves u, v, w unwritten (the first outer iteration: see _project)."""
        f = self.f
        self._rbuf.zero_()
        # mixed form: the pass also writes (float)su into the FP32 level-0 rhs,
        # the first defect of the solve that may follow
        rhs32 = self.mg.levels[0].rhs if hasattr(self.mg, "add_result") else None
        K.calvel3_calmas(self.ncells, f["ustar"], f["vstar"], f["wstar"], f["p"], f["u"],
                         f["v"], f["w"], f["su"], self._rbuf, self.dx, self.dy, self.dz,
                         self.dt, 1.0, NPL, rhs32, store_uvw=store_uvw)
        for name in ("u", "v", "w"):
            self.halo.exchange(f[name])
        self._read_calmas()

    def _project(self):
        """solver.py:178-212: outer Poisson/projection iterations.  With the
        fused kernels each iteration opens with the calvel/calmas pair in one
        pass, so the calvel that closes an iteration runs separately only when
        the loop ends without its break."""
        f = self.f
        eps = self.cfg.numerics.div_tolerance
        maxcy = self.cfg.numerics.max_outer_iters
        if not self.fused:
            self._calvel(1.0)
        iters = 0
        self.rmax_history = []
        for it in range(1, maxcy + 1):
            iters = it
            if self.fused:
                # the first iteration cannot exit the loop (it > 1 below), so
                # the velocities its calvel writes would be overwritten unread
                # by the next iteration's calvel (or the closing one): its pass
                # forms su and rmax without storing them
                self._calvel_calmas(store_uvw=it > 1)
            else:
                self._calmas()
            self.rmax_history.append(self.rmax)
            if self.rmax < eps and it > 1:
                break
            if self.rmax > 10.0:
                if self.fused and it == 1:
                    self._calvel(1.0)  # the velocities the reference leaves behind
                raise RuntimeError("projection diverged: rmax = %.6e at itime %d"
                                   % (self.rmax, self.itime))
            if self.fused and hasattr(self.mg, "add_result"):
                # mixed form: the pressure-gauge shift of pp is applied inside
                # p += pp (pp itself is left unshifted; p gets the same bits)
                self.mg.solve(f["su"], f["pp"], NPL, defer_mean=True, rhs_ready=True)
                self.mg.add_result(f["p"], f["pp"], NPL)
            else:
                self.mg.solve(f["su"], f["pp"], NPL)
                K.add_inplace(tuple(n + 2 for n in self.ncells), f["p"], f["pp"], NPL)
            self.halo.exchange(f["p"])
            self.bcs.apply_pbound(f["p"], self.ncells[0])
            if not self.fused:
                self._calvel(1.0)
        else:
            if self.fused:
                self._calvel(1.0)
            if self.topo.rank == 0:
                print("warning: projection did not converge, rmax = %.6e after %d "
                      "iterations" % (self.rmax, maxcy), flush=True)
        self.outer_iters = iters
        self.fill_bc_uvw(0)
        self._fill_stable = True

    # ------------------------------------------------------------------ #
    # one time step                                                      #
    # ------------------------------------------------------------------ #
    def _sgs(self):
        """solver.py:227-237"""
        f = self.f
        if self.cfg.physics.sgs_enabled:
            K.wale_viscosity(self.ncells, f["u"], f["v"], f["w"], f["vis"], self.dx,
                             self.dy, self.dz, self.rey, self.cfg.physics.sgs_cw, NPL)
            self.bcs.apply_vis(f["vis"], self.ncells[0], self.rey)
        else:
            K.fill_const(self.ncells, f["vis"], self.rey, NPL)

    def _sgs_checkdt(self):
        """compute_dt (solver.py:122-150) and the WALE SGS step
        (solver.py:227-237) in one pass; dt is finished as compute_dt does:
        at once with a variable dt, else with the first calmas read-back."""
        f = self.f
        K.wale_checkdt(self.ncells, f["u"], f["v"], f["w"], f["vis"], self._dbuf, self.dx,
                       self.dy, self.dz, self.rey, self.cfg.physics.sgs_cw, NPL)
        self.bcs.apply_vis(f["vis"], self.ncells[0], self.rey)
        if self.cfg.time.variable_dt:
            r = self._dbuf.numpy()
            self._finish_dt(float(r[0]), float(r[1]), float(r[2]), float(r[3]))
        else:
            self._dt_pending = True

    def _rk_stage(self, stage, alpha):
        """solver.py:249-265"""
        f = self.f
        dt = self.dt
        d3 = (self.dx, self.dy, self.dz)
        alfabc = 1 if stage == 0 else 0
        comps = (("u", K.conv4th_u, K.diff4th_u, K.convdiff4th_u),
                 ("v", K.conv4th_v, K.diff4th_v, K.convdiff4th_v),
                 ("w", K.conv4th_w, K.diff4th_w, K.convdiff4th_w))
        if self.fused and stage == 0 and self._fill_stable:
            # First stage, u, v, w as the end-of-step (or initial) fills left
            # them: a second fill changes only deep edge ghosts that no
            # conv4th/diff4th stencil reads, and a third fill changes nothing
            # more.  So the three conv+diff pairs run as one pass on the fields
            # as they are, with the values of the reference's fill/conv/fill/
            # diff sequence, and one fill after it leaves the ghosts as that
            # sequence's two fills do.
            self._fill_stable = False
            K.convdiff3_save(self.ncells, f["u"], f["v"], f["w"], f["vis"], f["ustar"],
                             f["vstar"], f["wstar"], f["uoo"], f["voo"], f["woo"], *d3, dt,
                             alpha, NPL)
            self.refill_bc_uvw(alfabc)
            return
        self._fill_stable = False
        if self.fused and stage > 0 and self.fused_uvw:
            # Second stage, the three pairs in one pass after the three fills.
            # conv_u reads v and w before their fills, which change only the
            # x-ghost planes -1 and ni of what conv_u reads; u's planes 0 and
            # ni-2 reach them, so those two run right after u's fill, and the
            # pass leaves them alone.  v's z-term reads w's cells from before
            # w's fill, which save_cross_ghosts keeps.
            self.fill_bc("u", alfabc)
            K.convdiff4th_u_edges(self.ncells, f["u"], f["v"], f["w"], f["uoo"], f["vis"],
                                  f["ustar"], *d3, dt, alpha, NPL)
            # w's cells before its fill, then v's and w's fills in one pair
            # of launches (independent fields, the reference's order within
            # each)
            K.save_cross_ghosts(self.ncells, f["w"], f["w_pre"], NPL)
            self.fill_bc_vw(alfabc)
            K.convdiff3(self.ncells, f["u"], f["v"], f["w"], f["uoo"], f["voo"], f["woo"],
                        f["vis"], f["w_pre"], f["ustar"], f["vstar"], f["wstar"], *d3, dt,
                        alpha, NPL)
            self.refill_bc_uvw(alfabc)
            return
        if self.fused and stage > 0 and self.fused_vw:
            # Second stage: u's pass right after u's fill (it reads v and w
            # before their fills, as the reference's conv_u does), then v's
            # and w's passes as one after both fills; v's z-term reads w's
            # cells from before w's fill, which save_cross_ghosts keeps.
            self.fill_bc("u", alfabc)
            K.convdiff4th_u(self._op_dim("u"), f["u"], f["v"], f["w"], f["uoo"], f["vis"],
                            f["ustar"], *d3, dt, alpha, NPL)
            self.fill_bc("v", alfabc)
            K.save_cross_ghosts(self.ncells, f["w"], f["w_pre"], NPL)
            self.fill_bc("w", alfabc)
            K.convdiff_vw(self.ncells, f["u"], f["v"], f["w"], f["voo"], f["woo"], f["vis"],
                          f["w_pre"], f["vstar"], f["wstar"], *d3, dt, alpha, NPL)
            self.fill_bc_uvw(alfabc)
            return
        if self.fused:
            # conv+diff of component c right after its first fill; the second
            # fills still run, in order, for their effect on the ghosts.  In
            # the first stage c_oo is c itself (step() skips the copies):
            # the kernel takes it from c and stores it into c_oo on its box.
            # (Second stage: the cross stencils read the other components'
            # ghosts as calvel left them, before their fills.)
            save = (K.convdiff4th_save_u, K.convdiff4th_save_v, K.convdiff4th_save_w)
            for q, (c, _, _, kcd) in enumerate(comps):
                self.fill_bc(c, alfabc)
                if stage == 0:
                    save[q](self._op_dim(c), f["u"], f["v"], f["w"], f[c + "oo"], f["vis"],
                            f[c + "star"], *d3, dt, alpha, NPL)
                else:
                    kcd(self._op_dim(c), f["u"], f["v"], f["w"], f[c + "oo"], f["vis"],
                        f[c + "star"], *d3, dt, alpha, NPL)
            self.fill_bc_uvw(alfabc)
            return
        for c, kconv, _, _ in comps:
            self.fill_bc(c, alfabc)
            kconv(self._op_dim(c), f["u"], f["v"], f["w"], f[c + "oo"], f[c + "star"],
                  *d3, dt, alpha, NPL)
        for c, _, kdiff, _ in comps:
            self.fill_bc(c, alfabc)
            kdiff(self._op_dim(c), f[c], f["vis"], f[c + "star"], *d3, dt, alpha, NPL)

    def step(self):
        """solver.py:217-271"""
        self.itime += 1
        f = self.f
        # fused, with WALE on: checkdt's reductions come out of the WALE pass
        # (nothing between them changes u, v, w or vis; each cell's vis is
        # read before WALE overwrites it)
        sgs_chk = self.fused and self.cfg.physics.sgs_enabled
        if sgs_chk:
            self._sgs_checkdt()
        else:
            self.compute_dt(defer=True)
        self.ctime += self.dt
        if not self.fused:
            copy(f["uoo"], f["u"])
            copy(f["voo"], f["v"])
            copy(f["woo"], f["w"])
        # fused: the first stage's conv+diff stores u, v, w into uoo, voo, woo
        # on the operator boxes, the only cells the second stage reads (the
        # copies' other cells are not refreshed); the SGS and fills before it
        # leave those cells of u, v, w as they are here
        if not sgs_chk:
            self._sgs()
        for stage, alpha in enumerate(ALFARK):
            self._rk_stage(stage, alpha)
            if stage + 1 < len(ALFARK):
                self._calvel(alpha)
        self._project()
        if self._dt_pending:  # no calmas ran (max_outer_iters < 1)
            self._dt_pending = False
            r = self._dbuf.numpy()
            self._finish_dt(float(r[0]), float(r[1]), float(r[2]), float(r[3]))
        return {"ctime": self.ctime, "dt": self.dt, "rmax": self.rmax,
                "iters": self.outer_iters, "cfl": self.cfl}
