/* warpflow_hip.h -- C ABI of the MI355X (gfx950) per-step hot path of the
 * warpflow-les lid-driven-cavity LES solver.
 *
 * The reference (dr-chunshui/warpflow-les) launches NVIDIA Warp kernels with
 *     wp.launch(kernel, dim=(d0, d1, d2), inputs=[arrays..., scalars...])
 * over C-ordered float64 wp.array3d fields.  Each entry below replaces one
 * such kernel (or one host method that issues a fixed launch sequence); the
 * citation names the reference file:line it stands in for.
 *
 * Conventions
 *  - Every entry returns an int: 0 on success, else a hipError_t code
 *    (wf_error_string() turns it into text).  Arguments are checked on the
 *    host before any launch; a bad shape returns WF_EINVAL (10001) and launches
 *    nothing.
 *  - Arrays are device pointers.  Fields use wf_grid: real cells ni, nj, nk,
 *    npl ghost layers, element strides sj (of j) and si (of i); k is
 *    unit-stride.  The launch box of a kernel is given as (d0, d1, d2) and
 *    starts at full index (npl, npl, npl), exactly like the reference's
 *    `i += npl` after wp.tid().
 *  - `stream` is a hipStream_t (NULL = the null stream).  Work is enqueued
 *    asynchronously; reductions write device scalars that the caller reads
 *    back (the reference's 1-element wp.array + .numpy()).
 *  - No entry point allocates device memory except wf_malloc and the
 *    multigrid set-up helpers that say so.
 */
#ifndef WARPFLOW_HIP_H
#define WARPFLOW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_EINVAL 10001
/* The call needs more workspace than the library's static scratch holds
 * (wf_mg_solve_mixed_static). */
#define WF_ESCRATCH 10002
/* Size of the static device scratch behind wf_mg_solve_mixed_static. */
#define WF_STATIC_SCRATCH_BYTES ((size_t)1 << 20)

typedef struct wf_grid {
  int32_t ni, nj, nk; /* real cells per axis */
  int32_t npl;        /* ghost layers per side */
  int64_t sj, si;     /* element strides of j and i */
} wf_grid;

/* One multigrid level: arrays of shape (n0+2, n1+2, n2+2) with a one-cell
 * ring (warpflow/multigrid.py:207-220).  prec = 8 (float64) or 4 (float32).
 * h0..h2 are the level spacings dx*2^l, dy*2^l, dz*2^l (multigrid.py:253).
 * Coefficients are matrix-free: 1/h^2 per face, 0 across the six physical
 * faces of the single cavity domain, ap = -(aw+ae+as+an+ab+at)
 * (multigrid.py:23-67).  tmp (nullable) is a second array of phi's shape,
 * zero-initialised: with it, smoothing runs as fused full sweeps that
 * ping-pong phi -> tmp -> phi (results always end in phi); without it, as
 * two in-place half-sweep launches per sweep. */
typedef struct wf_mg_level {
  void* phi;
  void* rhs;
  int32_t n0, n1, n2;
  int32_t prec;
  double h0, h1, h2;
  void* tmp;
  void* cop;          /* nullable: coarse operator of this level (see below) */
  int32_t cop_sweeps; /* sweeps the operator stands for */
  int32_t reserved;
} wf_mg_level;

/* ------------------------------------------------------------------------
 * Runtime and device arrays.  Replace wp.zeros / wp.empty (solver.py:47,
 * multigrid.py:212-220), wp.copy (solver.py:223-225), array.numpy()
 * (solver.py:119), array.assign() (tests/test_mg.py:45), array.zero_()
 * (multigrid.py:312) and array.fill_() (solver.py:77, :127).
 * ---------------------------------------------------------------------- */
int wf_device_count(int* n);
int wf_set_device(int device);
int wf_get_device_name(char* buf, int len);
int wf_get_device_attr(int* n_cu, int* max_lds_per_block, int* warp_size);
int wf_malloc(void** ptr, size_t bytes); /* zero-filled */
int wf_free(void* ptr);
int wf_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int wf_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int wf_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* Device-to-device copy by a kernel (wp.copy, solver.py:223-225): the same
 * call is safe inside a captured graph. */
int wf_copy(void* dst, const void* src, size_t bytes, void* stream);
int wf_memset(void* ptr, int value, size_t bytes, void* stream);
int wf_fill_f64(double* ptr, double value, size_t n, void* stream);
int wf_fill_f32(float* ptr, float value, size_t n, void* stream);
int wf_synchronize(void);
int wf_stream_create(void** stream);
int wf_stream_destroy(void* stream);
int wf_stream_synchronize(void* stream);
int wf_event_create(void** event);
int wf_event_destroy(void* event);
int wf_event_record(void* event, void* stream);
int wf_event_elapsed_ms(float* ms, void* start, void* stop);
const char* wf_error_string(int code);

/* ------------------------------------------------------------------------
 * Momentum, projection and SGS kernels (warpflow/kernels.py).
 * comp: 0 = u, 1 = v, 2 = w.
 * ---------------------------------------------------------------------- */

/* conv4th_u / conv4th_v / conv4th_w, kernels.py:53-89, :92-129, :132-168:
 * cstar = coo - dt*alpha*N(c), 4th-order staggered convection. */
int wf_conv4th(int comp, const double* u, const double* v, const double* w,
               const double* coo, double* cstar, wf_grid g, int d0, int d1,
               int d2, double dx, double dy, double dz, double dt, double alpha,
               void* stream);

/* diff4th_u / _v / _w, kernels.py:189-225: cstar += dt*alpha*vis_s*L4(c),
 * vis sampled at (i,j,k) for u, (i,j+1,k) for v, (i,j,k+1) for w. */
int wf_diff4th(int comp, const double* c, const double* vis, double* cstar,
               wf_grid g, int d0, int d1, int d2, double dx, double dy,
               double dz, double dt, double alpha, void* stream);

/* The conv4th_c + diff4th_c pair of one RK stage (solver.py:254-265) in one
 * pass: the same two roundings (cstar = coo - dt*alpha*N, then
 * cstar += dt*alpha*vis*L4), reading c, the other two components, coo and
 * vis once.  Equal to the pair because diff4th reads only face ghosts of c,
 * which the intervening fill_bc(c) leaves unchanged. */
int wf_convdiff4th(int comp, const double* u, const double* v, const double* w,
                   const double* coo, const double* vis, double* cstar,
                   wf_grid g, int d0, int d1, int d2, double dx, double dy,
                   double dz, double dt, double alpha, void* stream);

/* wf_convdiff4th for the first RK stage, where the base value c_oo is the
 * component itself (solver.py:223-225 copies u, v, w to uoo, voo, woo just
 * before): the base is taken from the component, and written to coo_save on
 * the operator box -- the cells the second stage reads as c_oo -- so no
 * separate full-field copy is needed.  coo_save's other cells are not
 * written.  No single reference kernel corresponds to it. */
int wf_convdiff4th_save(int comp, const double* u, const double* v, const double* w,
                        double* coo_save, const double* vis, double* cstar, wf_grid g,
                        int d0, int d1, int d2, double dx, double dy, double dz, double dt,
                        double alpha, void* stream);

/* wf_convdiff4th_save of u, v and w in one pass, each over its operator box
 * (solver.py:98-109: (ni-1, nj, nk), (ni, nj-1, nk), (ni, nj, nk-1)): the
 * first RK stage's fill/conv/diff sequence of solver.py:254-265.  The values
 * are those of the three single passes bit for bit, and the reference's when
 * u, v and w have been filled since their last change (as at the first stage
 * of every step: the projection ends with fill_bc of all three): a further
 * fill then changes only deep edge ghosts that no conv4th or diff4th stencil
 * reads, and a third fill changes nothing, so one fill of the three after
 * this call leaves the ghosts as the reference's sequence does.  It reads u,
 * v, w and vis once, where the three single passes read them 3 + 3 + 3 + 3
 * times.  No single reference kernel corresponds to it. */
int wf_convdiff3_save(const double* u, const double* v, const double* w, const double* vis,
                      double* ustar, double* vstar, double* wstar, double* uoo, double* voo,
                      double* woo, wf_grid g, double dx, double dy, double dz, double dt,
                      double alpha, void* stream);

/* The second RK stage's wf_convdiff4th of v and w in one pass, each over its
 * operator box.  In solver.py:254-265 conv_v reads w before w's fill and
 * conv_w after it; this pass runs after the fills of v and w, and v's z-term
 * reads w_pre at the cells outside w's operator box it reaches (rows -1 and
 * nj, columns -2, -1, nk-1, nk of every plane, real-cell coordinates), which
 * wf_save_cross_ghosts copied from w just before w's fill; everywhere else
 * it reads w.  The values are those of the two single passes in the
 * reference's order bit for bit.  No single reference kernel corresponds. */
int wf_convdiff_vw(const double* u, const double* v, const double* w, const double* voo,
                   const double* woo, const double* vis, const double* w_pre, double* vstar,
                   double* wstar, wf_grid g, double dx, double dy, double dz, double dt,
                   double alpha, void* stream);

/* The second RK stage's wf_convdiff4th of u, v and w in one pass, each over
 * its operator box, after the fills of all three (solver.py:254-265 fills and
 * convects one component at a time).  v's z-term reads w_pre as in
 * wf_convdiff_vw.  conv_u reads v and w before their fills; of the cells it
 * reads, those fills change only the x-ghost planes -1 and ni, which u's
 * planes 0 and ni-2 reach: this pass leaves those two planes of ustar
 * unwritten, and wf_convdiff4th_u_edges, called after u's fill and before the
 * fills of v and w, writes them.  The values are those of the single passes
 * in the reference's order bit for bit.  It reads u, v, w, uoo, voo, woo and
 * vis once.  No single reference kernel corresponds. */
int wf_convdiff3(const double* u, const double* v, const double* w, const double* uoo,
                 const double* voo, const double* woo, const double* vis, const double* w_pre,
                 double* ustar, double* vstar, double* wstar, wf_grid g, double dx, double dy,
                 double dz, double dt, double alpha, void* stream);

/* wf_convdiff4th of u on planes 0 and ni-2 of its operator box only (the
 * planes wf_convdiff3 leaves to it).  No single reference kernel corresponds. */
int wf_convdiff4th_u_edges(const double* u, const double* v, const double* w,
                           const double* uoo, const double* vis, double* ustar, wf_grid g,
                           double dx, double dy, double dz, double dt, double alpha,
                           void* stream);

/* w_pre <- w at the cells wf_convdiff_vw and wf_convdiff3 read from w_pre (see
 * wf_convdiff_vw); no other cell of w_pre is written.  No reference
 * counterpart. */
int wf_save_cross_ghosts(const double* w, double* w_pre, wf_grid g, void* stream);

/* calvel_u / _v / _w, kernels.py:232-264: c = cstar - dt*alpha*(p[+1]-p)/h
 * along the component's axis. */
int wf_calvel(int comp, const double* cstar, const double* p, double* c,
              wf_grid g, int d0, int d1, int d2, double h, double dt,
              double alpha, void* stream);

/* FlowSolver._calvel, solver.py:155-166: the three calvel launches over
 * their operator ranges (ni-1,nj,nk), (ni,nj-1,nk), (ni,nj,nk-1)
 * (solver.py:99-110) in one pass. */
int wf_calvel3(const double* ustar, const double* vstar, const double* wstar,
               const double* p, double* u, double* v, double* w, wf_grid g,
               double dx, double dy, double dz, double dt, double alpha,
               void* stream);

/* calmas, kernels.py:271-289: su = div/dt with the 4th-order divergence;
 * *rmax_dev = max(*rmax_dev, max|div*dx*dy*dz|) (caller zeroes it first,
 * solver.py:169). */
int wf_calmas(const double* u, const double* v, const double* w, double* su,
              double* rmax_dev, wf_grid g, int d0, int d1, int d2, double dx,
              double dy, double dz, double dt, void* stream);

/* wf_calvel3 followed by wf_calmas over the real cells, in one pass: the
 * pair _project launches back to back (solver.py:186-197, calvel_{u,v,w}
 * kernels.py:232-264 then calmas kernels.py:271-289; the halo exchange
 * between them is a no-op on one domain).  Writes u, v, w, su and
 * *rmax_dev = max(*rmax_dev, max|div*dx*dy*dz|) bit for bit as the two
 * entries do; calmas no longer re-reads u, v, w from memory.  rhs32
 * (nullable): an FP32 multigrid level array of the real cells ((n+2)^3,
 * one-cell ring) that also receives (float)su on its real cells -- the
 * first defect of wf_mg_solve_mixed (flags & 8).  No single reference
 * kernel corresponds to it. */
int wf_calvel3_calmas(const double* ustar, const double* vstar, const double* wstar,
                      const double* p, double* u, double* v, double* w, double* su,
                      double* rmax_dev, float* rhs32, wf_grid g, double dx, double dy,
                      double dz, double dt, double alpha, void* stream);

/* wf_calvel3_calmas without writing u, v, w: su, rmax and rhs32 are the
 * same, u, v, w are only read (where calvel does not write them, calmas reads
 * them).  For the first outer iteration of the projection loop, which never
 * exits the loop (solver.py:188: it > 1), so the velocities its calvel writes
 * are overwritten by the next iteration's calvel before anything reads them.
 * No single reference kernel corresponds. */
int wf_calmas_of_calvel3(const double* ustar, const double* vstar, const double* wstar,
                         const double* p, const double* u, const double* v, const double* w,
                         double* su, double* rmax_dev, float* rhs32, wf_grid g, double dx,
                         double dy, double dz, double dt, double alpha, void* stream);

/* wale_viscosity, kernels.py:297-367: vis = rrey + nut (WALE, cw), incl. the
 * reference's g2zx/g2zy terms (kernels.py:336, :338).  delta is
 * (dx*dy*dz)^(1/3) (kernels.py:360), computed once by the caller. */
int wf_wale(const double* u, const double* v, const double* w, double* vis,
            wf_grid g, int d0, int d1, int d2, double dx, double dy, double dz,
            double rrey, double cw, double delta, void* stream);

/* wf_checkdt then wf_wale in one pass over the real cells: the reductions of
 * compute_dt (solver.py:115-150) and the WALE viscosity that follows them in
 * a step (solver.py:227-237; nothing in between changes u, v, w or vis).
 * out4 and vis are written bit for bit as the two entries write them (each
 * cell's vis is read before it is overwritten).  No single reference kernel
 * corresponds to it. */
int wf_wale_checkdt(const double* u, const double* v, const double* w, double* vis,
                    double* out4_dev, wf_grid g, double dx, double dy, double dz,
                    double rrey, double cw, double delta, void* stream);

/* maxabs_reduce, kernels.py:374-381: *out = max(*out, max|f|). */
int wf_maxabs(const double* f, double* out_dev, wf_grid g, int d0, int d1,
              int d2, void* stream);

/* dtvisc_reduce, kernels.py:384-402: *out = min(*out, min dt-limit). */
int wf_dtvisc(const double* u, const double* v, const double* w,
              const double* vis, double* out_dev, wf_grid g, int d0, int d1,
              int d2, double dx, double dy, double dz, void* stream);

/* FlowSolver.compute_dt reductions, solver.py:115-134: out4 = {max|u|,
 * max|v|, max|w|, min dtvisc} over the real cells in one pass (the entry
 * initialises out4 to {0, 0, 0, 1e10} itself). */
int wf_checkdt(const double* u, const double* v, const double* w,
               const double* vis, double* out4_dev, wf_grid g, double dx,
               double dy, double dz, void* stream);

/* add_inplace, kernels.py:405-414, launched over ncells+2 (solver.py:198-201):
 * dst += src over the real cells plus one ghost ring. */
int wf_add_ring(double* dst, const double* src, wf_grid g, void* stream);

/* wf_add_ring with the mixed solve's deferred pressure gauge (flags & 4 of
 * wf_mg_solve_mixed leaves the mean of pp in work[2] unsubtracted): dst +=
 * src - *shift_dev on the real cells and the faces of the first ghost layer,
 * dst += src on its edges and corners -- the same bits as subtracting the mean
 * in the solve and then wf_add_ring.  No single reference kernel corresponds
 * to it. */
int wf_add_ring_shift(double* dst, const double* src, const double* shift_dev, wf_grid g,
                      void* stream);

/* fill_const, kernels.py:442-449: f = val over the launch box. */
int wf_fill_interior(double* f, double val, wf_grid g, int d0, int d1, int d2,
                     void* stream);

/* init_noise, kernels.py:452-471: uint32 hash of the full-array indices and
 * seed; u = v = amp*r, w = -amp*r. */
int wf_init_noise(double* u, double* v, double* w, double amp, int32_t seed,
                  wf_grid g, int d0, int d1, int d2, void* stream);

/* ------------------------------------------------------------------------
 * Ghost fills (warpflow/bcs.py).
 * ---------------------------------------------------------------------- */

/* _mk_lin(axis, high), bcs.py:42-106: fld[ghost m] = sign*fld[source m],
 * m < nlay; low side ghost g0-m / source s0+m, high side g0+m / s0-m; the
 * perpendicular axes (ascending) run over [a0, a0+na) x [b0, b0+nb). */
int wf_bc_lin(double* fld, wf_grid g, int axis, int high, int g0, int s0,
              double sign, int a0, int na, int b0, int nb, int nlay,
              void* stream);

/* _mk_const(axis, high), bcs.py:109-163: fld[ghost m] = val, m < nlay. */
int wf_bc_const(double* fld, wf_grid g, int axis, int high, int g0, double val,
                int a0, int na, int b0, int nb, int nlay, void* stream);

/* BoundaryKernels.apply_* (bcs.py:201-311) for the cavity boundary set
 * (W/E/B wall, S/N slip, T moving lid).  which: 0 = apply_u (param = lid
 * velocity), 1 = apply_v, 2 = apply_w, 3 = apply_vis (param = 1/Re),
 * 4 = apply_pbound, 5 = apply_p_mirror (param = number of layers).  The six
 * side launches of the reference run as three launches, one per axis: a
 * thread applies the low-side then the high-side operations of its
 * perpendicular column in the reference order, which is the same sequence
 * per cell (each side operation only touches its own column). */
int wf_bc_fill(double* fld, wf_grid g, int which, double param, void* stream);

/* apply_u(u, lid velocity ulid), apply_v(v), apply_w(w) (bcs.py:201-265) --
 * the three fills solver.py issues back to back (boundu, boundv, boundw) --
 * on three fields of one geometry, as two launches (the x ops, then the y and
 * z ops) that each cover the three fields.  The fields are independent, so
 * the result is that of three wf_bc_fill calls.  u may be NULL: then v's and
 * w's fills only. */
int wf_bc_fill_uvw(double* u, double* v, double* w, wf_grid g, double ulid, void* stream);

/* wf_bc_fill_uvw again on fields filled since their last change (as the RK
 * stages fill u, v, w a second time): the same result, running only the lines
 * that lie in another axis's ghost band -- the only lines whose sources a later
 * op of the first fill can have changed.  No reference counterpart. */
int wf_bc_refill_uvw(double* u, double* v, double* w, wf_grid g, double ulid, void* stream);

/* ------------------------------------------------------------------------
 * Multigrid (warpflow/multigrid.py).  Level arrays follow wf_mg_level.
 * ---------------------------------------------------------------------- */

/* _mgcoeff, multigrid.py:23-67: write the seven coefficient arrays of a level
 * (shape n+2 per axis, real cells only) -- for inspection; the solver itself
 * is matrix-free. */
int wf_mg_coeff(const wf_mg_level* lev, void* aw, void* ae, void* as,
                void* an, void* ab, void* at, void* ap, void* stream);

/* _rbgs, multigrid.py:70-90: one half-sweep on cells with
 * (i+j+k+color) % 2 == 1 (0-based real indices), in place. */
int wf_mg_rbgs(const wf_mg_level* lev, int color, void* stream);

/* Multigrid._refill, multigrid.py:264-287 with bcs._mk_neumann
 * (bcs.py:326-360): one-layer Neumann mirror of the six faces of the ring
 * (faces only; edge and corner ring cells are not written). */
int wf_mg_refill(const wf_mg_level* lev, void* stream);

/* Multigrid._relax, multigrid.py:289-299: nsweeps x (colour 0, colour 1,
 * refill).  Levels whose arrays fit the workgroup's LDS run all sweeps in one
 * single-workgroup launch. */
int wf_mg_relax(const wf_mg_level* lev, int nsweeps, void* stream);

/* _restrict, multigrid.py:93-119, followed by the coarse phi zeroing of
 * multigrid.py:312 (whole coarse array incl. the ring faces). */
int wf_mg_restrict(const wf_mg_level* fine, const wf_mg_level* coarse,
                   void* stream);

/* _prolong, multigrid.py:122-146, followed by the fine-level refill of
 * multigrid.py:321. */
int wf_mg_prolong(const wf_mg_level* coarse, const wf_mg_level* fine,
                  void* stream);

/* _copy_in, multigrid.py:149-154: field su (wf_grid g) -> level rhs. */
int wf_mg_copy_in(const double* src, wf_grid g, const wf_mg_level* lev,
                  void* stream);

/* _copy_out_ring, multigrid.py:157-163: level phi incl. ring -> field at
 * offset npl-1. */
int wf_mg_copy_out_ring(const wf_mg_level* lev, double* dst, wf_grid g,
                        void* stream);

/* Multigrid._vcycle, multigrid.py:301-322, over levels[0..nlev). */
int wf_mg_vcycle(const wf_mg_level* levels, int nlev, int n_pre, int n_post,
                 int n_coarse, void* stream);

/* Multigrid.solve, multigrid.py:324-337: copy-in, zero phi0, n_vcycles
 * V-cycles, copy-out incl. ring.  All levels float64.  The reference also
 * zeroes pp first; here pp's real cells and whole ring are written by the
 * copy-out and its deeper ghost layers are left as they are (keep them zero,
 * as a freshly allocated field is, for the reference's pp).  phi0 is not
 * zeroed when the first sweep can take the zero iterate unread. */
int wf_mg_solve(const wf_mg_level* levels, int nlev, const double* su,
                double* pp, wf_grid g, int n_pre, int n_post, int n_coarse,
                int n_vcycles, void* stream);

/* Coarse operator of a small level: nsweeps RBGS sweeps started from
 * phi = 0 are a fixed linear map phi = M rhs of the level's real cells.
 * wf_mg_coarse_op_build forms M (column-major, ncell x ncell in the level
 * precision; ncell = n0*n1*n2 <= 4096) by running exactly those sweeps on
 * every unit rhs, one workgroup per column.  The V-cycle applies it to the
 * coarsest level, whose phi restriction has just zeroed, when the level's cop
 * is set and cop_sweeps equals the coarsest sweep count -- the same result as
 * the sweeps (multigrid.py:313-316) up to rounding, as one matrix-vector
 * product instead of 2*nsweeps dependent half-sweeps. */
int wf_mg_coarse_op_bytes(const wf_mg_level* lev, size_t* bytes);
int wf_mg_coarse_op_build(const wf_mg_level* lev, int nsweeps, void* op, void* stream);
/* phi = M rhs on the real cells (+ face-ring images), M from the build. */
int wf_mg_coarse_op_apply(const wf_mg_level* lev, void* stream);

/* max|rhs - A phi| over the real cells of one level (FP64 or FP32 level),
 * written to *out_dev -- the residual norm of tests/test_mg.py:16-32. */
int wf_mg_residual_max(const wf_mg_level* lev, double* out_dev, void* stream);

/* ------------------------------------------------------------------------
 * Mixed-precision pressure solve.  No single counterpart in the reference:
 * it stands in for Multigrid.solve (multigrid.py:324-337) when the
 * V-cycle runs in FP32 on the error equation inside FP64 defect correction.
 * ---------------------------------------------------------------------- */

/* Bytes of device workspace wf_mg_solve_mixed / wf_field_mean need for a
 * field of grid g (a few scalars plus one partial sum per workgroup). */
int wf_mg_mixed_workspace_bytes(wf_grid g, size_t* bytes);

/* pp = (approximate) A^-1 su.  levels: an all-FP32 hierarchy whose level 0
 * has the field's real-cell counts.  Each cycle: FP64 defect r = su - A pp
 * (matrix-free on the field), optional FP64 mean removal (flags & 1), FP32
 * V-cycle from zero on r, pp += correction in FP64.  rtol <= 0: exactly
 * max_cycles cycles, no host synchronisation.  rtol > 0: stop once
 * max|r| <= rtol*max|su| (one host read per cycle), at most max_cycles.
 * flags & 2: subtract the FP64 mean of pp at the end (pressure gauge); with
 * flags & 4 as well, only form that mean in work[2] and leave pp unshifted
 * (for wf_add_ring_shift).  flags & 8: levels[0].rhs already holds (float)su
 * on the real cells (wf_calvel3_calmas' rhs32), so the first defect is not
 * formed again -- honoured with a fixed cycle count and no defect mean.  The
 * first ghost layer of pp gets the Neumann face mirror on its faces and zero
 * on its edges and corners; the deeper ghost layers are not written (the
 * reference's solve leaves them zero: keep them zero, as a freshly allocated
 * field is).  *cycles_out / *res_out (nullable) report the V-cycles run
 * and, in tolerance mode, the last relative residual.  work: device buffer
 * of wf_mg_mixed_workspace_bytes. */
int wf_mg_solve_mixed(const wf_mg_level* levels, int nlev, const double* su,
                      double* pp, wf_grid g, int n_pre, int n_post,
                      int n_coarse, int max_cycles, double rtol, int flags,
                      double* work, size_t work_bytes, int* cycles_out,
                      double* res_out, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): wf_mg_solve_mixed with its workspace taken from a static
 * device array of WF_STATIC_SCRATCH_BYTES (1 MiB) that the library holds on
 * each GPU it runs on until the process ends, instead of a `work` argument.
 * Same arguments, results and bits as wf_mg_solve_mixed otherwise; it
 * allocates nothing.  Limits: the grid's workspace
 * (wf_mg_mixed_workspace_bytes) must fit WF_STATIC_SCRATCH_BYTES, else
 * WF_ESCRATCH is returned before anything is launched (up to about 2^17
 * partial sums: 320^3 fits, 256^3 needs 512 KiB); flags & 4 (deferred gauge
 * mean) is rejected with WF_EINVAL, since the mean would stay in the
 * library's scratch.  All calls on one GPU share the one array: calls that
 * overlap on a GPU (two streams or threads) race. */
int wf_mg_solve_mixed_static(const wf_mg_level* levels, int nlev, const double* su,
                             double* pp, wf_grid g, int n_pre, int n_post,
                             int n_coarse, int max_cycles, double rtol, int flags,
                             int* cycles_out, double* res_out, void* stream);

/* *out_dev = mean of f over the real cells (deterministic FP64 sum). */
int wf_field_mean(const double* f, wf_grid g, double* work, size_t work_bytes,
                  double* out_dev, void* stream);

/* ------------------------------------------------------------------------
 * Ensembles: many independent cavities of different sizes advanced
 * together, one launch per operator for all members.  No counterpart in the
 * reference, which runs one cavity per process (solver.py:29-271); each
 * batched entry applies the single-domain entry named beside it to every
 * member, with the same arithmetic.
 *
 * Layout.  wf_batch_create packs the members' arrays into pools:
 * every field (u, v, w, p, ...) is one pool of field_elems doubles with
 * member m's (n0+6)(n1+6)(n2+6) array at field_off[m]; multigrid level l is
 * one FP32 pool per level array (phi, rhs, tmp) with member m's
 * (n0/2^l+2)(n1/2^l+2)(n2/2^l+2) array at level_off[l*nmem+m]; the coarsest
 * level's operators (wf_mg_coarse_op_build, column-major) share one FP32
 * pool at cop_off[m] (-1 when the coarsest level has more than 4096 cells).
 * wf_batch_layout reports the offsets and sizes; the caller allocates pools.
 *
 * Masks.  `active` (nullable device int[nmem]) limits a call to the members
 * still iterating the projection loop; wf_batch_project_mask updates it on
 * the device, so a whole step can be captured into a graph.
 * ---------------------------------------------------------------------- */
typedef struct {
  int32_t n0, n1, n2;  /* real cells, each divisible by 2^(nlev-1) */
  uint32_t seed;       /* init_noise seed */
  double dx, dy, dz;   /* cell sizes */
  double rey;          /* 1/Re */
  double ulid;         /* lid velocity */
  double ls;           /* WALE length scale cw * delta (as wf_wale's caller forms it) */
  double noise;        /* initial noise amplitude */
  double dt;           /* fixed time step */
} wf_member;

typedef struct wf_batch_s* wf_batch;

int wf_batch_create(const wf_member* members, int nmem, int nlev, wf_batch* out);
int wf_batch_destroy(wf_batch b);
int wf_batch_layout(wf_batch b, int64_t* field_off, int64_t* field_elems, int64_t* level_off,
                    int64_t* level_elems, int64_t* cop_off, int64_t* cop_elems);
/* wf_init_noise over the real cells of every member. */
int wf_batch_init_noise(wf_batch b, double* u, double* v, double* w, void* stream);
/* wf_fill_interior(vis, 1/Re) of every member (WALE off, solver.py:236). */
int wf_batch_fill_vis(wf_batch b, double* vis, void* stream);
/* wf_wale. */
int wf_batch_wale(wf_batch b, const double* u, const double* v, const double* w, double* vis,
                  void* stream);
/* wf_checkdt: out4[4m .. 4m+3] = umax, vmax, wmax, dt limit of member m. */
int wf_batch_checkdt(wf_batch b, const double* u, const double* v, const double* w,
                     const double* vis, double* out4, void* stream);
/* wf_bc_fill(which) with param = lid velocity (u), 1/Re (vis), 3 (p mirror). */
int wf_batch_bc_fill(wf_batch b, double* f, int which, const int* active, void* stream);
/* wf_bc_fill_uvw: the fills of u, v and w (u may be NULL: v and w only), each
 * axis launch covering the fields. */
int wf_batch_bc_fill_uvw(wf_batch b, double* u, double* v, double* w, const int* active,
                         void* stream);
/* wf_convdiff4th over the operator box of component comp. */
int wf_batch_convdiff(wf_batch b, int comp, const double* u, const double* v, const double* w,
                      const double* coo, const double* vis, double* cstar, double alpha,
                      void* stream);
/* wf_convdiff4th_save: the first stage's form, base value from the component,
 * stored into coo_save on the operator box. */
int wf_batch_convdiff_save(wf_batch b, int comp, const double* u, const double* v,
                           const double* w, double* coo_save, const double* vis, double* cstar,
                           double alpha, void* stream);
/* wf_convdiff3_save: the first RK stage's three pairs in one pass (valid on
 * fields filled since their last change). */
int wf_batch_convdiff3_save(wf_batch b, const double* u, const double* v, const double* w,
                            const double* vis, double* ustar, double* vstar, double* wstar,
                            double* uoo, double* voo, double* woo, double alpha, void* stream);
/* wf_convdiff3: the second RK stage's three pairs in one pass after the three
 * fills (u's planes 0 and ni-2 from wf_batch_convdiff4th_u_edges). */
int wf_batch_convdiff3(wf_batch b, const double* u, const double* v, const double* w,
                       const double* uoo, const double* voo, const double* woo,
                       const double* vis, const double* w_pre, double* ustar, double* vstar,
                       double* wstar, double alpha, void* stream);
/* wf_convdiff4th_u_edges. */
int wf_batch_convdiff4th_u_edges(wf_batch b, const double* u, const double* v, const double* w,
                                 const double* uoo, const double* vis, double* ustar,
                                 double alpha, void* stream);
/* wf_save_cross_ghosts. */
int wf_batch_save_cross_ghosts(wf_batch b, const double* w, double* w_pre, void* stream);
/* wf_calvel3. */
int wf_batch_calvel3(wf_batch b, const double* us, const double* vs, const double* ws,
                     const double* p, double* u, double* v, double* w, double alpha,
                     const int* active, void* stream);
/* wf_calmas: rmax[m] of member m. */
int wf_batch_calmas(wf_batch b, const double* u, const double* v, const double* w, double* su,
                    double* rmax, const int* active, void* stream);
/* wf_calvel3_calmas: calvel3 then calmas (rmax[m] of member m) in one pass. */
int wf_batch_calvel3_calmas(wf_batch b, const double* us, const double* vs, const double* ws,
                            const double* p, double* u, double* v, double* w, double* su,
                            double* rmax, double alpha, const int* active, void* stream);
/* wf_add_ring. */
int wf_batch_add_ring(wf_batch b, double* p, const double* pp, const int* active,
                      void* stream);
/* wf_mg_solve_mixed with a fixed ncycles (no tolerance test, no host sync):
 * phi/rhs/tmp = FP32 level pools (nlev each), cop = operator pool or NULL
 * (NULL: n_coarse sweeps on the coarsest level), flags 2 = remove the mean. */
int wf_batch_mg_solve_mixed(wf_batch b, float* const* phi, float* const* rhs, float* const* tmp,
                            const float* cop, const double* su, double* pp, int ncycles,
                            int n_pre, int n_post, int n_coarse, int flags, const int* active,
                            void* stream);
/* Outer-loop bookkeeping of solver.py:185-197 for iteration it: each active
 * member records iters[m] = it, then leaves when rmax[m] < eps and it > 1;
 * rmax[m] > 10 sets status[m] = 1 (diverged) and leaves. */
int wf_batch_project_mask(wf_batch b, const double* rmax, double eps, int it, int* active,
                          int* iters, int* status, void* stream);
/* p[0 .. n) = value on the device (graph-safe int fill). */
int wf_batch_set_int(int* p, int n, int value, void* stream);

/* ------------------------------------------------------------------------
 * Launch-shape tuning (no counterpart in the reference, whose Warp launches
 * have no shape parameters).  A setting applies to launches of exactly the
 * shape it names -- the real cells (n0, n1, n2) of the field (convdiff) or of
 * the multigrid level (sweeps) -- so tuning one grid never changes another.
 * value 0 removes the entry (the built-in choice); results never depend on
 * any setting.
 *   WF_TUNE_CD_ICH              i-planes per workgroup of wf_convdiff4th
 *   WF_TUNE_SWEEP_TILE_F64/F32  fused RBGS sweep tile (j x k): 1 = 8 x 64,
 *                               2 = 16 x 32, 3 = 16 x 64, 4 = 32 x 32
 *   WF_TUNE_SWEEP_ICH_F64/F32   i-planes per workgroup of the fused sweep;
 *                               -2 / -4: chunks of 2 / 4 planes, each in one
 *                               pass without the plane pipeline (small levels)
 *   WF_TUNE_SWEEP2_TILE_F64/F32 tile of the two-sweep kernel (pairs of sweeps
 *                               in a V-cycle): 1 = 8 x 64, 2 = 16 x 32; FP32
 *                               also 3 = 16 x 64, 4 = 32 x 32
 *   WF_TUNE_SWEEP2_ICH_F64/F32  i-planes per workgroup of the two-sweep kernel
 *   WF_TUNE_SWEEP_PAIRS_F64/F32 pairs of sweeps of this level as 1 = two
 *                               one-sweep launches, 2 = one two-sweep launch
 *                               (built-in: 2 for FP64 levels of >= 2^23
 *                               cells, else 1)
 *   WF_TUNE_PROJ_ICH            i-planes per workgroup of wf_calvel3_calmas
 *                               (raised where needed to keep the workgroup
 *                               count within the reduction's partial slots)
 *   WF_TUNE_CD3_ICH             i-planes per workgroup of wf_convdiff3_save
 *   WF_TUNE_CDVW_ICH            i-planes per workgroup of wf_convdiff_vw
 *   WF_TUNE_CD3S2_ICH           i-planes per workgroup of wf_convdiff3
 * At most 128 entries.
 * ---------------------------------------------------------------------- */
enum {
  WF_TUNE_CD_ICH = 0,
  WF_TUNE_SWEEP_TILE_F64 = 1,
  WF_TUNE_SWEEP_TILE_F32 = 2,
  WF_TUNE_SWEEP_ICH_F64 = 3,
  WF_TUNE_SWEEP_ICH_F32 = 4,
  WF_TUNE_SWEEP2_TILE_F64 = 5,
  WF_TUNE_SWEEP2_TILE_F32 = 6,
  WF_TUNE_SWEEP2_ICH_F64 = 7,
  WF_TUNE_SWEEP2_ICH_F32 = 8,
  WF_TUNE_SWEEP_PAIRS_F64 = 9,
  WF_TUNE_PROJ_ICH = 10,
  WF_TUNE_SWEEP_PAIRS_F32 = 11,
  WF_TUNE_CD3_ICH = 12,
  WF_TUNE_CDVW_ICH = 13,
  WF_TUNE_CD3S2_ICH = 14,
  WF_TUNE_COUNT = 15
};
int wf_tune_set(int knob, int n0, int n1, int n2, int value);
int wf_tune_get(int knob, int n0, int n1, int n2, int* value);
int wf_tune_reset(void);

/* ------------------------------------------------------------------------
 * HIP graphs: capture the work enqueued on a (non-null) stream between begin
 * and end, then replay it.  Replaces the reference's per-launch Python
 * dispatch of a whole step.
 * ---------------------------------------------------------------------- */
int wf_graph_begin(void* stream);
int wf_graph_end(void* stream, void** exec);
int wf_graph_launch(void* exec, void* stream);
int wf_graph_destroy(void* exec);

#ifdef __cplusplus
}
#endif

#endif /* WARPFLOW_HIP_H */
