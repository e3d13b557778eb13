/* wf_oracle.c -- plain C (+ OpenMP) restatement of the warpflow-les per-step
 * hot path.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): used by the
 * tests as a fast second transcription of the reference schedule and timed
 * as the CPU baseline by bench.py.  Nothing in the product links it.
 *
 * It follows oracle/np_oracle.py function for function (which in turn cites
 * the reference file:line of every piece):
 *   kernels.py  conv4th_* :53-168, diff4th_* :189-225, calvel_* :232-264,
 *               calmas :271-289, wale_viscosity :297-367, maxabs_reduce :374,
 *               dtvisc_reduce :384, add_inplace :405, fill_const :442,
 *               init_noise :452
 *   bcs.py      apply_u/v/w/vis/pbound/p_mirror :201-311 (one axis at a time,
 *               low side then high side per perpendicular line -- the same
 *               per-cell order as the six side launches)
 *   multigrid.py _rbgs :70, _restrict :93, _prolong :122, _relax :289,
 *               _vcycle :301, solve :324 -- matrix-free coefficients
 *   solver.py   compute_dt :122, _calvel :155, _calmas :168, _project :178,
 *               step :217
 * All per-cell work is independent within a sweep (RBGS colours read only the
 * other colour), and the reductions are max/min, so results do not depend on
 * the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NPL 3

typedef struct {
  int n0, n1, n2;      /* real cells */
  long long s0, s1;    /* strides of i, j (k unit) for a field with NPL ghosts */
  long long size;
} FGrid;

typedef struct {
  int n0, n1, n2;
  long long s0, s1, size;
  double h0, h1, h2;
  double *phi, *rhs;
} Level;

typedef struct {
  FGrid g;
  double dx, dy, dz, rey, ulid, cw;
  int sgs, variable_dt, max_outer, nlev, npre, npost, ncoarse, nvcycles, seed;
  double eps, dt, dt_growth, safety, noise;
  double ctime, rmax, cfl;
  int itime, outer_iters;
  double *f[13];
  Level lev[16];
  double rmax_hist[64];
  int nhist;
} Solver;

enum { U, V, W, P, US, VS, WS, UOO, VOO, WOO, VIS, SU, PP };

#define IDX(g, i, j, k) ((long long)(i) * (g).s0 + (long long)(j) * (g).s1 + (k))

/* ------------------------------------------------------------ kernels.py */
static inline double I4(const double *f, long long c, long long s) {
  return (-f[c + s] + 9.0 * f[c] + 9.0 * f[c - s] - f[c - 2 * s]) / 16.0;
}
static inline double D4(double sp, double sn, double sm, double smm, double h) {
  return (-sp + 27.0 * sn - 27.0 * sm + smm) / (24.0 * h);
}
static inline double conv_self(const double *c, long long x, long long sa, double ha) {
  double p2 = I4(c, x + 2 * sa, sa), p1 = I4(c, x + sa, sa), p0 = I4(c, x, sa),
         m1 = I4(c, x - sa, sa);
  return D4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, ha);
}
static inline double conv_cross(const double *c, const double *q, long long x, long long sa,
                                long long sb, double hb) {
  double sp1 = I4(c, x + 2 * sb, sb) * I4(q, x + sa + sb, sa);
  double sn = I4(c, x + sb, sb) * I4(q, x + sa, sa);
  double sm1 = I4(c, x, sb) * I4(q, x + sa - sb, sa);
  double sm2 = I4(c, x - sb, sb) * I4(q, x + sa - 2 * sb, sa);
  return D4(sp1, sn, sm1, sm2, hb);
}

static void conv4th(Solver *S, int comp, double dt, double alpha) {
  const FGrid g = S->g;
  const double *u = S->f[U], *v = S->f[V], *w = S->f[W], *coo = S->f[UOO + comp];
  double *cs = S->f[US + comp];
  const int d0 = g.n0 - (comp == 0), d1 = g.n1 - (comp == 1), d2 = g.n2 - (comp == 2);
  const long long sx = g.s0, sy = g.s1, sz = 1;
  const double dta = dt * alpha, dx = S->dx, dy = S->dy, dz = S->dz;
#pragma omp parallel for schedule(static)
  for (int i = NPL; i < NPL + d0; ++i)
    for (int j = NPL; j < NPL + d1; ++j)
      for (int k = NPL; k < NPL + d2; ++k) {
        const long long x = IDX(g, i, j, k);
        double n;
        if (comp == 0)
          n = conv_self(u, x, sx, dx) + conv_cross(u, v, x, sx, sy, dy) +
              conv_cross(u, w, x, sx, sz, dz);
        else if (comp == 1)
          n = conv_cross(v, u, x, sy, sx, dx) + conv_self(v, x, sy, dy) +
              conv_cross(v, w, x, sy, sz, dz);
        else
          n = conv_cross(w, u, x, sz, sx, dx) + conv_cross(w, v, x, sz, sy, dy) +
              conv_self(w, x, sz, dz);
        cs[x] = coo[x] - dta * n;
      }
}

static void diff4th(Solver *S, int comp, double dt, double alpha) {
  const FGrid g = S->g;
  const double *c = S->f[U + comp], *vis = S->f[VIS];
  double *cs = S->f[US + comp];
  const int d0 = g.n0 - (comp == 0), d1 = g.n1 - (comp == 1), d2 = g.n2 - (comp == 2);
  const long long voff = comp == 0 ? 0 : (comp == 1 ? g.s1 : 1);
  const double dta = dt * alpha, rx = 12.0 * S->dx * S->dx, ry = 12.0 * S->dy * S->dy,
               rz = 12.0 * S->dz * S->dz;
  const long long sx = g.s0, sy = g.s1;
#pragma omp parallel for schedule(static)
  for (int i = NPL; i < NPL + d0; ++i)
    for (int j = NPL; j < NPL + d1; ++j)
      for (int k = NPL; k < NPL + d2; ++k) {
        const long long x = IDX(g, i, j, k);
        const double f0 = c[x];
        double lx = (-c[x + 2 * sx] + 16.0 * c[x + sx] - 30.0 * f0 + 16.0 * c[x - sx] -
                     c[x - 2 * sx]) / rx;
        double ly = (-c[x + 2 * sy] + 16.0 * c[x + sy] - 30.0 * f0 + 16.0 * c[x - sy] -
                     c[x - 2 * sy]) / ry;
        double lz = (-c[x + 2] + 16.0 * c[x + 1] - 30.0 * f0 + 16.0 * c[x - 1] - c[x - 2]) / rz;
        double diff = vis[x + voff] * (lx + ly + lz);
        cs[x] = cs[x] + dta * diff;
      }
}

static void calvel(Solver *S, double alpha) {
  const FGrid g = S->g;
  const double dta = S->dt * alpha;
  const double *p = S->f[P];
#pragma omp parallel for schedule(static)
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) {
        const long long x = IDX(g, i, j, k);
        if (i < NPL + g.n0 - 1) S->f[U][x] = S->f[US][x] - dta * (p[x + g.s0] - p[x]) / S->dx;
        if (j < NPL + g.n1 - 1) S->f[V][x] = S->f[VS][x] - dta * (p[x + g.s1] - p[x]) / S->dy;
        if (k < NPL + g.n2 - 1) S->f[W][x] = S->f[WS][x] - dta * (p[x + 1] - p[x]) / S->dz;
      }
}

static double calmas(Solver *S) {
  const FGrid g = S->g;
  const double *u = S->f[U], *v = S->f[V], *w = S->f[W];
  double *su = S->f[SU];
  double rmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : rmax)
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) {
        const long long x = IDX(g, i, j, k);
        double dudx = D4(u[x + g.s0], u[x], u[x - g.s0], u[x - 2 * g.s0], S->dx);
        double dvdy = D4(v[x + g.s1], v[x], v[x - g.s1], v[x - 2 * g.s1], S->dy);
        double dwdz = D4(w[x + 1], w[x], w[x - 1], w[x - 2], S->dz);
        double fact = dudx + dvdy + dwdz;
        su[x] = fact / S->dt;
        double a = fabs(fact * S->dx * S->dy * S->dz);
        if (a > rmax) rmax = a;
      }
  return rmax;
}

static void wale(Solver *S) {
  const FGrid g = S->g;
  const double *u = S->f[U], *v = S->f[V], *w = S->f[W];
  double *vis = S->f[VIS];
  const double dx = S->dx, dy = S->dy, dz = S->dz;
  const double delta = pow(dx * dy * dz, 1.0 / 3.0), ls = S->cw * delta;
  const long long X = g.s0, Y = g.s1, Z = 1;
#pragma omp parallel for schedule(static)
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) {
        const long long x = IDX(g, i, j, k);
        double dudx = (u[x] - u[x - X]) / dx;
        double dudy = (0.25 * (u[x + Y] + u[x - X + Y] - u[x - Y] - u[x - X - Y])) / dy;
        double dudz = (0.25 * (u[x + Z] + u[x - X + Z] - u[x - Z] - u[x - X - Z])) / dz;
        double dvdx = (0.25 * (v[x + X] + v[x + X - Y] - v[x - X] - v[x - X - Y])) / dx;
        double dvdy = (v[x] - v[x - Y]) / dy;
        double dvdz = (0.25 * (v[x + Z] + v[x - Y + Z] - v[x - Z] - v[x - Y - Z])) / dz;
        double dwdx = (0.25 * (w[x + X] + w[x + X - Z] - w[x - X] - w[x - X - Z])) / dx;
        double dwdy = (0.25 * (w[x + Y] + w[x + Y - Z] - w[x - Y] - w[x - Y - Z])) / dy;
        double dwdz = (w[x] - w[x - Z]) / dz;
        double g2xx = dudx * dudx + dudy * dvdx + dudz * dwdx;
        double g2yy = dvdx * dudy + dvdy * dvdy + dvdz * dwdy;
        double g2zz = dwdx * dudz + dwdy * dvdz + dwdz * dwdz;
        double g2xy = dudx * dudy + dudy * dvdy + dudz * dwdy;
        double g2yx = dvdx * dudx + dvdy * dvdx + dvdz * dwdx;
        double g2xz = dudx * dudz + dudy * dvdz + dudz * dwdz;
        double g2zx = dwdx * dudx + dwdy * dvdx + dwdz * dudz;
        double g2yz = dvdx * dudz + dvdy * dvdz + dvdz * dwdz;
        double g2zy = dwdx * dudy + dwdy * dvdy + dwdz * dudy;
        double tr = (g2xx + g2yy + g2zz) / 3.0;
        double sdxx = g2xx - tr, sdyy = g2yy - tr, sdzz = g2zz - tr;
        double sdxy = 0.5 * (g2xy + g2yx), sdxz = 0.5 * (g2xz + g2zx), sdyz = 0.5 * (g2yz + g2zy);
        double sdsd = sdxx * sdxx + sdyy * sdyy + sdzz * sdzz +
                      2.0 * (sdxy * sdxy + sdxz * sdxz + sdyz * sdyz);
        double sxy = 0.5 * (dudy + dvdx), sxz = 0.5 * (dudz + dwdx), syz = 0.5 * (dvdz + dwdy);
        double ss = dudx * dudx + dvdy * dvdy + dwdz * dwdz +
                    2.0 * (sxy * sxy + sxz * sxz + syz * syz);
        double nut = 0.0;
        if (sdsd > 1.0e-8)
          nut = ls * ls * pow(sdsd, 1.5) / (pow(fabs(ss), 2.5) + pow(sdsd, 1.25));
        vis[x] = S->rey + nut;
      }
}

static void fill_interior(Solver *S, double *f, double val) {
  const FGrid g = S->g;
#pragma omp parallel for schedule(static)
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) f[IDX(g, i, j, k)] = val;
}

static void init_noise(Solver *S, double amp, uint32_t seed) {
  const FGrid g = S->g;
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) {
        uint32_t h = ((uint32_t)i * 73856093u) ^ ((uint32_t)j * 19349663u) ^
                     ((uint32_t)k * 83492791u) ^ seed;
        h ^= h >> 13;
        h *= 1274126177u;
        h ^= h >> 16;
        double r = ((double)h / 4294967295.0) * 2.0 - 1.0;
        long long x = IDX(g, i, j, k);
        S->f[U][x] = amp * r;
        S->f[V][x] = amp * r;
        S->f[W][x] = -amp * r;
      }
}

/* ---------------------------------------------------------------- bcs.py */
typedef struct { int lin, high, g0, s0, nlay; double val; } BcOp;

static void bc_axis(const FGrid g, double *f, int axis, int a0, int na, int b0, int nb,
                    const BcOp *ops, int nops) {
  for (int aa = 0; aa < na; ++aa)
    for (int bb = 0; bb < nb; ++bb) {
      int a = a0 + aa, b = b0 + bb;
      long long base, s;
      if (axis == 0) { base = IDX(g, 0, a, b); s = g.s0; }
      else if (axis == 1) { base = IDX(g, a, 0, b); s = g.s1; }
      else { base = IDX(g, a, b, 0); s = 1; }
      for (int q = 0; q < nops; ++q)
        for (int m = 0; m < ops[q].nlay; ++m) {
          int gi = ops[q].high ? ops[q].g0 + m : ops[q].g0 - m;
          if (ops[q].lin) {
            int si = ops[q].high ? ops[q].s0 - m : ops[q].s0 + m;
            f[base + gi * s] = ops[q].val * f[base + si * s];
          } else {
            f[base + gi * s] = ops[q].val;
          }
        }
    }
}

#define CO(h, g0, v, nl) {0, h, g0, 0, nl, v}
#define LI(h, g0, s0, sg, nl) {1, h, g0, s0, nl, sg}

static void apply_fill(const FGrid g, double *f, int which, double param) {
  const int N = NPL, ni = g.n0, nj = g.n1, nk = g.n2;
  if (which == 0) { /* apply_u, bcs.py:201-223 */
    BcOp x[2] = {CO(0, N - 1, 0.0, 3), CO(1, N + ni - 1, 0.0, 3)};
    BcOp y[2] = {LI(0, N - 1, N, 1.0, 3), LI(1, N + nj, N + nj - 1, 1.0, 3)};
    BcOp z[2] = {LI(0, N - 1, N, -1.0, 3), CO(1, N + nk, param, 3)};
    bc_axis(g, f, 0, N - 1, nj + 2, N - 1, nk + 2, x, 2);
    bc_axis(g, f, 1, N - 1, ni + 1, N - 1, nk + 2, y, 2);
    bc_axis(g, f, 2, N - 1, ni + 1, N - 1, nj + 2, z, 2);
  } else if (which == 1) { /* apply_v, bcs.py:225-244 */
    BcOp x[2] = {LI(0, N - 1, N, -1.0, 3), LI(1, N + ni, N + ni - 1, -1.0, 3)};
    BcOp y[2] = {CO(0, N - 1, 0.0, 3), CO(1, N + nj - 1, 0.0, 3)};
    BcOp z[2] = {LI(0, N - 1, N, -1.0, 3), LI(1, N + nk, N + nk - 1, 1.0, 3)};
    bc_axis(g, f, 0, N - 1, nj + 1, N - 1, nk + 2, x, 2);
    bc_axis(g, f, 1, N - 1, ni + 2, N - 1, nk + 2, y, 2);
    bc_axis(g, f, 2, N - 1, ni + 2, N - 1, nj + 1, z, 2);
  } else if (which == 2) { /* apply_w, bcs.py:246-265 */
    BcOp x[2] = {LI(0, N - 1, N, -1.0, 3), LI(1, N + ni, N + ni - 1, -1.0, 3)};
    BcOp y[2] = {LI(0, N - 1, N, 1.0, 3), LI(1, N + nj, N + nj - 1, 1.0, 3)};
    BcOp z[2] = {CO(0, N - 1, 0.0, 3), CO(1, N + nk - 1, 0.0, 3)};
    bc_axis(g, f, 0, N - 1, nj + 2, N - 1, nk + 1, x, 2);
    bc_axis(g, f, 1, N - 1, ni + 2, N - 1, nk + 1, y, 2);
    bc_axis(g, f, 2, N - 1, ni + 2, N - 1, nj + 2, z, 2);
  } else if (which == 3) { /* apply_vis, bcs.py:267-293 */
    BcOp x[4] = {CO(0, N, param, 1), LI(0, N - 1, N, 1.0, 1), CO(1, N + ni - 1, param, 1),
                 LI(1, N + ni, N + ni - 1, 1.0, 1)};
    BcOp y[2] = {LI(0, N - 1, N, 1.0, 1), LI(1, N + nj, N + nj - 1, 1.0, 1)};
    BcOp z[3] = {CO(0, N, param, 1), LI(0, N - 1, N, 1.0, 1), LI(1, N + nk, N + nk - 1, 1.0, 1)};
    bc_axis(g, f, 0, N - 1, nj + 2, N - 1, nk + 2, x, 4);
    bc_axis(g, f, 1, N - 1, ni + 2, N - 1, nk + 2, y, 2);
    bc_axis(g, f, 2, N - 1, ni + 2, N - 1, nj + 2, z, 3);
  } else { /* 4: apply_pbound (bcs.py:295-302); 5: apply_p_mirror (:304-311) */
    const int n3[3] = {ni, nj, nk};
    const int nl = which == 4 ? 1 : (int)param;
    for (int ax = 0; ax < 3; ++ax) {
      int pa = ax == 0 ? nj : ni, pb = ax == 2 ? nj : nk;
      BcOp o[2] = {LI(0, N - 1, N, 1.0, nl), LI(1, N + n3[ax], N + n3[ax] - 1, 1.0, nl)};
      if (which == 4)
        bc_axis(g, f, ax, N - 1, pa + 3, N - 1, pb + 3, o, 2);
      else
        bc_axis(g, f, ax, 0, pa + 2 * N, 0, pb + 2 * N, o, 2);
    }
  }
}

/* ---------------------------------------------------------- multigrid.py */
#define LIDX(L, i, j, k) ((long long)((i) + 1) * (L)->s0 + (long long)((j) + 1) * (L)->s1 + ((k) + 1))

typedef struct { double aw, ae, as, an, ab, at, ap; } Coef;

static inline Coef coef(const Level *L, int i, int j, int k) {
  const double cx = 1.0 / (L->h0 * L->h0), cy = 1.0 / (L->h1 * L->h1), cz = 1.0 / (L->h2 * L->h2);
  Coef c;
  c.aw = i == 0 ? 0.0 : cx;
  c.ae = i == L->n0 - 1 ? 0.0 : cx;
  c.as = j == 0 ? 0.0 : cy;
  c.an = j == L->n1 - 1 ? 0.0 : cy;
  c.ab = k == 0 ? 0.0 : cz;
  c.at = k == L->n2 - 1 ? 0.0 : cz;
  c.ap = -(c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

static inline double nbsum(const Level *L, const double *phi, long long x, const Coef *c) {
  return c->aw * phi[x - L->s0] + c->ae * phi[x + L->s0] + c->as * phi[x - L->s1] +
         c->an * phi[x + L->s1] + c->ab * phi[x - 1] + c->at * phi[x + 1];
}

static void rbgs(Level *L, int color) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < L->n0; ++i)
    for (int j = 0; j < L->n1; ++j)
      for (int k = (i + j + color + 1) & 1; k < L->n2; k += 2) {
        long long x = LIDX(L, i, j, k);
        Coef c = coef(L, i, j, k);
        L->phi[x] = (L->rhs[x] - nbsum(L, L->phi, x, &c)) / c.ap;
      }
}

static void refill(Level *L) { /* multigrid.py:264-287, faces only */
  for (int j = 0; j < L->n1; ++j)
    for (int k = 0; k < L->n2; ++k) {
      L->phi[LIDX(L, -1, j, k)] = L->phi[LIDX(L, 0, j, k)];
      L->phi[LIDX(L, L->n0, j, k)] = L->phi[LIDX(L, L->n0 - 1, j, k)];
    }
  for (int i = 0; i < L->n0; ++i)
    for (int k = 0; k < L->n2; ++k) {
      L->phi[LIDX(L, i, -1, k)] = L->phi[LIDX(L, i, 0, k)];
      L->phi[LIDX(L, i, L->n1, k)] = L->phi[LIDX(L, i, L->n1 - 1, k)];
    }
  for (int i = 0; i < L->n0; ++i)
    for (int j = 0; j < L->n1; ++j) {
      L->phi[LIDX(L, i, j, -1)] = L->phi[LIDX(L, i, j, 0)];
      L->phi[LIDX(L, i, j, L->n2)] = L->phi[LIDX(L, i, j, L->n2 - 1)];
    }
}

static void relax(Level *L, int n) {
  for (int s = 0; s < n; ++s) {
    rbgs(L, 0);
    rbgs(L, 1);
    refill(L);
  }
}

static void restrict_(Level *F, Level *C) {
#pragma omp parallel for schedule(static)
  for (int ci = 0; ci < C->n0; ++ci)
    for (int cj = 0; cj < C->n1; ++cj)
      for (int ck = 0; ck < C->n2; ++ck) {
        double acc = 0.0;
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
              int i = 2 * ci + di, j = 2 * cj + dj, k = 2 * ck + dk;
              long long x = LIDX(F, i, j, k);
              Coef c = coef(F, i, j, k);
              acc = acc + (F->rhs[x] - (nbsum(F, F->phi, x, &c) + c.ap * F->phi[x]));
            }
        C->rhs[LIDX(C, ci, cj, ck)] = acc * 0.125;
      }
  memset(C->phi, 0, sizeof(double) * C->size);
}

static void prolong(Level *C, Level *F) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n0; ++i)
    for (int j = 0; j < F->n1; ++j)
      for (int k = 0; k < F->n2; ++k) {
        int c0x = (i % 2 == 0) ? i / 2 - 1 : i / 2, c0y = (j % 2 == 0) ? j / 2 - 1 : j / 2,
            c0z = (k % 2 == 0) ? k / 2 - 1 : k / 2;
        double wx0 = (i % 2 == 0) ? 0.25 : 0.75, wy0 = (j % 2 == 0) ? 0.25 : 0.75,
               wz0 = (k % 2 == 0) ? 0.25 : 0.75;
        double acc = 0.0;
        for (int di = 0; di < 2; ++di) {
          double wxi = di == 0 ? wx0 : 1.0 - wx0;
          for (int dj = 0; dj < 2; ++dj) {
            double wyj = dj == 0 ? wy0 : 1.0 - wy0;
            for (int dk = 0; dk < 2; ++dk) {
              double wzk = dk == 0 ? wz0 : 1.0 - wz0;
              acc = acc + wxi * wyj * wzk * C->phi[LIDX(C, c0x + di, c0y + dj, c0z + dk)];
            }
          }
        }
        long long x = LIDX(F, i, j, k);
        F->phi[x] = F->phi[x] + acc;
      }
  refill(F);
}

static void vcycle(Solver *S) {
  const int nl = S->nlev;
  for (int l = 0; l < nl; ++l) {
    if (l > 0) restrict_(&S->lev[l - 1], &S->lev[l]);
    relax(&S->lev[l], l < nl - 1 ? S->npre : S->ncoarse);
  }
  for (int l = nl - 2; l >= 0; --l) {
    prolong(&S->lev[l + 1], &S->lev[l]);
    relax(&S->lev[l], S->npost);
  }
}

static void mg_solve(Solver *S) {
  const FGrid g = S->g;
  Level *L = &S->lev[0];
  for (int i = 0; i < L->n0; ++i)
    for (int j = 0; j < L->n1; ++j)
      for (int k = 0; k < L->n2; ++k) L->rhs[LIDX(L, i, j, k)] = S->f[SU][IDX(g, i + NPL, j + NPL, k + NPL)];
  memset(S->f[PP], 0, sizeof(double) * g.size);
  memset(L->phi, 0, sizeof(double) * L->size);
  for (int c = 0; c < S->nvcycles; ++c) vcycle(S);
  for (int i = 0; i < L->n0 + 2; ++i)
    for (int j = 0; j < L->n1 + 2; ++j)
      for (int k = 0; k < L->n2 + 2; ++k)
        S->f[PP][IDX(g, i + NPL - 1, j + NPL - 1, k + NPL - 1)] =
            L->phi[(long long)i * L->s0 + (long long)j * L->s1 + k];
}

/* ------------------------------------------------------------- solver.py */
static int compute_dt(Solver *S) {
  const FGrid g = S->g;
  double mu = 0, mv = 0, mw = 0, dtv = 1.0e10;
  const double rsum = 1.0 / (S->dx * S->dx) + 1.0 / (S->dy * S->dy) + 1.0 / (S->dz * S->dz);
  const double *u = S->f[U], *v = S->f[V], *w = S->f[W], *vis = S->f[VIS];
#pragma omp parallel for schedule(static) reduction(max : mu, mv, mw) reduction(min : dtv)
  for (int i = NPL; i < NPL + g.n0; ++i)
    for (int j = NPL; j < NPL + g.n1; ++j)
      for (int k = NPL; k < NPL + g.n2; ++k) {
        long long x = IDX(g, i, j, k);
        if (fabs(u[x]) > mu) mu = fabs(u[x]);
        if (fabs(v[x]) > mv) mv = fabs(v[x]);
        if (fabs(w[x]) > mw) mw = fabs(w[x]);
        double uc = 0.5 * (u[x] + u[x - g.s0]), vc = 0.5 * (v[x] + v[x - g.s1]),
               wc = 0.5 * (w[x] + w[x - 1]);
        double val = 1.0 / (fabs(uc) / S->dx + fabs(vc) / S->dy + fabs(wc) / S->dz +
                            2.0 * vis[x] * rsum + 1.0e-30);
        if (val < dtv) dtv = val;
      }
  double big = 1.0e10, dtcfl = big;
  if (mu > 0.0 && S->dx / mu < dtcfl) dtcfl = S->dx / mu;
  if (mv > 0.0 && S->dy / mv < dtcfl) dtcfl = S->dy / mv;
  if (mw > 0.0 && S->dz / mw < dtcfl) dtcfl = S->dz / mw;
  double dtmax = dtv < dtcfl ? dtv : dtcfl;
  if (dtmax <= 0.0 || !(dtmax < big)) return -1;
  S->cfl = S->dt / dtmax;
  if (S->variable_dt) {
    double a = S->dt * S->dt_growth, b = S->safety * dtmax;
    S->dt = a < b ? a : b;
  }
  return 0;
}

static int project(Solver *S) {
  calvel(S, 1.0);
  int iters = 0;
  S->nhist = 0;
  for (int it = 1; it <= S->max_outer; ++it) {
    iters = it;
    S->rmax = calmas(S);
    if (S->nhist < 64) S->rmax_hist[S->nhist++] = S->rmax;
    if (S->rmax < S->eps && it > 1) break;
    if (S->rmax > 10.0) return -2;
    mg_solve(S);
    const FGrid g = S->g;
    for (int i = NPL - 1; i < NPL + g.n0 + 1; ++i)
      for (int j = NPL - 1; j < NPL + g.n1 + 1; ++j)
        for (int k = NPL - 1; k < NPL + g.n2 + 1; ++k) {
          long long x = IDX(g, i, j, k);
          S->f[P][x] = S->f[P][x] + S->f[PP][x];
        }
    apply_fill(S->g, S->f[P], 4, 0.0);
    calvel(S, 1.0);
  }
  S->outer_iters = iters;
  apply_fill(S->g, S->f[U], 0, S->ulid);
  apply_fill(S->g, S->f[V], 1, 0.0);
  apply_fill(S->g, S->f[W], 2, 0.0);
  return 0;
}

static void fill_bc(Solver *S, int c) {
  apply_fill(S->g, S->f[U + c], c, c == 0 ? S->ulid : 0.0);
}

/* ------------------------------------------------------------- C ABI --- */
void *wfo_create(int n0, int n1, int n2, double lx, double ly, double lz, double reynolds,
                 int sgs, double cw, double noise, int seed, double ulid, int mg_levels,
                 int max_outer, double eps, int variable_dt, double dt_init, double safety,
                 double growth) {
  Solver *S = (Solver *)calloc(1, sizeof(Solver));
  if (!S || mg_levels < 1 || mg_levels > 16) { free(S); return 0; }
  FGrid g;
  g.n0 = n0; g.n1 = n1; g.n2 = n2;
  g.s1 = n2 + 2 * NPL;
  g.s0 = (long long)(n1 + 2 * NPL) * g.s1;
  g.size = (long long)(n0 + 2 * NPL) * g.s0;
  S->g = g;
  for (int q = 0; q < 13; ++q) S->f[q] = (double *)calloc((size_t)g.size, sizeof(double));
  S->dx = lx / n0; S->dy = ly / n1; S->dz = lz / n2;
  S->rey = 1.0 / reynolds; S->sgs = sgs; S->cw = cw; S->ulid = ulid;
  S->nlev = mg_levels; S->max_outer = max_outer; S->eps = eps;
  S->variable_dt = variable_dt; S->dt = dt_init; S->safety = safety; S->dt_growth = growth;
  S->npre = 6; S->npost = 6; S->ncoarse = 300; S->nvcycles = 4;
  for (int l = 0; l < mg_levels; ++l) {
    Level *L = &S->lev[l];
    L->n0 = n0 >> l; L->n1 = n1 >> l; L->n2 = n2 >> l;
    L->s1 = L->n2 + 2;
    L->s0 = (long long)(L->n1 + 2) * L->s1;
    L->size = (long long)(L->n0 + 2) * L->s0;
    L->h0 = S->dx * pow(2.0, l); L->h1 = S->dy * pow(2.0, l); L->h2 = S->dz * pow(2.0, l);
    L->phi = (double *)calloc((size_t)L->size, sizeof(double));
    L->rhs = (double *)calloc((size_t)L->size, sizeof(double));
  }
  if (noise > 0.0) init_noise(S, noise, (uint32_t)seed);
  for (long long q = 0; q < g.size; ++q) S->f[VIS][q] = S->rey;
  apply_fill(g, S->f[U], 0, ulid);
  apply_fill(g, S->f[V], 1, 0.0);
  apply_fill(g, S->f[W], 2, 0.0);
  apply_fill(g, S->f[P], 5, 3.0);
  return S;
}

void wfo_destroy(void *h) {
  Solver *S = (Solver *)h;
  if (!S) return;
  for (int q = 0; q < 13; ++q) free(S->f[q]);
  for (int l = 0; l < S->nlev; ++l) { free(S->lev[l].phi); free(S->lev[l].rhs); }
  free(S);
}

/* info: ctime, dt, rmax, iters, cfl.  Returns 0, -1 (dt collapse), -2 (diverged). */
int wfo_step(void *h, double *info) {
  Solver *S = (Solver *)h;
  S->itime += 1;
  if (compute_dt(S)) return -1;
  S->ctime += S->dt;
  const long long n = S->g.size;
  memcpy(S->f[UOO], S->f[U], sizeof(double) * n);
  memcpy(S->f[VOO], S->f[V], sizeof(double) * n);
  memcpy(S->f[WOO], S->f[W], sizeof(double) * n);
  if (S->sgs) {
    wale(S);
    apply_fill(S->g, S->f[VIS], 3, S->rey);
  } else {
    fill_interior(S, S->f[VIS], S->rey);
  }
  const double alf[2] = {0.5, 1.0};
  for (int st = 0; st < 2; ++st) {
    for (int c = 0; c < 3; ++c) { fill_bc(S, c); conv4th(S, c, S->dt, alf[st]); }
    for (int c = 0; c < 3; ++c) { fill_bc(S, c); diff4th(S, c, S->dt, alf[st]); }
    if (st == 0) calvel(S, alf[st]);
  }
  int r = project(S);
  if (info) {
    info[0] = S->ctime; info[1] = S->dt; info[2] = S->rmax;
    info[3] = S->outer_iters; info[4] = S->cfl;
  }
  return r;
}

/* field ids: u v w p ustar vstar wstar uoo voo woo vis su pp (0..12) */
int wfo_get_field(void *h, int id, double *out) {
  Solver *S = (Solver *)h;
  if (id < 0 || id > 12) return -1;
  memcpy(out, S->f[id], sizeof(double) * S->g.size);
  return 0;
}

int wfo_set_field(void *h, int id, const double *in) {
  Solver *S = (Solver *)h;
  if (id < 0 || id > 12) return -1;
  memcpy(S->f[id], in, sizeof(double) * S->g.size);
  return 0;
}

int wfo_rmax_history(void *h, double *out, int cap) {
  Solver *S = (Solver *)h;
  int n = S->nhist < cap ? S->nhist : cap;
  memcpy(out, S->rmax_hist, sizeof(double) * n);
  return n;
}

int wfo_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads for the following steps (timing legs: 1 thread, all cores). */
void wfo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
