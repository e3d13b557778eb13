// Momentum, projection and SGS kernels of one RK2 fractional step
// (the reference's warpflow/kernels.py), FP64, for gfx950.
//
// Every kernel maps one thread to one cell of its launch box: 64 lanes along
// the unit-stride k axis, 4 j-rows per workgroup, one i-plane per grid.z, so
// each wavefront reads and writes whole 512-byte k-rows.  Expressions keep
// the reference's evaluation order (left to right) so results agree with
// the FP64 oracle to a few ulps.
#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "wf_common.h"

namespace wf {

// 4th-order midpoint interpolant along stride s at c (kernels.py:31-46):
// (-f[c+s] + 9 f[c] + 9 f[c-s] - f[c-2s]) / 16
__device__ inline double I4(const double* __restrict__ f, int64_t c, int64_t s) {
  return (-f[c + s] + 9.0 * f[c] + 9.0 * f[c - s] - f[c - 2 * s]) / 16.0;
}

// Staggered 4-point derivative (kernels.py:69, :80, :281-286):
// (-s[+] + 27 s[0] - 27 s[-] + s[--]) / (24 h)
__device__ inline double D4(double sp, double sn, double sm, double smm, double h) {
  return (-sp + 27.0 * sn - 27.0 * sm + smm) / (24.0 * h);
}
// The same with the division replaced by r24h = 1 / (24 h) formed once.
__device__ inline double D4r(double sp, double sn, double sm, double smm, double r24h) {
  return (-sp + 27.0 * sn - 27.0 * sm + smm) * r24h;
}

// d(c c)/dx_a for the component c living on axis a (kernels.py:64-69,
// :103-108, :143-148).
__device__ inline double conv_self(const double* __restrict__ c, int64_t x, int64_t sa,
                                   double ha) {
  const double p2 = I4(c, x + 2 * sa, sa), p1 = I4(c, x + sa, sa);
  const double p0 = I4(c, x, sa), m1 = I4(c, x - sa, sa);
  return D4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, ha);
}

// d(c q)/dx_b for component c (axis a) and the component q along the cross
// axis b: c is interpolated along b at b+2..b-1, q along a at a+1 and
// b+1..b-2, so the derivative lands on c's face (kernels.py:71-87,
// :110-127, :150-166).
__device__ inline double conv_cross(const double* __restrict__ c,
                                    const double* __restrict__ q, int64_t x,
                                    int64_t sa, int64_t sb, double hb) {
  const double sp1 = I4(c, x + 2 * sb, sb) * I4(q, x + sa + sb, sa);
  const double sn = I4(c, x + sb, sb) * I4(q, x + sa, sa);
  const double sm1 = I4(c, x, sb) * I4(q, x + sa - sb, sa);
  const double sm2 = I4(c, x - sb, sb) * I4(q, x + sa - 2 * sb, sa);
  return D4(sp1, sn, sm1, sm2, hb);
}

// N(c) = x-term + y-term + z-term, the reference's summation order for all
// three components (kernels.py:89, :129, :168).
template <int C>
__device__ inline double conv_sum(const double* __restrict__ u,
                                  const double* __restrict__ v,
                                  const double* __restrict__ w, int64_t x,
                                  const wf_grid& g, double dx, double dy, double dz) {
  const int64_t sx = g.si, sy = g.sj, sz = 1;
  if (C == 0) {
    return conv_self(u, x, sx, dx) + conv_cross(u, v, x, sx, sy, dy) +
           conv_cross(u, w, x, sx, sz, dz);
  } else if (C == 1) {
    return conv_cross(v, u, x, sy, sx, dx) + conv_self(v, x, sy, dy) +
           conv_cross(v, w, x, sy, sz, dz);
  } else {
    return conv_cross(w, u, x, sz, sx, dx) + conv_cross(w, v, x, sz, sy, dy) +
           conv_self(w, x, sz, dz);
  }
}

// 4th-order Laplacian (kernels.py:177-186); r12xx = 12*dx*dx etc.
__device__ inline double lap4(const double* __restrict__ f, int64_t x, const wf_grid& g,
                              double r12xx, double r12yy, double r12zz) {
  const int64_t sx = g.si, sy = g.sj;
  const double f0 = f[x];
  const double lx = (-f[x + 2 * sx] + 16.0 * f[x + sx] - 30.0 * f0 + 16.0 * f[x - sx] -
                     f[x - 2 * sx]) / r12xx;
  const double ly = (-f[x + 2 * sy] + 16.0 * f[x + sy] - 30.0 * f0 + 16.0 * f[x - sy] -
                     f[x - 2 * sy]) / r12yy;
  const double lz =
      (-f[x + 2] + 16.0 * f[x + 1] - 30.0 * f0 + 16.0 * f[x - 1] - f[x - 2]) / r12zz;
  return lx + ly + lz;
}

template <int C>
__device__ inline int64_t vis_offset(const wf_grid& g) {
  return C == 0 ? 0 : (C == 1 ? g.sj : 1);  // kernels.py:198, :211, :224
}

template <int C>
__global__ void __launch_bounds__(256)
conv4th_kernel(const double* __restrict__ u, const double* __restrict__ v,
               const double* __restrict__ w, const double* __restrict__ coo,
               double* __restrict__ cstar, wf_grid g, Box3 b, double dx, double dy,
               double dz, double dta, int ich) {
  for_planes(b.d0, ich, [&](Blk B) {
    int i, j, k;
    if (!box_thread(b, B, i, j, k)) return;
    const int64_t x = at(g, i, j, k);
    cstar[x] = coo[x] - dta * conv_sum<C>(u, v, w, x, g, dx, dy, dz);
  });
}

template <int C>
__global__ void __launch_bounds__(256)
diff4th_kernel(const double* __restrict__ c, const double* __restrict__ vis,
               double* __restrict__ cstar, wf_grid g, Box3 b, double r12xx,
               double r12yy, double r12zz, double dta, int ich) {
  for_planes(b.d0, ich, [&](Blk B) {
    int i, j, k;
    if (!box_thread(b, B, i, j, k)) return;
    const int64_t x = at(g, i, j, k);
    const double diff = vis[x + vis_offset<C>(g)] * lap4(c, x, g, r12xx, r12yy, r12zz);
    cstar[x] = cstar[x] + dta * diff;
  });
}

// The pressure correction of one velocity value, c = cstar - dt*alpha*dp/h
// (kernels.py:239, :250, :261), with 1/h formed on the host.  The rounding
// is fixed here (one product, one fused multiply-add) rather than left to
// the compiler, so that every kernel computing a calvel value -- the three
// calvel launches, calvel3 and the fused calvel/calmas pass, single-domain
// or batched -- produces the same bits.
__device__ inline double calvel_val(double cstar, double pn, double p0, double dta, double rh) {
#pragma clang fp contract(off)
  return __builtin_fma(-rh, dta * (pn - p0), cstar);
}
// The bracket of the staggered 4-point derivative, -s[+] + 27 s[0] - 27 s[-]
// + s[--] (kernels.py:281-286), and the divergence dudx + dvdy + dwdz from
// the three brackets and 1/(24 h): fixed roundings, shared by calmas and the
// fused calvel/calmas pass.
__device__ inline double d4_bracket(double sp, double sn, double sm, double smm) {
#pragma clang fp contract(off)
  return __builtin_fma(-27.0, sm, __builtin_fma(27.0, sn, -sp)) + smm;
}
__device__ inline double div4(double tx, double ty, double tz, double r24x, double r24y,
                              double r24z) {
#pragma clang fp contract(off)
  return __builtin_fma(tz, r24z, __builtin_fma(ty, r24y, tx * r24x));
}

// x / d rounded to nearest, from rd = 1/d rounded to nearest: q = x rd is
// within an ulp of x/d, the FMA gives the remainder x - q d exactly, and one
// correction step returns the correctly rounded quotient (Markstein's
// theorem; it holds for finite x whose quotient is normal or zero).  Three
// FP64 operations in place of the ~11 of a full division, same bits.
__device__ inline double div_rn(double x, double d, double rd) {
#pragma clang fp contract(off)
  const double q = x * rd;
  const double r = __builtin_fma(-q, d, x);
  return __builtin_fma(r, rd, q);
}

// The viscous/advective time-step bound of one cell (kernels.py:384-402) is
// 1 / X, X the denominator below, from the velocities at the cell and its
// low neighbours and the cell's vis, with fixed roundings (no contraction)
// so that checkdt and the fused WALE pass agree bit for bit; rdx, rdy, rdz
// are 1/dx, 1/dy, 1/dz rounded to nearest.  X > 0 and 1/x rounded is
// monotone in x, so the reductions take the min of the bounds as one
// reciprocal of the max of X: the same bits, one division per workgroup
// instead of one per cell.
__device__ inline double dtvisc_den(double u0, double um, double v0, double vm, double w0,
                                    double wm, double vis, double dx, double dy, double dz,
                                    double rdx, double rdy, double rdz, double rsum) {
#pragma clang fp contract(off)
  const double uc = 0.5 * (u0 + um);
  const double vc = 0.5 * (v0 + vm);
  const double wc = 0.5 * (w0 + wm);
  return div_rn(fabs(uc), dx, rdx) + div_rn(fabs(vc), dy, rdy) + div_rn(fabs(wc), dz, rdz) +
         2.0 * vis * rsum + 1.0e-30;
}

// min(1e10, 1 / X over the cells) from xmax = the max of X (0 for no cells),
// NaN if any X is.
__device__ inline double dtvisc_of_max(double xmax) { return nan_min(1.0e10, 1.0 / xmax); }
__global__ void __launch_bounds__(256)
calvel_kernel(const double* __restrict__ cstar, const double* __restrict__ p,
              double* __restrict__ c, wf_grid g, Box3 b, int64_t s, double rh, double dta,
              int ich) {
  for_planes(b.d0, ich, [&](Blk B) {
    int i, j, k;
    if (!box_thread(b, B, i, j, k)) return;
    const int64_t x = at(g, i, j, k);
    c[x] = calvel_val(cstar[x], p[x + s], p[x], dta, rh);
  });
}

// All three calvel launches over the real cells; component c is written
// only inside its operator range (solver.py:99-110).
__device__ inline void calvel3_body(const double* __restrict__ us, const double* __restrict__ vs,
                                    const double* __restrict__ ws, const double* __restrict__ p,
                                    double* __restrict__ u, double* __restrict__ v,
                                    double* __restrict__ w, const wf_grid& g, const Box3& b,
                                    Blk B, double rdx, double rdy, double rdz, double dta) {
  int i, j, k;
  if (!box_thread(b, B, i, j, k)) return;
  const int64_t x = at(g, i, j, k);
  const double p0 = p[x];
  if (i < g.npl + g.ni - 1) u[x] = calvel_val(us[x], p[x + g.si], p0, dta, rdx);
  if (j < g.npl + g.nj - 1) v[x] = calvel_val(vs[x], p[x + g.sj], p0, dta, rdy);
  if (k < g.npl + g.nk - 1) w[x] = calvel_val(ws[x], p[x + 1], p0, dta, rdz);
}

__global__ void __launch_bounds__(256)
calvel3_kernel(const double* __restrict__ us, const double* __restrict__ vs,
               const double* __restrict__ ws, const double* __restrict__ p,
               double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
               wf_grid g, Box3 b, double rdx, double rdy, double rdz, double dta, int ich) {
  for_planes(b.d0, ich,
             [&](Blk B) { calvel3_body(us, vs, ws, p, u, v, w, g, b, B, rdx, rdy, rdz, dta); });
}

// Block-wide reductions for 256-thread (4-wave) workgroups; every thread of
// the block must call them.
__device__ inline double block_max(double v) {
  __shared__ double part[4];
  v = wave_max(v);
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) part[t >> 6] = v;
  __syncthreads();
  double r = part[0];
  for (int q = 1; q < 4; ++q) r = nan_max(r, part[q]);
  return r;
}
__device__ inline bool block_leader() { return threadIdx.x == 0 && threadIdx.y == 0; }

// Reductions: a workgroup covers a 64 x 4 (k, j) tile over a chunk of `ich`
// i-planes, reduces it in registers and LDS, and stores one value per
// quantity in a partials slot of a static device array (referenced only by
// device code); one workgroup then folds the slots.  No atomics on a shared
// address, and max/min are order-independent, so results are deterministic.
constexpr int RED_MAX_BLOCKS = 4096;
constexpr int RED_NQ = 4;
__device__ double g_red_part[RED_NQ * RED_MAX_BLOCKS];

__device__ inline int red_block() {
  return (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

__device__ inline bool red_rows(const Box3& b, Blk B, int ich, int& j, int& k, int& ibeg,
                                int& iend) {
  k = B.x * BX + threadIdx.x;
  j = B.y * BY + threadIdx.y;
  ibeg = B.z * ich;
  iend = min(ibeg + ich, b.d0);
  return k < b.d2 && j < b.d1;
}
__device__ inline bool red_rows(const Box3& b, int ich, int& j, int& k, int& ibeg, int& iend) {
  return red_rows(b, this_blk(), ich, j, k, ibeg, iend);
}

// max|div dV| of one workgroup's rows into part[slot]; su = div / dt.
__device__ inline void calmas_body(const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ w, double* __restrict__ su,
                                   const wf_grid& g, const Box3& b, Blk B, int ich, double dx,
                                   double dy, double dz, double dt, double* part, int slot) {
  const double r24x = 1.0 / (24.0 * dx), r24y = 1.0 / (24.0 * dy), r24z = 1.0 / (24.0 * dz);
  const double rdt = 1.0 / dt;
  int j, k, ibeg, iend;
  double val = 0.0;
  if (red_rows(b, B, ich, j, k, ibeg, iend)) {
    const int64_t sx = g.si, sy = g.sj;
    for (int ii = ibeg; ii < iend; ++ii) {
      const int64_t x = at(g, b.lo0 + ii, b.lo1 + j, b.lo2 + k);
      // dudx + dvdy + dwdz (kernels.py:281-287)
      const double fact = div4(d4_bracket(u[x + sx], u[x], u[x - sx], u[x - 2 * sx]),
                               d4_bracket(v[x + sy], v[x], v[x - sy], v[x - 2 * sy]),
                               d4_bracket(w[x + 1], w[x], w[x - 1], w[x - 2]), r24x, r24y,
                               r24z);
      su[x] = fact * rdt;
      val = nan_max(val, fabs(fact * dx * dy * dz));
    }
  }
  const double r = block_max(val);
  if (block_leader()) part[slot] = r;
}

__global__ void __launch_bounds__(256)
calmas_kernel(const double* __restrict__ u, const double* __restrict__ v,
              const double* __restrict__ w, double* __restrict__ su, wf_grid g, Box3 b,
              int ich, double dx, double dy, double dz, double dt) {
  calmas_body(u, v, w, su, g, b, this_blk(), ich, dx, dy, dz, dt, g_red_part, red_block());
}

// calvel3 followed by calmas in one pass.  The projection loop always runs
// calmas right after a calvel of all three components (solver.py:186-197;
// the halo exchange between them does nothing on one domain), and calmas
// then re-reads the u, v, w calvel has just written.  Here a workgroup owns
// an 8 x 32 (j, k) tile of the real cells and streams a chunk of i-planes:
//   * u is corrected one plane ahead in the thread's own column and kept in
//     a four-plane register window (the x-derivative needs planes i+1 .. i-2);
//   * v and w of plane i are corrected at the tile cells and at the halo the
//     y- and z-derivatives reach (rows -2, -1, +8 of v; columns -2, -1, +32
//     of w), staged in LDS (double-buffered by plane parity, one barrier per
//     plane), and the divergence of plane i is formed from them.
// A halo value is recomputed with the same expression as the tile that owns
// it, and a position outside calvel's write range (walls, ghosts; calvel3_body
// above) contributes its value in memory, as calmas would read it, so u, v,
// w, su and rmax equal calvel3 + calmas bit for bit.  Each value is written
// once, by the workgroup owning its cell and plane.  HBM traffic falls from
// 88 to 64 bytes per cell (calmas no longer reads u, v, w).
namespace pj {
// v: rows -2 .. TJ, columns 0 .. TK-1; w: rows 0 .. TJ-1, columns -2 .. TK
template <int TJ, int TK>
struct Tile {
  static constexpr int NT = TJ * TK, VW = TK, VR = TJ + 3, WW = TK + 3;
  static constexpr int NVH = 3 * TK, NWH = 3 * TJ;
  static_assert(NVH <= NT && NWH <= NT, "one halo position of each kind per thread");
  __device__ static int vix(int r, int c) { return (r + 2) * VW + c; }
  __device__ static int wix(int r, int c) { return r * WW + (c + 2); }
};
struct Consts {
  double rdx, rdy, rdz, dta, r24x, r24y, r24z, rdt, dx, dy, dz;
};
}  // namespace pj

// One (k tile, j tile, i chunk) work item of the fused pass on a TJ x TK
// tile (TJ * TK threads, one per tile cell); returns the item's max|div dV|
// to every thread.  STORE = false: u, v, w are corrected and used but not
// written (su, rmax and the FP32 rhs are the same).  DEPTH: a plane's loads
// are issued DEPTH planes ahead (2: two register banks, the loop takes two
// planes per trip).
template <int TJ, int TK, bool STORE = true, int DEPTH = 1>
__device__ inline double calvel3_calmas_body(
    const double* __restrict__ us, const double* __restrict__ vs,
    const double* __restrict__ ws, const double* __restrict__ p, double* __restrict__ u,
    double* __restrict__ v, double* __restrict__ w, double* __restrict__ su, const wf_grid& g,
    int ich, int ktile, int jtile, int chunk, const pj::Consts& C,
    float* __restrict__ r32 = nullptr) {
  using T = pj::Tile<TJ, TK>;
  __shared__ double Vs[2][T::VR * T::VW];
  __shared__ double Ws[2][TJ * T::WW];
  const int tid = threadIdx.x, tj = tid / TK, tk = tid % TK;
  const int npl = g.npl;
  const int NI = g.ni + 2 * npl, NJ = g.nj + 2 * npl, NK = g.nk + 2 * npl;
  const int J0 = npl + jtile * TJ, K0 = npl + ktile * TK;
  const int I0 = npl + chunk * ich, I1 = npl + min(chunk * ich + ich, g.ni);
  const int J = J0 + tj, K = K0 + tk;
  const bool live = J < npl + g.nj && K < npl + g.nk;
  const double dta = C.dta;
  // in-plane offset of (j, k), clamped into the array (clamped positions are
  // never read by a live cell)
  auto at2 = [=](int j, int k) -> int64_t {
    return (int64_t)min(max(j, 0), NJ - 1) * g.sj + min(max(k, 0), NK - 1);
  };
  auto plane = [=](int i) -> int64_t { return (int64_t)min(max(i, 0), NI - 1) * g.si; };
  // calvel's write ranges (calvel3_body)
  auto u_in = [=](int i) { return i >= npl && i < npl + g.ni - 1; };
  auto v_in = [=](int j, int k) {
    return j >= npl && j < npl + g.nj - 1 && k >= npl && k < npl + g.nk;
  };
  auto w_in = [=](int j, int k) {
    return k >= npl && k < npl + g.nk - 1 && j >= npl && j < npl + g.nj;
  };

  const int64_t xo = at2(J, K), xoj = at2(J + 1, K), xok = at2(J, K + 1);
  const bool vo_in = v_in(J, K), wo_in = w_in(J, K);
  const bool hv = tid < T::NVH, hw = tid < T::NWH;
  int vr = 0, vc = 0, wr = 0, wc = 0;
  if (hv) {
    const int q = tid / TK;
    vr = q < 2 ? q - 2 : TJ;
    vc = tid % TK;
  }
  if (hw) {
    const int q = tid / TJ;
    wr = tid % TJ;
    wc = q < 2 ? q - 2 : TK;
  }
  const int64_t xv = at2(J0 + vr, K0 + vc), xv1 = at2(J0 + vr + 1, K0 + vc);
  const int64_t xw = at2(J0 + wr, K0 + wc), xw1 = at2(J0 + wr, K0 + wc + 1);
  const bool vh_in = hv && v_in(J0 + vr, K0 + vc), wh_in = hw && w_in(J0 + wr, K0 + wc);
  const double* __restrict__ vos = vo_in ? vs : v;  // source of the own v value
  const double* __restrict__ wos = wo_in ? ws : w;
  const double* __restrict__ vhs = vh_in ? vs : v;
  const double* __restrict__ whs = wh_in ? ws : w;

  // u at plane i of the own column: corrected inside calvel's range, else
  // the value in memory
  auto u_at = [=](int i, double a, double pn, double p0) {
    return u_in(i) ? calvel_val(a, pn, p0, dta, C.rdx) : a;
  };
  auto u_src = [=](int i) { return u_in(i) ? us : u; };

  // prologue: u(I0), u(I0-1), u(I0-2) in uw[0..2]; at the top of plane i the
  // window shifts and uw[0..3] = u at planes i+1, i, i-1, i-2
  double uw[4];
  {
    double pq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) pq[q] = p[plane(I0 - 2 + q) + xo];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int i = I0 - 2 + q;
      uw[2 - q] = u_at(i, u_src(i)[plane(i) + xo], pq[q + 1], pq[q]);
    }
    if (STORE && live && u_in(I0)) u[plane(I0) + xo] = uw[0];
    uw[3] = 0.0;
  }
  double p0 = p[plane(I0) + xo], p1 = p[plane(I0 + 1) + xo];

  // loads of plane i (issued during plane i-1)
  struct Grp {
    double ua, p2, va, pv, wa, pw, vh, vh0, vh1, wh, wh0, wh1;
  };
  auto issue = [&](Grp& G, int i) {
    i = min(i, I1 - 1);
    const int64_t pi = plane(i);
    G.ua = u_src(i + 1)[plane(i + 1) + xo];
    G.p2 = p[plane(i + 2) + xo];
    G.va = vos[pi + xo];
    G.pv = p[pi + xoj];
    G.wa = wos[pi + xo];
    G.pw = p[pi + xok];
    if (hv) {
      G.vh = vhs[pi + xv];
      G.vh0 = p[pi + xv];
      G.vh1 = p[pi + xv1];
    }
    if (hw) {
      G.wh = whs[pi + xw];
      G.wh0 = p[pi + xw];
      G.wh1 = p[pi + xw1];
    }
  };
  Grp A, B;
  A.vh = A.vh0 = A.vh1 = A.wh = A.wh0 = A.wh1 = 0.0;
  B.vh = B.vh0 = B.vh1 = B.wh = B.wh0 = B.wh1 = 0.0;
  issue(A, I0);
  if constexpr (DEPTH == 2) issue(B, I0 + 1);

  double val = 0.0;
  // plane i on the loads N holds; then N takes those of plane i + DEPTH
  auto step = [&](const int i, Grp& N) {
    const bool on = DEPTH == 1 || i < I1;  // the second bank's plane past I1
    const int buf = (i - I0) & 1;
    // u of plane i+1 into the window
    uw[3] = uw[2];
    uw[2] = uw[1];
    uw[1] = uw[0];
    uw[0] = u_at(i + 1, N.ua, N.p2, p1);
    const double vo = vo_in ? calvel_val(N.va, N.pv, p0, dta, C.rdy) : N.va;
    const double wo = wo_in ? calvel_val(N.wa, N.pw, p0, dta, C.rdz) : N.wa;
    double* __restrict__ V = Vs[buf];
    double* __restrict__ W = Ws[buf];
    V[T::vix(tj, tk)] = vo;
    W[T::wix(tj, tk)] = wo;
    if (hv) V[T::vix(vr, vc)] = vh_in ? calvel_val(N.vh, N.vh1, N.vh0, dta, C.rdy) : N.vh;
    if (hw) W[T::wix(wr, wc)] = wh_in ? calvel_val(N.wh, N.wh1, N.wh0, dta, C.rdz) : N.wh;
    if (STORE && live && on) {
      if (u_in(i + 1) && i + 1 < I1) u[plane(i + 1) + xo] = uw[0];
      if (vo_in) v[plane(i) + xo] = vo;
      if (wo_in) w[plane(i) + xo] = wo;
    }
    p0 = p1;
    p1 = N.p2;
    issue(N, i + DEPTH);
    __syncthreads();
    if (live && on) {
      // dudx + dvdy + dwdz at (i, J, K) (kernels.py:281-287)
      const double fact =
          div4(d4_bracket(uw[0], uw[1], uw[2], uw[3]),
               d4_bracket(V[T::vix(tj + 1, tk)], V[T::vix(tj, tk)], V[T::vix(tj - 1, tk)],
                          V[T::vix(tj - 2, tk)]),
               d4_bracket(W[T::wix(tj, tk + 1)], W[T::wix(tj, tk)], W[T::wix(tj, tk - 1)],
                          W[T::wix(tj, tk - 2)]),
               C.r24x, C.r24y, C.r24z);
      const double sv = fact * C.rdt;
      su[plane(i) + xo] = sv;
      if (r32)  // (n+2)^3 level array, one-cell ring: real cell (i, J, K) - npl + 1
        r32[((int64_t)(i - npl + 1) * (g.nj + 2) + (J - npl + 1)) * (g.nk + 2) + (K - npl + 1)] =
            (float)sv;
      val = nan_max(val, fabs(fact * C.dx * C.dy * C.dz));
    }
  };
  if constexpr (DEPTH == 1) {
    for (int i = I0; i < I1; ++i) step(i, A);
  } else {
    for (int i = I0; i < I1; i += 2) {
      step(i, A);
      step(i + 1, B);
    }
  }
  return block_max(val);
}

// Single domain: 8 x 32 tiles; the workgroup ids are remapped over the XCDs
// (k tile fastest, then j tile, then i-chunk).
constexpr int PJ_TJ = 8, PJ_TK = 32;
template <bool STORE>
__global__ void __launch_bounds__(256)
calvel3_calmas_kernel(const double* __restrict__ us, const double* __restrict__ vs,
                      const double* __restrict__ ws, const double* __restrict__ p,
                      double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                      double* __restrict__ su, wf_grid g, int ich, int kt, int jt, pj::Consts K,
                      float* __restrict__ r32) {
  const int nb = kt * jt * (int)gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  const double r = calvel3_calmas_body<PJ_TJ, PJ_TK, STORE, 2>(
      us, vs, ws, p, u, v, w, su, g, ich, L % kt, (L / kt) % jt, L / (kt * jt), K, r32);
  if (threadIdx.x == 0) g_red_part[blockIdx.x + blockIdx.y * gridDim.x] = r;
}

// Fold the partials of nq quantities: bit q of min_mask selects min (else
// max); combine != 0 folds the previous *out in too (the reference's
// atomic_max/atomic_min into a pre-set buffer).
__global__ void __launch_bounds__(1024)
red_fold_kernel(double* __restrict__ out, int nblk, int nq, int min_mask, int combine) {
  __shared__ double part[16];
  const int t = threadIdx.x;
  for (int q = 0; q < nq; ++q) {
    const bool mn = (min_mask >> q) & 1;
    double v = mn ? 1.0e10 : 0.0;
    for (int p = t; p < nblk; p += 1024) {
      const double y = g_red_part[q * RED_MAX_BLOCKS + p];
      v = mn ? nan_min(v, y) : nan_max(v, y);
    }
    v = mn ? wave_min(v) : wave_max(v);
    if ((t & 63) == 0) part[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
      double r = part[0];
      for (int w = 1; w < 16; ++w) r = mn ? nan_min(r, part[w]) : nan_max(r, part[w]);
      out[q] = combine ? (mn ? nan_min(out[q], r) : nan_max(out[q], r)) : r;
    }
    __syncthreads();
  }
}

// WALE viscosity at one cell (kernels.py:297-367) from accessors U, V, W
// returning the velocity component at plane/row/column offset (di, dj, dk).
template <class FU, class FV, class FW>
__device__ inline double wale_value(const FU& U, const FV& V, const FW& W, double rdx,
                                    double rdy, double rdz, double rrey, double ls) {
  // Roundings are fixed here -- each sum of products accumulates with fused
  // multiply-adds left to right, nothing else is contracted -- so that every
  // kernel inlining this (WALE alone, with checkdt, batched) gives the same bits.
#pragma clang fp contract(off)
  // kernels.py:311-327
  const double dudx = (U(0, 0, 0) - U(-1, 0, 0)) * rdx;
  const double dudy = (0.25 * (U(0, 1, 0) + U(-1, 1, 0) - U(0, -1, 0) - U(-1, -1, 0))) * rdy;
  const double dudz = (0.25 * (U(0, 0, 1) + U(-1, 0, 1) - U(0, 0, -1) - U(-1, 0, -1))) * rdz;
  const double dvdx = (0.25 * (V(1, 0, 0) + V(1, -1, 0) - V(-1, 0, 0) - V(-1, -1, 0))) * rdx;
  const double dvdy = (V(0, 0, 0) - V(0, -1, 0)) * rdy;
  const double dvdz = (0.25 * (V(0, 0, 1) + V(0, -1, 1) - V(0, 0, -1) - V(0, -1, -1))) * rdz;
  const double dwdx = (0.25 * (W(1, 0, 0) + W(1, 0, -1) - W(-1, 0, 0) - W(-1, 0, -1))) * rdx;
  const double dwdy = (0.25 * (W(0, 1, 0) + W(0, 1, -1) - W(0, -1, 0) - W(0, -1, -1))) * rdy;
  const double dwdz = (W(0, 0, 0) - W(0, 0, -1)) * rdz;
  // a*b + c*d + e*f, left to right
  auto dot3 = [](double a, double b, double c, double d, double e, double f) {
    return __builtin_fma(e, f, __builtin_fma(c, d, a * b));
  };
  // kernels.py:330-338, g2zx and g2zy with the reference's last terms
  const double g2xx = dot3(dudx, dudx, dudy, dvdx, dudz, dwdx);
  const double g2yy = dot3(dvdx, dudy, dvdy, dvdy, dvdz, dwdy);
  const double g2zz = dot3(dwdx, dudz, dwdy, dvdz, dwdz, dwdz);
  const double g2xy = dot3(dudx, dudy, dudy, dvdy, dudz, dwdy);
  const double g2yx = dot3(dvdx, dudx, dvdy, dvdx, dvdz, dwdx);
  const double g2xz = dot3(dudx, dudz, dudy, dvdz, dudz, dwdz);
  const double g2zx = dot3(dwdx, dudx, dwdy, dvdx, dwdz, dudz);
  const double g2yz = dot3(dvdx, dudz, dvdy, dvdz, dvdz, dwdz);
  const double g2zy = dot3(dwdx, dudy, dwdy, dvdy, dwdz, dudy);
  // kernels.py:340-358
  const double tr = div_rn(g2xx + g2yy + g2zz, 3.0, 1.0 / 3.0);
  const double sdxx = g2xx - tr, sdyy = g2yy - tr, sdzz = g2zz - tr;
  const double sdxy = 0.5 * (g2xy + g2yx);
  const double sdxz = 0.5 * (g2xz + g2zx);
  const double sdyz = 0.5 * (g2yz + g2zy);
  const double sdsd = __builtin_fma(2.0, dot3(sdxy, sdxy, sdxz, sdxz, sdyz, sdyz),
                                    dot3(sdxx, sdxx, sdyy, sdyy, sdzz, sdzz));
  const double sxy = 0.5 * (dudy + dvdx);
  const double sxz = 0.5 * (dudz + dwdx);
  const double syz = 0.5 * (dvdz + dwdy);
  const double ss = __builtin_fma(2.0, dot3(sxy, sxy, sxz, sxz, syz, syz),
                                  dot3(dudx, dudx, dvdy, dvdy, dwdz, dwdz));
  // kernels.py:360-367 with x^1.5 = x*sqrt(x), |s|^2.5 = s*s*sqrt(|s|),
  // x^1.25 = x*sqrt(sqrt(x)) (exact identities; a few ulps from pow)
  double nut = 0.0;
  if (sdsd > 1.0e-8) {
    const double rs = sqrt(sdsd);
    const double sdsd15 = sdsd * rs;
    const double ass = fabs(ss);
    const double denom = __builtin_fma(sdsd, sqrt(rs), ass * ass * sqrt(ass));
    nut = ls * ls * sdsd15 / denom;
  }
  return rrey + nut;
}

// WALE streamed along i: an 8 x 32 (j, k) tile walks a chunk of planes with
// u, v, w staged in four rolling LDS planes (tile + 1-cell halo), so every
// global value is read once per tile and the plane written at iteration i
// (plane i+1) is the one last read at iteration i-2 -- one barrier per plane.
constexpr int WS_TJ = 8, WS_TK = 32, WS_RK = WS_TK + 2, WS_RN = (WS_TJ + 2) * WS_RK;
constexpr int WS_NL = (WS_RN + 255) / 256;  // staged values per thread per field

// CHK: the checkdt reductions of the same cells in the same pass (max |u|,
// |v|, |w| and min dtvisc with the vis read before it is overwritten), from
// the values staged for WALE, into part[q * RED_MAX_BLOCKS + slot].
struct ChkArgs {
  double dx, dy, dz, rsum;  // with the rdx, rdy, rdz of the WALE gradients
  double* part;
  int slot;
};
template <bool CHK = false>
__device__ inline void wale_stream_body(const double* __restrict__ u,
                                        const double* __restrict__ v,
                                        const double* __restrict__ w, double* __restrict__ vis,
                                        const wf_grid& g, const Box3& b, int ich, int ktile,
                                        int jtile, int chunk, double rdx, double rdy, double rdz,
                                        double rrey, double ls, ChkArgs ck = ChkArgs{}) {
  __shared__ double S[4][3][WS_RN];
  const int tid = threadIdx.x;
  const int tj = tid / WS_TK, tk = tid % WS_TK;
  const int j0 = jtile * WS_TJ, k0 = ktile * WS_TK;
  const int i0 = chunk * ich, i1 = min(i0 + ich, b.d0);
  if (i0 >= i1) {  // uniform per workgroup
    if (CHK && tid == 0) {
      ck.part[0 * RED_MAX_BLOCKS + ck.slot] = 0.0;
      ck.part[1 * RED_MAX_BLOCKS + ck.slot] = 0.0;
      ck.part[2 * RED_MAX_BLOCKS + ck.slot] = 0.0;
      ck.part[3 * RED_MAX_BLOCKS + ck.slot] = 1.0e10;
    }
    return;
  }
  double mu = 0.0, mv = 0.0, mw = 0.0, xmax = 0.0;
  const int nI = g.ni + 2 * g.npl, nJ = g.nj + 2 * g.npl, nK = g.nk + 2 * g.npl;
  int64_t off[WS_NL];
  int lds[WS_NL];
  bool on[WS_NL];
#pragma unroll
  for (int e = 0; e < WS_NL; ++e) {
    const int l = tid + 256 * e;
    on[e] = l < WS_RN;
    const int lc = on[e] ? l : 0;
    const int jf = min(max(b.lo1 + j0 + lc / WS_RK - 1, 0), nJ - 1);
    const int kf = min(max(b.lo2 + k0 + lc % WS_RK - 1, 0), nK - 1);
    off[e] = (int64_t)jf * g.sj + kf;
    lds[e] = lc;
  }
  auto plane = [&](int i) { return (int64_t)min(max(b.lo0 + i, 0), nI - 1) * g.si; };
  double ru[WS_NL], rv[WS_NL], rw[WS_NL];
  auto load = [&](int i) {
    const int64_t pb = plane(i);
#pragma unroll
    for (int e = 0; e < WS_NL; ++e) {
      ru[e] = u[pb + off[e]];
      rv[e] = v[pb + off[e]];
      rw[e] = w[pb + off[e]];
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int e = 0; e < WS_NL; ++e)
      if (on[e]) {
        S[slot][0][lds[e]] = ru[e];
        S[slot][1][lds[e]] = rv[e];
        S[slot][2][lds[e]] = rw[e];
      }
  };
  const bool live = j0 + tj < b.d1 && k0 + tk < b.d2;
  const int own = (tj + 1) * WS_RK + (tk + 1);
  // slot of plane p: (p - i0 + 1) mod 4
  load(i0 - 1);
  store(0);
  load(i0);
  store(1);
  load(i0 + 1);
  const int64_t xrow = (int64_t)(b.lo1 + j0 + tj) * g.sj + (b.lo2 + k0 + tk);
  // CHK reads each cell's vis before WALE overwrites it, one plane ahead and
  // from every thread (clamped into the box), so the load never sits in a
  // branch that the compiler closes with a wait for every load in flight --
  // the u, v, w planes prefetched two planes ahead included
  const int64_t xrow_ld = (int64_t)(b.lo1 + min(j0 + tj, b.d1 - 1)) * g.sj +
                          (b.lo2 + min(k0 + tk, b.d2 - 1));
  double vnext = 0.0;
  if (CHK) vnext = vis[plane(i0) + xrow_ld];
  auto step = [&](auto Rc, int i) {
    constexpr int R = decltype(Rc)::value;  // (i - i0) mod 4
    constexpr int SPv = R, S0v = (R + 1) % 4, SNv = (R + 2) % 4;
    store(SNv);  // plane i+1
    double vold = vnext;  // this step's vis of the cell, before WALE overwrites it
    // unconditional (the last step re-reads plane i1, an L2 hit): a load
    // under a branch makes the compiler wait for it where the paths join
    load(min(i + 2, i1));
    if (CHK) vnext = vis[plane(i + 1) + xrow_ld];
    __syncthreads();
    // the vis value is first used here, after the barrier: the prefetched
    // planes stay in flight through it
    if (CHK) asm volatile("" : "+v"(vold));
    if (live) {
      auto F = [&](int f, int di, int dj, int dk) {
        const int sl = di < 0 ? SPv : (di > 0 ? SNv : S0v);
        return S[sl][f][own + dj * WS_RK + dk];
      };
      auto U = [&](int di, int dj, int dk) { return F(0, di, dj, dk); };
      auto V = [&](int di, int dj, int dk) { return F(1, di, dj, dk); };
      auto W = [&](int di, int dj, int dk) { return F(2, di, dj, dk); };
      if constexpr (CHK) {  // checkdt_body's quantities of cell (i, j, k)
        mu = nan_max(mu, fabs(U(0, 0, 0)));
        mv = nan_max(mv, fabs(V(0, 0, 0)));
        mw = nan_max(mw, fabs(W(0, 0, 0)));
        xmax = nan_max(xmax, dtvisc_den(U(0, 0, 0), U(-1, 0, 0), V(0, 0, 0), V(0, -1, 0),
                                        W(0, 0, 0), W(0, 0, -1), vold, ck.dx, ck.dy, ck.dz,
                                        rdx, rdy, rdz, ck.rsum));
      }
      vis[plane(i) + xrow] = wale_value(U, V, W, rdx, rdy, rdz, rrey, ls);
    }
  };
  for (int t = 0; t < i1 - i0; t += 4) {
    step(std::integral_constant<int, 0>{}, i0 + t);
    if (t + 1 < i1 - i0) step(std::integral_constant<int, 1>{}, i0 + t + 1);
    if (t + 2 < i1 - i0) step(std::integral_constant<int, 2>{}, i0 + t + 2);
    if (t + 3 < i1 - i0) step(std::integral_constant<int, 3>{}, i0 + t + 3);
  }
  if constexpr (CHK) {
    __syncthreads();  // block_max reuses its own LDS; the planes are done with
    mu = block_max(mu);
    __syncthreads();
    mv = block_max(mv);
    __syncthreads();
    mw = block_max(mw);
    __syncthreads();
    xmax = block_max(xmax);
    if (tid == 0) {
      ck.part[0 * RED_MAX_BLOCKS + ck.slot] = mu;
      ck.part[1 * RED_MAX_BLOCKS + ck.slot] = mv;
      ck.part[2 * RED_MAX_BLOCKS + ck.slot] = mw;
      ck.part[3 * RED_MAX_BLOCKS + ck.slot] = dtvisc_of_max(xmax);
    }
  }
}

__global__ void __launch_bounds__(256)
wale_stream_kernel(const double* __restrict__ u, const double* __restrict__ v,
                   const double* __restrict__ w, double* __restrict__ vis, wf_grid g, Box3 b,
                   int ich, int kt, int jt, double rdx, double rdy, double rdz, double rrey,
                   double ls) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  wale_stream_body(u, v, w, vis, g, b, ich, L % kt, (L / kt) % jt, L / (kt * jt), rdx, rdy, rdz,
                   rrey, ls);
}

// WALE with the checkdt reductions of the same step in one pass.
__global__ void __launch_bounds__(256)
wale_checkdt_kernel(const double* __restrict__ u, const double* __restrict__ v,
                    const double* __restrict__ w, double* __restrict__ vis, wf_grid g, Box3 b,
                    int ich, int kt, int jt, double rdx, double rdy, double rdz, double rrey,
                    double ls, double dx, double dy, double dz, double rsum) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  wale_stream_body<true>(u, v, w, vis, g, b, ich, L % kt, (L / kt) % jt, L / (kt * jt), rdx,
                         rdy, rdz, rrey, ls,
                         ChkArgs{dx, dy, dz, rsum, g_red_part,
                                 (int)(blockIdx.x + blockIdx.y * gridDim.x)});
}

__global__ void __launch_bounds__(256)
maxabs_kernel(const double* __restrict__ f, wf_grid g, Box3 b, int ich) {
  int j, k, ibeg, iend;
  double val = 0.0;
  if (red_rows(b, ich, j, k, ibeg, iend))
    for (int ii = ibeg; ii < iend; ++ii)
      val = nan_max(val, fabs(f[at(g, b.lo0 + ii, b.lo1 + j, b.lo2 + k)]));
  const double r = block_max(val);
  if (block_leader()) g_red_part[red_block()] = r;
}

// checkdt.for dt-limit denominator at one cell (kernels.py:394-401); rsum =
// 1/dx^2 + 1/dy^2 + 1/dz^2 computed once by the host in the reference's order.
__device__ inline double dtvisc_cell(const double* __restrict__ u,
                                     const double* __restrict__ v,
                                     const double* __restrict__ w,
                                     const double* __restrict__ vis, int64_t x,
                                     const wf_grid& g, double dx, double dy, double dz,
                                     double rdx, double rdy, double rdz, double rsum) {
  return dtvisc_den(u[x], u[x - g.si], v[x], v[x - g.sj], w[x], w[x - 1], vis[x], dx, dy, dz,
                    rdx, rdy, rdz, rsum);
}

__global__ void __launch_bounds__(256)
dtvisc_kernel(const double* __restrict__ u, const double* __restrict__ v,
              const double* __restrict__ w, const double* __restrict__ vis, wf_grid g,
              Box3 b, int ich, double dx, double dy, double dz, double rsum) {
  const double rdx = 1.0 / dx, rdy = 1.0 / dy, rdz = 1.0 / dz;
  int j, k, ibeg, iend;
  double xmax = 0.0;
  if (red_rows(b, ich, j, k, ibeg, iend))
    for (int ii = ibeg; ii < iend; ++ii)
      xmax = nan_max(xmax, dtvisc_cell(u, v, w, vis, at(g, b.lo0 + ii, b.lo1 + j, b.lo2 + k),
                                       g, dx, dy, dz, rdx, rdy, rdz, rsum));
  const double r = dtvisc_of_max(block_max(xmax));
  if (block_leader()) g_red_part[red_block()] = r;
}

// umax, vmax, wmax, min dt of one workgroup's rows into part[q * stride + slot].
__device__ inline void checkdt_body(const double* __restrict__ u, const double* __restrict__ v,
                                    const double* __restrict__ w,
                                    const double* __restrict__ vis, const wf_grid& g,
                                    const Box3& b, Blk B, int ich, double dx, double dy,
                                    double dz, double rsum, double* part, int stride, int slot) {
  const double rdx = 1.0 / dx, rdy = 1.0 / dy, rdz = 1.0 / dz;
  int j, k, ibeg, iend;
  double mu = 0.0, mv = 0.0, mw = 0.0, xmax = 0.0;
  if (red_rows(b, B, ich, j, k, ibeg, iend)) {
    for (int ii = ibeg; ii < iend; ++ii) {
      const int64_t x = at(g, b.lo0 + ii, b.lo1 + j, b.lo2 + k);
      mu = nan_max(mu, fabs(u[x]));
      mv = nan_max(mv, fabs(v[x]));
      mw = nan_max(mw, fabs(w[x]));
      xmax = nan_max(xmax, dtvisc_cell(u, v, w, vis, x, g, dx, dy, dz, rdx, rdy, rdz, rsum));
    }
  }
  mu = block_max(mu);
  __syncthreads();
  mv = block_max(mv);
  __syncthreads();
  mw = block_max(mw);
  __syncthreads();
  xmax = block_max(xmax);
  if (block_leader()) {
    part[0 * stride + slot] = mu;
    part[1 * stride + slot] = mv;
    part[2 * stride + slot] = mw;
    part[3 * stride + slot] = dtvisc_of_max(xmax);
  }
}

__global__ void __launch_bounds__(256)
checkdt_kernel(const double* __restrict__ u, const double* __restrict__ v,
               const double* __restrict__ w, const double* __restrict__ vis, wf_grid g,
               Box3 b, int ich, double dx, double dy, double dz, double rsum) {
  checkdt_body(u, v, w, vis, g, b, this_blk(), ich, dx, dy, dz, rsum, g_red_part,
               RED_MAX_BLOCKS, red_block());
}

__device__ inline void add_body(double* __restrict__ dst, const double* __restrict__ src,
                                const wf_grid& g, const Box3& b, Blk B) {
  int i, j, k;
  if (!box_thread(b, B, i, j, k)) return;
  const int64_t x = at(g, i, j, k);
  dst[x] = dst[x] + src[x];
}
// dst += src - *shift on the real cells and the faces of the first ghost
// layer, dst += src on its edges and corners: p += pp with the mixed
// solve's pressure-gauge shift of pp applied on the fly (the solve leaves its
// edges and corners zero and mirrors the shifted cells onto the faces, so
// this equals shifting pp and then adding, bit for bit).
__device__ inline void add_shift_body(double* __restrict__ dst, const double* __restrict__ src,
                                      const double* __restrict__ shift, const wf_grid& g,
                                      const Box3& b, Blk B) {
  int i, j, k;
  if (!box_thread(b, B, i, j, k)) return;
  const int64_t x = at(g, i, j, k);
  const int lo = g.npl - 1;
  const int nring = (i == lo || i == g.npl + g.ni) + (j == lo || j == g.npl + g.nj) +
                    (k == lo || k == g.npl + g.nk);
  const double v = src[x];
  dst[x] = dst[x] + (nring <= 1 ? v - *shift : v);
}
__global__ void __launch_bounds__(256)
add_shift_kernel(double* __restrict__ dst, const double* __restrict__ src,
                 const double* __restrict__ shift, wf_grid g, Box3 b, int ich) {
  for_planes(b.d0, ich, [&](Blk B) { add_shift_body(dst, src, shift, g, b, B); });
}

__global__ void __launch_bounds__(256)
add_kernel(double* __restrict__ dst, const double* __restrict__ src, wf_grid g, Box3 b, int ich) {
  for_planes(b.d0, ich, [&](Blk B) { add_body(dst, src, g, b, B); });
}

__device__ inline void fill_box_body(double* __restrict__ f, double val, const wf_grid& g,
                                     const Box3& b, Blk B) {
  int i, j, k;
  if (!box_thread(b, B, i, j, k)) return;
  f[at(g, i, j, k)] = val;
}
__global__ void __launch_bounds__(256)
fill_box_kernel(double* __restrict__ f, double val, wf_grid g, Box3 b, int ich) {
  for_planes(b.d0, ich, [&](Blk B) { fill_box_body(f, val, g, b, B); });
}

// kernels.py:460-471; uint32 wrap-around arithmetic on the full-array indices.
__device__ inline void noise_body(double* __restrict__ u, double* __restrict__ v,
                                  double* __restrict__ w, double amp, uint32_t seed,
                                  const wf_grid& g, const Box3& b, Blk B) {
  int i, j, k;
  if (!box_thread(b, B, i, j, k)) return;
  uint32_t h = ((uint32_t)i * 73856093u) ^ ((uint32_t)j * 19349663u) ^
               ((uint32_t)k * 83492791u) ^ seed;
  h = h ^ (h >> 13);
  h = h * 1274126177u;
  h = h ^ (h >> 16);
  const double r = ((double)h / 4294967295.0) * 2.0 - 1.0;
  const int64_t x = at(g, i, j, k);
  u[x] = amp * r;
  v[x] = amp * r;
  w[x] = -amp * r;
}
__global__ void __launch_bounds__(256)
noise_kernel(double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
             double amp, uint32_t seed, wf_grid g, Box3 b, int ich) {
  for_planes(b.d0, ich, [&](Blk B) { noise_body(u, v, w, amp, seed, g, b, B); });
}

// ---------------------------------------------------------------------------
// Batched (ensemble) kernels: a persistent grid walks the work items of a
// plan; each item is one block of one member, run through the same body as
// the single-domain kernel on that member's arrays.  Members whose `active`
// flag is 0 are skipped.  SYNC: a barrier after every item, for bodies that
// use LDS.
// ---------------------------------------------------------------------------
template <bool SYNC, class F>
__device__ inline void for_items(const Plan& P, const BView& V, const int* active, F&& f) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    if (active && !active[m]) continue;
    f(m, it - P.start[m], V.mem[m]);
    if (SYNC) __syncthreads();
  }
}

__device__ inline Box3 cells_box(const wf_grid& g) {
  return Box3{g.npl, g.npl, g.npl, g.ni, g.nj, g.nk};
}

__global__ void __launch_bounds__(256)
batch_noise_kernel(BView V, Plan P, double* u, double* v, double* w) {
  for_items<false>(P, V, nullptr, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    noise_body(u + M.off, v + M.off, w + M.off, M.noise, M.seed, M.g, b,
               cells_blk(q, b.d1, b.d2));
  });
}

__global__ void __launch_bounds__(256) batch_fill_vis_kernel(BView V, Plan P, double* vis) {
  for_items<false>(P, V, nullptr, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    fill_box_body(vis + M.off, M.rey, M.g, b, cells_blk(q, b.d1, b.d2));
  });
}

__global__ void __launch_bounds__(256)
batch_wale_kernel(BView V, Plan P, const double* u, const double* v, const double* w,
                  double* vis) {
  for_items<true>(P, V, nullptr, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    const int kt = (b.d2 + WS_TK - 1) / WS_TK, jt = (b.d1 + WS_TJ - 1) / WS_TJ;
    wale_stream_body(u + M.off, v + M.off, w + M.off, vis + M.off, M.g, b, BATCH_CD_ICH, q % kt,
                     (q / kt) % jt, q / (kt * jt), 1.0 / M.dx, 1.0 / M.dy, 1.0 / M.dz, M.rey,
                     M.ls);
  });
}

__global__ void __launch_bounds__(256)
batch_calvel3_kernel(BView V, Plan P, const double* us, const double* vs, const double* ws,
                     const double* p, double* u, double* v, double* w, double alpha,
                     const int* active) {
  for_items<false>(P, V, active, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    calvel3_body(us + M.off, vs + M.off, ws + M.off, p + M.off, u + M.off, v + M.off,
                 w + M.off, M.g, b, cells_blk(q, b.d1, b.d2), 1.0 / M.dx, 1.0 / M.dy, 1.0 / M.dz,
                 M.dt * alpha);
  });
}

__global__ void __launch_bounds__(256)
batch_add_ring_kernel(BView V, Plan P, double* dst, const double* src, const int* active) {
  for_items<false>(P, V, active, [&](int, int q, const MemDev& M) {
    const wf_grid& g = M.g;
    const Box3 b{g.npl - 1, g.npl - 1, g.npl - 1, g.ni + 2, g.nj + 2, g.nk + 2};
    add_body(dst + M.off, src + M.off, g, b, cells_blk(q, b.d1, b.d2));
  });
}

__global__ void __launch_bounds__(256)
batch_calmas_kernel(BView V, Plan P, const double* u, const double* v, const double* w,
                    double* su, double* part, const int* active) {
  for_items<true>(P, V, active, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    calmas_body(u + M.off, v + M.off, w + M.off, su + M.off, M.g, b, cells_blk(q, b.d1, b.d2),
                BATCH_RED_ICH, M.dx, M.dy, M.dz, M.dt, part, M.red0 + q);
  });
}

// The fused calvel/calmas pass per member on the reduction items (64 x 4
// (k, j) tiles x BATCH_RED_ICH planes), so its partials fill the slots
// batch_red_fold_kernel folds, as batch_calmas_kernel's do.
__global__ void __launch_bounds__(256)
batch_calvel3_calmas_kernel(BView V, Plan P, const double* us, const double* vs,
                            const double* ws, const double* p, double* u, double* v, double* w,
                            double* su, double* part, double alpha, const int* active) {
  for_items<true>(P, V, active, [&](int, int q, const MemDev& M) {
    const wf_grid& g = M.g;
    const int kt = (g.nk + 63) / 64, jt = (g.nj + 3) / 4;
    const pj::Consts K{1.0 / M.dx, 1.0 / M.dy, 1.0 / M.dz, M.dt * alpha,
                       1.0 / (24.0 * M.dx), 1.0 / (24.0 * M.dy), 1.0 / (24.0 * M.dz),
                       1.0 / M.dt, M.dx, M.dy, M.dz};
    const double r = calvel3_calmas_body<4, 64>(
        us + M.off, vs + M.off, ws + M.off, p + M.off, u + M.off, v + M.off, w + M.off,
        su + M.off, g, BATCH_RED_ICH, q % kt, (q / kt) % jt, q / (kt * jt), K);
    if (threadIdx.x == 0) part[M.red0 + q] = r;
  });
}

__global__ void __launch_bounds__(256)
batch_checkdt_kernel(BView V, Plan P, const double* u, const double* v, const double* w,
                     const double* vis, double* part, int stride) {
  for_items<true>(P, V, nullptr, [&](int, int q, const MemDev& M) {
    const Box3 b = cells_box(M.g);
    checkdt_body(u + M.off, v + M.off, w + M.off, vis + M.off, M.g, b,
                 cells_blk(q, b.d1, b.d2), BATCH_RED_ICH, M.dx, M.dy, M.dz, M.rsum, part,
                 stride, M.red0 + q);
  });
}

// One workgroup per member folds its reduction partials: out[m*nq + q] for
// quantity q (bit q of min_mask: min, else max).
__global__ void __launch_bounds__(1024)
batch_red_fold_kernel(BView V, const double* part, int stride, int nq, int min_mask,
                      double* out, const int* active) {
  const int m = blockIdx.x;
  if (active && !active[m]) return;  // uniform per workgroup
  const MemDev& M = V.mem[m];
  const int n = red_blocks(M.g.ni, M.g.nj, M.g.nk);
  __shared__ double sh[16];
  const int t = threadIdx.x;
  for (int q = 0; q < nq; ++q) {
    const bool mn = (min_mask >> q) & 1;
    double val = mn ? 1.0e10 : 0.0;
    for (int p = t; p < n; p += 1024) {
      const double y = part[q * stride + M.red0 + p];
      val = mn ? nan_min(val, y) : nan_max(val, y);
    }
    val = mn ? wave_min(val) : wave_max(val);
    if ((t & 63) == 0) sh[t >> 6] = val;
    __syncthreads();
    if (t == 0) {
      double r = sh[0];
      for (int w = 1; w < 16; ++w) r = mn ? nan_min(r, sh[w]) : nan_max(r, sh[w]);
      out[m * nq + q] = r;
    }
    __syncthreads();
  }
}

int batch_noise(const Batch& B, double* u, double* v, double* w, hipStream_t s) {
  const Plan& P = B.plan[PK_CELLS];
  hipLaunchKernelGGL(batch_noise_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, u, v, w);
  WF_LAUNCHED();
  return 0;
}

int batch_fill_vis(const Batch& B, double* vis, hipStream_t s) {
  const Plan& P = B.plan[PK_CELLS];
  hipLaunchKernelGGL(batch_fill_vis_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, vis);
  WF_LAUNCHED();
  return 0;
}

int batch_wale(const Batch& B, const double* u, const double* v, const double* w, double* vis,
               hipStream_t s) {
  const Plan& P = B.plan[PK_WALE];
  hipLaunchKernelGGL(batch_wale_kernel, dim3(P.grid), dim3(256), 0, s, B.view(), P, u, v, w,
                     vis);
  WF_LAUNCHED();
  return 0;
}

int batch_calvel3(const Batch& B, const double* us, const double* vs, const double* ws,
                  const double* p, double* u, double* v, double* w, double alpha,
                  const int* active, hipStream_t s) {
  const Plan& P = B.plan[PK_CELLS];
  hipLaunchKernelGGL(batch_calvel3_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, us, vs,
                     ws, p, u, v, w, alpha, active);
  WF_LAUNCHED();
  return 0;
}

int batch_add_ring(const Batch& B, double* p, const double* pp, const int* active,
                   hipStream_t s) {
  const Plan& P = B.plan[PK_RING];
  hipLaunchKernelGGL(batch_add_ring_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, p, pp,
                     active);
  WF_LAUNCHED();
  return 0;
}

int batch_calmas(const Batch& B, const double* u, const double* v, const double* w, double* su,
                 double* rmax, const int* active, hipStream_t s) {
  const Plan& P = B.plan[PK_RED];
  hipLaunchKernelGGL(batch_calmas_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, u, v, w,
                     su, B.red_part, active);
  WF_LAUNCHED();
  hipLaunchKernelGGL(batch_red_fold_kernel, dim3(B.nmem), dim3(1024), 0, s, B.view(),
                     (const double*)B.red_part, B.nred, 1, 0, rmax, active);
  WF_LAUNCHED();
  return 0;
}

int batch_calvel3_calmas(const Batch& B, const double* us, const double* vs, const double* ws,
                         const double* p, double* u, double* v, double* w, double* su,
                         double* rmax, double alpha, const int* active, hipStream_t s) {
  const Plan& P = B.plan[PK_RED];
  hipLaunchKernelGGL(batch_calvel3_calmas_kernel, dim3(P.grid), dim3(256), 0, s, B.view(), P, us,
                     vs, ws, p, u, v, w, su, B.red_part, alpha, active);
  WF_LAUNCHED();
  hipLaunchKernelGGL(batch_red_fold_kernel, dim3(B.nmem), dim3(1024), 0, s, B.view(),
                     (const double*)B.red_part, B.nred, 1, 0, rmax, active);
  WF_LAUNCHED();
  return 0;
}

int batch_checkdt(const Batch& B, const double* u, const double* v, const double* w,
                  const double* vis, double* out4, hipStream_t s) {
  const Plan& P = B.plan[PK_RED];
  hipLaunchKernelGGL(batch_checkdt_kernel, dim3(P.grid), box_block(), 0, s, B.view(), P, u, v, w,
                     vis, B.red_part, B.nred);
  WF_LAUNCHED();
  hipLaunchKernelGGL(batch_red_fold_kernel, dim3(B.nmem), dim3(1024), 0, s, B.view(),
                     (const double*)B.red_part, B.nred, 4, 8, out4, (const int*)nullptr);
  WF_LAUNCHED();
  return 0;
}

}  // namespace wf

using namespace wf;

// ---------------------------------------------------------------------------
// Host-side argument checks: the launch box plus the stencil reach must stay
// inside the allocated (ni+2npl, nj+2npl, nk+2npl) array.
// ---------------------------------------------------------------------------
static bool grid_ok(const wf_grid& g) {
  return g.ni > 0 && g.nj > 0 && g.nk > 0 && g.npl >= 0 &&
         g.sj >= (int64_t)(g.nk + 2 * g.npl) &&
         g.si >= g.sj * (int64_t)(g.nj + 2 * g.npl);
}
// Box [lo, lo+d) widened by `reach` must lie in [0, n+2npl) on every axis.
static bool box_ok(const wf_grid& g, const Box3& b, int reach_lo, int reach_hi) {
  if (!grid_ok(g) || b.d0 < 0 || b.d1 < 0 || b.d2 < 0) return false;
  const int lo[3] = {b.lo0, b.lo1, b.lo2};
  const int d[3] = {b.d0, b.d1, b.d2};
  const int n[3] = {g.ni, g.nj, g.nk};
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0) continue;
    if (lo[a] - reach_lo < 0) return false;
    if (lo[a] + d[a] - 1 + reach_hi > n[a] + 2 * g.npl - 1) return false;
  }
  return true;
}
struct RedGeom {
  dim3 grid;
  int ich, nblk;
};
// Reduction launch shape: (k, j) tiles x i-chunks, at most RED_MAX_BLOCKS
// workgroups; false when a single i-plane already needs more tiles.
static bool red_geom(const Box3& b, RedGeom& r) {
  const int kt = div_up(b.d2, BX), jt = div_up(b.d1, BY);
  const int64_t tiles = (int64_t)kt * jt;
  if (tiles > RED_MAX_BLOCKS || b.d0 <= 0) return false;
  const int ic_max = (int)(RED_MAX_BLOCKS / tiles);
  r.ich = div_up(b.d0, ic_max);
  const int ic = div_up(b.d0, r.ich);
  r.grid = dim3(kt, jt, ic);
  r.nblk = (int)(tiles * ic);
  return true;
}
static int fold(double* out, const RedGeom& r, int nq, int min_mask, int combine,
                hipStream_t s) {
  hipLaunchKernelGGL(red_fold_kernel, dim3(1), dim3(1024), 0, s, out, r.nblk, nq, min_mask,
                     combine);
  WF_LAUNCHED();
  return 0;
}

static Box3 npl_box(const wf_grid& g, int d0, int d1, int d2) {
  return Box3{g.npl, g.npl, g.npl, d0, d1, d2};
}
static bool empty(const Box3& b) { return b.d0 == 0 || b.d1 == 0 || b.d2 == 0; }
#define S_(s) ((hipStream_t)(s))
#define CH(b) plane_chunk((b).d0, (b).d1, (b).d2)
#define CG(b) chunk_grid((b).d0, (b).d1, (b).d2, CH(b))

extern "C" {

int wf_conv4th(int comp, const double* u, const double* v, const double* w,
               const double* coo, double* cstar, wf_grid g, int d0, int d1, int d2,
               double dx, double dy, double dz, double dt, double alpha, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (comp < 0 || comp > 2 || !box_ok(g, b, 3, 3)) return WF_EINVAL;
  if (empty(b)) return 0;
  const double dta = dt * alpha;
  if (comp == 0)
    hipLaunchKernelGGL(conv4th_kernel<0>, CG(b), box_block(), 0, S_(stream), u, v, w,
                       coo, cstar, g, b, dx, dy, dz, dta, CH(b));
  else if (comp == 1)
    hipLaunchKernelGGL(conv4th_kernel<1>, CG(b), box_block(), 0, S_(stream), u, v, w,
                       coo, cstar, g, b, dx, dy, dz, dta, CH(b));
  else
    hipLaunchKernelGGL(conv4th_kernel<2>, CG(b), box_block(), 0, S_(stream), u, v, w,
                       coo, cstar, g, b, dx, dy, dz, dta, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_diff4th(int comp, const double* c, const double* vis, double* cstar, wf_grid g,
               int d0, int d1, int d2, double dx, double dy, double dz, double dt,
               double alpha, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (comp < 0 || comp > 2 || !box_ok(g, b, 2, 2)) return WF_EINVAL;
  if (empty(b)) return 0;
  const double dta = dt * alpha;
  const double rx = 12.0 * dx * dx, ry = 12.0 * dy * dy, rz = 12.0 * dz * dz;
  if (comp == 0)
    hipLaunchKernelGGL(diff4th_kernel<0>, CG(b), box_block(), 0, S_(stream), c, vis,
                       cstar, g, b, rx, ry, rz, dta, CH(b));
  else if (comp == 1)
    hipLaunchKernelGGL(diff4th_kernel<1>, CG(b), box_block(), 0, S_(stream), c, vis,
                       cstar, g, b, rx, ry, rz, dta, CH(b));
  else
    hipLaunchKernelGGL(diff4th_kernel<2>, CG(b), box_block(), 0, S_(stream), c, vis,
                       cstar, g, b, rx, ry, rz, dta, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_convdiff4th(int comp, const double* u, const double* v, const double* w,
                   const double* coo, const double* vis, double* cstar, wf_grid g, int d0,
                   int d1, int d2, double dx, double dy, double dz, double dt, double alpha,
                   void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (comp < 0 || comp > 2 || !box_ok(g, b, 3, 3)) return WF_EINVAL;
  if (empty(b)) return 0;
  return convdiff_launch(comp, u, v, w, coo, vis, cstar, g, b, dx, dy, dz, dt * alpha, 0,
                         S_(stream));
}

int wf_convdiff4th_save(int comp, const double* u, const double* v, const double* w,
                        double* coo_save, const double* vis, double* cstar, wf_grid g, int d0,
                        int d1, int d2, double dx, double dy, double dz, double dt,
                        double alpha, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (comp < 0 || comp > 2 || !box_ok(g, b, 3, 3) || !coo_save) return WF_EINVAL;
  if (empty(b)) return 0;
  return convdiff_launch(comp, u, v, w, nullptr, vis, cstar, g, b, dx, dy, dz, dt * alpha, 0,
                         S_(stream), coo_save);
}

int wf_convdiff3_save(const double* u, const double* v, const double* w, const double* vis,
                      double* ustar, double* vstar, double* wstar, double* uoo, double* voo,
                      double* woo, wf_grid g, double dx, double dy, double dz, double dt,
                      double alpha, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 3, 3) || !u || !v || !w || !vis || !ustar || !vstar || !wstar || !uoo ||
      !voo || !woo)
    return WF_EINVAL;
  return convdiff3_save_launch(u, v, w, vis, ustar, vstar, wstar, uoo, voo, woo, g, dx, dy, dz,
                               dt * alpha, 0, S_(stream));
}

int wf_convdiff_vw(const double* u, const double* v, const double* w, const double* voo,
                   const double* woo, const double* vis, const double* w_pre, double* vstar,
                   double* wstar, wf_grid g, double dx, double dy, double dz, double dt,
                   double alpha, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 3, 3) || !u || !v || !w || !voo || !woo || !vis || !w_pre || !vstar ||
      !wstar)
    return WF_EINVAL;
  return convdiff_vw_launch(u, v, w, voo, woo, vis, w_pre, vstar, wstar, g, dx, dy, dz,
                            dt * alpha, 0, S_(stream));
}

int wf_convdiff3(const double* u, const double* v, const double* w, const double* uoo,
                 const double* voo, const double* woo, const double* vis, const double* w_pre,
                 double* ustar, double* vstar, double* wstar, wf_grid g, double dx, double dy,
                 double dz, double dt, double alpha, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 3, 3) || !u || !v || !w || !uoo || !voo || !woo || !vis || !w_pre ||
      !ustar || !vstar || !wstar)
    return WF_EINVAL;
  return convdiff3_launch(u, v, w, uoo, voo, woo, vis, w_pre, ustar, vstar, wstar, g, dx, dy,
                          dz, dt * alpha, 0, S_(stream));
}

int wf_convdiff4th_u_edges(const double* u, const double* v, const double* w,
                           const double* uoo, const double* vis, double* ustar, wf_grid g,
                           double dx, double dy, double dz, double dt, double alpha,
                           void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 3, 3) || !u || !v || !w || !uoo || !vis || !ustar) return WF_EINVAL;
  return convdiff_u_edges_launch(u, v, w, uoo, vis, ustar, g, dx, dy, dz, dt * alpha,
                                 S_(stream));
}

int wf_save_cross_ghosts(const double* w, double* w_pre, wf_grid g, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 2, 2) || !w || !w_pre) return WF_EINVAL;
  return save_cross_ghosts_launch(w, w_pre, g, S_(stream));
}

int wf_calvel(int comp, const double* cstar, const double* p, double* c, wf_grid g, int d0,
              int d1, int d2, double h, double dt, double alpha, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (comp < 0 || comp > 2 || !box_ok(g, b, 0, 1)) return WF_EINVAL;
  if (empty(b)) return 0;
  const int64_t s = comp == 0 ? g.si : (comp == 1 ? g.sj : 1);
  hipLaunchKernelGGL(calvel_kernel, CG(b), box_block(), 0, S_(stream), cstar, p, c, g,
                     b, s, 1.0 / h, dt * alpha, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_calvel3(const double* ustar, const double* vstar, const double* wstar,
               const double* p, double* u, double* v, double* w, wf_grid g, double dx,
               double dy, double dz, double dt, double alpha, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 0, 1)) return WF_EINVAL;
  hipLaunchKernelGGL(calvel3_kernel, CG(b), box_block(), 0, S_(stream), ustar, vstar,
                     wstar, p, u, v, w, g, b, 1.0 / dx, 1.0 / dy, 1.0 / dz, dt * alpha, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_calmas(const double* u, const double* v, const double* w, double* su,
              double* rmax_dev, wf_grid g, int d0, int d1, int d2, double dx, double dy,
              double dz, double dt, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  RedGeom r;
  if (!box_ok(g, b, 2, 1) || !rmax_dev) return WF_EINVAL;
  if (empty(b)) return 0;
  if (!red_geom(b, r)) return WF_EINVAL;
  hipLaunchKernelGGL(calmas_kernel, r.grid, box_block(), 0, S_(stream), u, v, w, su, g, b,
                     r.ich, dx, dy, dz, dt);
  WF_LAUNCHED();
  return fold(rmax_dev, r, 1, 0, 1, S_(stream));
}

static int calvel3_calmas_launch(const double* ustar, const double* vstar, const double* wstar,
                                 const double* p, double* u, double* v, double* w, double* su,
                                 double* rmax_dev, float* rhs32, wf_grid g, double dx,
                                 double dy, double dz, double dt, double alpha, void* stream,
                                 bool store) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 2, 1) || !rmax_dev) return WF_EINVAL;
  if (empty(b)) return 0;
  const int kt = div_up(g.nk, PJ_TK), jt = div_up(g.nj, PJ_TJ);
  const int64_t tiles = (int64_t)kt * jt;
  if (tiles > RED_MAX_BLOCKS) {  // more tiles than partial slots: the two passes
    if (rhs32) return WF_EINVAL;  // the two-pass form has no FP32 rhs output
    const int e = wf_calvel3(ustar, vstar, wstar, p, u, v, w, g, dx, dy, dz, dt, alpha, stream);
    if (e) return e;
    return wf_calmas(u, v, w, su, rmax_dev, g, g.ni, g.nj, g.nk, dx, dy, dz, dt, stream);
  }
  // built-in: >= 2048 workgroups where the grid allows; at most RED_MAX_BLOCKS
  int ich = tune_get(WF_TUNE_PROJ_ICH, g.ni, g.nj, g.nk);
  if (ich <= 0) ich = (int)std::max<int64_t>(8, (tiles * g.ni + 2047) / 2048);
  ich = std::max(ich, div_up(g.ni, (int)(RED_MAX_BLOCKS / tiles)));
  const int nch = div_up(g.ni, ich);
  RedGeom r;
  r.grid = dim3((unsigned)tiles, (unsigned)nch);
  r.ich = ich;
  r.nblk = (int)(tiles * nch);
  const pj::Consts K{1.0 / dx, 1.0 / dy, 1.0 / dz, dt * alpha,
                     1.0 / (24.0 * dx), 1.0 / (24.0 * dy), 1.0 / (24.0 * dz),
                     1.0 / dt, dx, dy, dz};
  if (store)
    hipLaunchKernelGGL(calvel3_calmas_kernel<true>, r.grid, dim3(PJ_TJ * PJ_TK), 0, S_(stream),
                       ustar, vstar, wstar, p, u, v, w, su, g, ich, kt, jt, K, rhs32);
  else
    hipLaunchKernelGGL(calvel3_calmas_kernel<false>, r.grid, dim3(PJ_TJ * PJ_TK), 0, S_(stream),
                       ustar, vstar, wstar, p, u, v, w, su, g, ich, kt, jt, K, rhs32);
  WF_LAUNCHED();
  return fold(rmax_dev, r, 1, 0, 1, S_(stream));
}

int wf_calvel3_calmas(const double* ustar, const double* vstar, const double* wstar,
                      const double* p, double* u, double* v, double* w, double* su,
                      double* rmax_dev, float* rhs32, wf_grid g, double dx, double dy,
                      double dz, double dt, double alpha, void* stream) {
  return calvel3_calmas_launch(ustar, vstar, wstar, p, u, v, w, su, rmax_dev, rhs32, g, dx, dy,
                               dz, dt, alpha, stream, true);
}

int wf_calmas_of_calvel3(const double* ustar, const double* vstar, const double* wstar,
                         const double* p, const double* u, const double* v, const double* w,
                         double* su, double* rmax_dev, float* rhs32, wf_grid g, double dx,
                         double dy, double dz, double dt, double alpha, void* stream) {
  // the kernel only reads u, v, w when STORE is false
  return calvel3_calmas_launch(ustar, vstar, wstar, p, const_cast<double*>(u),
                               const_cast<double*>(v), const_cast<double*>(w), su, rmax_dev,
                               rhs32, g, dx, dy, dz, dt, alpha, stream, false);
}

int wf_wale(const double* u, const double* v, const double* w, double* vis, wf_grid g,
            int d0, int d1, int d2, double dx, double dy, double dz, double rrey,
            double cw, double delta, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (!box_ok(g, b, 1, 1)) return WF_EINVAL;
  if (empty(b)) return 0;
  const int kt = div_up(b.d2, WS_TK), jt = div_up(b.d1, WS_TJ);
  const int ich = (int)std::max<int64_t>(16, ((int64_t)kt * jt * b.d0 + 2047) / 2048);
  hipLaunchKernelGGL(wale_stream_kernel, dim3(kt * jt, div_up(b.d0, ich)), dim3(256), 0,
                     S_(stream), u, v, w, vis, g, b, ich, kt, jt, 1.0 / dx, 1.0 / dy, 1.0 / dz,
                     rrey, cw * delta);
  WF_LAUNCHED();
  return 0;
}

int wf_maxabs(const double* f, double* out_dev, wf_grid g, int d0, int d1, int d2,
              void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  RedGeom r;
  if (!box_ok(g, b, 0, 0) || !out_dev) return WF_EINVAL;
  if (empty(b)) return 0;
  if (!red_geom(b, r)) return WF_EINVAL;
  hipLaunchKernelGGL(maxabs_kernel, r.grid, box_block(), 0, S_(stream), f, g, b, r.ich);
  WF_LAUNCHED();
  return fold(out_dev, r, 1, 0, 1, S_(stream));
}

static double inv_sq_sum(double dx, double dy, double dz) {
  return 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
}

int wf_dtvisc(const double* u, const double* v, const double* w, const double* vis,
              double* out_dev, wf_grid g, int d0, int d1, int d2, double dx, double dy,
              double dz, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  RedGeom r;
  if (!box_ok(g, b, 1, 0) || !out_dev) return WF_EINVAL;
  if (empty(b)) return 0;
  if (!red_geom(b, r)) return WF_EINVAL;
  hipLaunchKernelGGL(dtvisc_kernel, r.grid, box_block(), 0, S_(stream), u, v, w, vis, g, b,
                     r.ich, dx, dy, dz, inv_sq_sum(dx, dy, dz));
  WF_LAUNCHED();
  return fold(out_dev, r, 1, 1, 1, S_(stream));
}

int wf_checkdt(const double* u, const double* v, const double* w, const double* vis,
               double* out4_dev, wf_grid g, double dx, double dy, double dz, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  RedGeom r;
  if (!box_ok(g, b, 1, 0) || !out4_dev || !red_geom(b, r)) return WF_EINVAL;
  hipLaunchKernelGGL(checkdt_kernel, r.grid, box_block(), 0, S_(stream), u, v, w, vis, g, b,
                     r.ich, dx, dy, dz, inv_sq_sum(dx, dy, dz));
  WF_LAUNCHED();
  return fold(out4_dev, r, 4, 8, 0, S_(stream));
}

int wf_wale_checkdt(const double* u, const double* v, const double* w, double* vis,
                    double* out4_dev, wf_grid g, double dx, double dy, double dz, double rrey,
                    double cw, double delta, void* stream) {
  const Box3 b = npl_box(g, g.ni, g.nj, g.nk);
  if (!box_ok(g, b, 1, 1) || !out4_dev) return WF_EINVAL;
  if (empty(b)) return 0;
  const int kt = div_up(b.d2, WS_TK), jt = div_up(b.d1, WS_TJ);
  const int64_t tiles = (int64_t)kt * jt;
  int ich = (int)std::max<int64_t>(16, (tiles * b.d0 + 2047) / 2048);
  if (tiles > RED_MAX_BLOCKS) {  // more tiles than partial slots: the two passes
    const int e = wf_checkdt(u, v, w, vis, out4_dev, g, dx, dy, dz, stream);
    if (e) return e;
    return wf_wale(u, v, w, vis, g, g.ni, g.nj, g.nk, dx, dy, dz, rrey, cw, delta, stream);
  }
  ich = std::max(ich, div_up(b.d0, (int)(RED_MAX_BLOCKS / tiles)));
  RedGeom r;
  r.grid = dim3((unsigned)tiles, (unsigned)div_up(b.d0, ich));
  r.ich = ich;
  r.nblk = (int)(tiles * r.grid.y);
  hipLaunchKernelGGL(wale_checkdt_kernel, r.grid, dim3(256), 0, S_(stream), u, v, w, vis, g, b,
                     ich, kt, jt, 1.0 / dx, 1.0 / dy, 1.0 / dz, rrey, cw * delta, dx, dy, dz,
                     inv_sq_sum(dx, dy, dz));
  WF_LAUNCHED();
  return fold(out4_dev, r, 4, 8, 0, S_(stream));
}

int wf_add_ring(double* dst, const double* src, wf_grid g, void* stream) {
  if (g.npl < 1) return WF_EINVAL;
  const Box3 b{g.npl - 1, g.npl - 1, g.npl - 1, g.ni + 2, g.nj + 2, g.nk + 2};
  if (!box_ok(g, b, 0, 0)) return WF_EINVAL;
  hipLaunchKernelGGL(add_kernel, CG(b), box_block(), 0, S_(stream), dst, src, g, b, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_add_ring_shift(double* dst, const double* src, const double* shift_dev, wf_grid g,
                      void* stream) {
  if (g.npl < 1 || !shift_dev) return WF_EINVAL;
  const Box3 b{g.npl - 1, g.npl - 1, g.npl - 1, g.ni + 2, g.nj + 2, g.nk + 2};
  if (!box_ok(g, b, 0, 0)) return WF_EINVAL;
  hipLaunchKernelGGL(add_shift_kernel, CG(b), box_block(), 0, S_(stream), dst, src, shift_dev,
                     g, b, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_fill_interior(double* f, double val, wf_grid g, int d0, int d1, int d2,
                     void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (!box_ok(g, b, 0, 0)) return WF_EINVAL;
  if (empty(b)) return 0;
  hipLaunchKernelGGL(fill_box_kernel, CG(b), box_block(), 0, S_(stream), f, val, g,
                     b, CH(b));
  WF_LAUNCHED();
  return 0;
}

int wf_init_noise(double* u, double* v, double* w, double amp, int32_t seed, wf_grid g,
                  int d0, int d1, int d2, void* stream) {
  const Box3 b = npl_box(g, d0, d1, d2);
  if (!box_ok(g, b, 0, 0)) return WF_EINVAL;
  if (empty(b)) return 0;
  hipLaunchKernelGGL(noise_kernel, CG(b), box_block(), 0, S_(stream), u, v, w, amp,
                     (uint32_t)seed, g, b, CH(b));
  WF_LAUNCHED();
  return 0;
}

}  // extern "C"
