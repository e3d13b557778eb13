// conv4th + diff4th of one velocity component in one pass, streamed along i
// (the reference's conv4th_{u,v,w} + diff4th_{u,v,w}, kernels.py:53-225,
// fused as solver.py:256-265 launches them back to back).
//
// Work decomposition.  A 256-thread workgroup owns a TJ x TK = 8 x 32 (j, k)
// tile of the operator box and walks a chunk of i-planes; thread (tj, tk)
// computes cell (i, j0 + tj, k0 + tk) of every plane in turn.
//
// Where each stencil tap comes from.  The component c = u, v or w:
//   * along i: a per-thread register window of its own column, planes
//     i-3 .. i+3 (the x-interpolants and x-Laplacian of c never touch LDS);
//   * along j and k: plane i of c, staged in LDS with a 3-cell halo in the
//     shape of a cross (no tap of c is off both axes at once).
// The two other components enter only through products of interpolants.
//   * C = 0 (u): the y- and z-terms need v and w interpolated along x at
//     plane i+1 on four rows / columns; each thread keeps a 4-plane register
//     window of v and w at its tile position (and at one halo position) and
//     stages the interpolant, so the compute reads 4 + 4 values from LDS.
//   * C = 1 (v), C = 2 (w): the x-term is D4 over Q(p) = I4x(c)(p+1) *
//     I4y(u)(p) (I4z for w) at planes i+1 .. i-2; Q of the newest plane is
//     formed from plane i+1 of u staged in LDS and rolled through registers.
//     The remaining cross term reads plane i of the other component from LDS.
// The cross-quantity LDS is double-buffered by plane parity, so one barrier
// per plane orders the staging of plane i against the reads of plane i-2;
// every global load for plane i+1 is issued before the compute of plane i.
// The own-component plane has four slots: a thread's halo value of plane
// i+3 is loaded during plane i and stored into its slot during plane i+1, so
// that neighbouring tiles read a halo line close to when its owner tile
// fetched it for its register window (the line is still in L2).
//
// Arithmetic.  Each interpolant, product, D4 and Laplacian is the reference's
// expression term for term (summation order included), except that the
// divisions by 24h and 12h^2 are multiplications by reciprocals formed once
// on the host (a rounding difference of about one ulp) and /16 is *0.0625
// (exact).  Roundings are fixed by the source: contraction is off in this
// file and the multiply-adds that are fused are written as fma (the
// weighted terms of each stencil and the two RK updates), so every kernel
// here -- one component, all three, batched -- gives the same bits.
//
// Workgroup order.  Consecutive workgroup ids are dispatched round-robin over
// the 8 XCDs, each with its own L2; the ids are remapped so that each XCD
// walks a contiguous run of tiles (k fastest, then j, then the i-chunk) and
// neighbouring tiles share their halo rows in the same L2.
#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "wf_common.h"

#pragma clang fp contract(off)

namespace wf {
namespace cdv {

constexpr int TJ = 8, TK = 32, NT = TJ * TK;
static_assert(NT == 256, "one thread per tile cell");

// A staged (j, k) region around the tile: rows [R0, R1), columns [C0, C1)
// relative to the tile origin.  CROSS drops the four corner blocks.  Every
// region position that is not a tile cell is a halo position; there are at
// most NT of them, so each thread stages at most one.
template <int R0, int R1, int C0, int C1, bool CROSS>
struct Region {
  static constexpr int W = C1 - C0, N = (R1 - R0) * W;
  static constexpr int BW = CROSS ? TK : W;  // width of the top and bottom bands
  static constexpr int BC0 = CROSS ? 0 : C0;
  static constexpr int NTOP = -R0 * BW, NBOT = (R1 - TJ) * BW;
  static constexpr int NLEFT = -C0 * TJ, NRIGHT = (C1 - TK) * TJ;
  static constexpr int NHALO = NTOP + NBOT + NLEFT + NRIGHT;
  static_assert(R0 <= 0 && R1 >= TJ && C0 <= 0 && C1 >= TK, "region covers the tile");
  static_assert(NHALO <= NT, "at most one halo position per thread");
  __device__ static int idx(int r, int c) { return (r - R0) * W + (c - C0); }
  // halo position h (0 <= h < NHALO) -> (r, c)
  __device__ static void halo(int h, int& r, int& c) {
    if (h < NTOP) {
      r = R0 + h / BW;
      c = BC0 + h % BW;
      return;
    }
    h -= NTOP;
    if (h < NBOT) {
      r = TJ + h / BW;
      c = BC0 + h % BW;
      return;
    }
    h -= NBOT;
    if constexpr (NLEFT > 0) {
      if (h < NLEFT) {
        r = h / (-C0);
        c = C0 + h % (-C0);
        return;
      }
    }
    h -= NLEFT;
    if constexpr (NRIGHT > 0) {
      r = h / (C1 - TK);
      c = TK + h % (C1 - TK);
    } else {
      r = 0;
      c = 0;
    }
  }
};

// Where one staged value lives: in-plane offset of its (clamped) array
// position and its LDS index.  `on` is false for threads without a halo
// position in this region.
struct Pos {
  int64_t off;
  int lds;
  bool on;
};

struct Geo {
  int64_t sj, si;
  int nI, nJ, nK;
  int jb, kb;  // array row / column of the tile origin
  int ib;      // array plane of output plane 0 of the box
  __device__ int64_t plane(int i) const {  // output plane i -> clamped plane offset
    return (int64_t)min(max(ib + i, 0), nI - 1) * si;
  }
  __device__ int64_t inplane(int r, int c) const {
    const int jf = min(max(jb + r, 0), nJ - 1), kf = min(max(kb + c, 0), nK - 1);
    return (int64_t)jf * sj + kf;
  }
};

template <class R>
__device__ inline Pos interior(const Geo& G, int tj, int tk) {
  return Pos{G.inplane(tj, tk), R::idx(tj, tk), true};
}

template <class R>
__device__ inline Pos halo_pos(const Geo& G, int tid) {
  Pos p{0, 0, tid < R::NHALO};
  int r = 0, c = 0;
  if (p.on) R::halo(tid, r, c);
  p.off = G.inplane(r, c);
  p.lds = R::idx(r, c);
  return p;
}

// (-f[+1] + 9 f[0] + 9 f[-1] - f[-2]) / 16  (kernels.py:31-46)
__device__ inline double i4(double p1, double c0, double m1, double m2) {
  return (__builtin_fma(9.0, m1, __builtin_fma(9.0, c0, -p1)) - m2) * 0.0625;
}
// (-s[+] + 27 s[0] - 27 s[-] + s[--]) / (24 h)  (kernels.py:69)
__device__ inline double d4(double sp, double sn, double sm, double smm, double r24h) {
  return (__builtin_fma(-27.0, sm, __builtin_fma(27.0, sn, -sp)) + smm) * r24h;
}
// (-f[+2] + 16 f[+1] - 30 f[0] + 16 f[-1] - f[-2]) / (12 h^2)  (kernels.py:177-186)
__device__ inline double l4(double p2, double p1, double c, double m1, double m2,
                            double ir12) {
  return (__builtin_fma(16.0, m1, __builtin_fma(-30.0, c, __builtin_fma(16.0, p1, -p2))) -
          m2) *
         ir12;
}
// The two roundings of the RK update (kernels.py:89, :199):
// c_oo - dt alpha N, then + dt alpha (vis L4).
__device__ inline double rk_update(double coo, double nsum, double vis, double lap,
                                   double dta) {
  const double t = __builtin_fma(-dta, nsum, coo);
  return __builtin_fma(dta, vis * lap, t);
}

struct Coefs {
  double r24x, r24y, r24z;     // 1 / (24 h)
  double ir12x, ir12y, ir12z;  // 1 / (12 h^2)
  double dta;                  // dt * alpha
};

// Own-component region: plane i with a cross-shaped 3-cell halo.
using OwnR = Region<-3, TJ + 3, -3, TK + 3, true>;
// C = 0: v interpolated along x on rows -2..+1, w on columns -2..+1.
using IvR = Region<-2, TJ + 1, 0, TK, false>;
using IwR = Region<0, TJ, -2, TK + 1, false>;
// C = 1: u on rows -1..+2 (y-interpolant), w on rows -1..+2 x columns -2..+1.
using U1R = Region<-1, TJ + 2, 0, TK, false>;
using W1R = Region<-1, TJ + 2, -2, TK + 1, false>;
// C = 2: u on columns -1..+2 (z-interpolant), v on rows -2..+1 x columns -1..+2.
using U2R = Region<0, TJ, -1, TK + 2, false>;
using V2R = Region<-2, TJ + 1, -1, TK + 2, false>;

template <int C>
struct Lay {  // regions of the two staged cross quantities
  using A = typename std::conditional<C == 0, IvR, typename std::conditional<C == 1, U1R, U2R>::type>::type;
  using B = typename std::conditional<C == 0, IwR, typename std::conditional<C == 1, W1R, V2R>::type>::type;
};

// One (k tile, j tile, i chunk) work item; every thread of the workgroup
// calls it with the same item.  SAVE (the first RK stage): the step's base
// value c_oo is the component itself, taken from the thread's register
// window instead of being loaded, and written to coo_save on the box -- the
// cells the second stage reads as c_oo -- in place of a full-field copy.
template <int C, bool SAVE = false>
__device__ inline void convdiff_body(const double* __restrict__ u, const double* __restrict__ v,
                                     const double* __restrict__ w,
                                     const double* __restrict__ coo,
                                     const double* __restrict__ vis,
                                     double* __restrict__ cstar, const wf_grid& g,
                                     const Box3& b, int ich, int ktile, int jtile, int chunk,
                                     const Coefs& K, double* __restrict__ coo_save = nullptr) {
  using RA = typename Lay<C>::A;
  using RB = typename Lay<C>::B;
  __shared__ double Ls[4][OwnR::N];
  __shared__ double La[2][RA::N];
  __shared__ double Lb[2][RB::N];

  const double* __restrict__ F = C == 0 ? u : (C == 1 ? v : w);  // own component
  // C = 0: A <- v (x-interpolated), B <- w (x-interpolated)
  // C = 1: A <- u (plane i+1),      B <- w (plane i)
  // C = 2: A <- u (plane i+1),      B <- v (plane i)
  const double* __restrict__ FA = C == 0 ? v : u;
  const double* __restrict__ FB = C == 2 ? v : w;

  const int tid = threadIdx.x;
  const int tj = tid / TK, tk = tid % TK;
  const int j0 = jtile * TJ, k0 = ktile * TK;
  const int i0 = chunk * ich, i1 = min(i0 + ich, b.d0);
  if (i0 >= i1) return;  // uniform per workgroup: no barrier is skipped by part of it

  Geo G;
  G.sj = g.sj;
  G.si = g.si;
  G.nI = g.ni + 2 * g.npl;
  G.nJ = g.nj + 2 * g.npl;
  G.nK = g.nk + 2 * g.npl;
  G.jb = b.lo1 + j0;
  G.kb = b.lo2 + k0;
  G.ib = b.lo0;

  const Pos own = interior<OwnR>(G, tj, tk);  // also the tile cell of every region
  const Pos hs = halo_pos<OwnR>(G, tid);
  const Pos ia = interior<RA>(G, tj, tk), ha = halo_pos<RA>(G, tid);
  const Pos ib = interior<RB>(G, tj, tk), hb = halo_pos<RB>(G, tid);
  const bool live = j0 + tj < b.d1 && k0 + tk < b.d2;
  // vis at the component's face (kernels.py:198, :211, :224), clamped like
  // every other position so that threads outside the box stay in the array
  const int64_t visoff = G.inplane(tj + (C == 1), tk + (C == 2));

  // ---- prologue ----
  double f[7];  // own column, planes i-3 .. i+3 after the shift at the loop top
#pragma unroll
  for (int q = 1; q < 7; ++q) f[q] = F[G.plane(i0 - 4 + q) + own.off];
  f[0] = 0.0;
  double fnew = F[G.plane(i0 + 3) + own.off];
  // halo of planes i0, i0+1 straight into their slots; plane i0+2 in flight
  if (hs.on) {
    Ls[0][hs.lds] = F[G.plane(i0) + hs.off];
    Ls[1][hs.lds] = F[G.plane(i0 + 1) + hs.off];
  }
  double fhalo = hs.on ? F[G.plane(i0 + 2) + hs.off] : 0.0;
  double cnew = 0.0;
  if constexpr (!SAVE) cnew = coo[G.plane(i0) + own.off];
  double vnew = vis[G.plane(i0) + visoff];

  // C = 0: windows of v and w at the tile cell and at the halo position,
  // planes i-1 .. i+2 after the shift
  double va[4], vh[4], wa[4], wh[4];
  double van = 0, vhn = 0, wan = 0, whn = 0;
  // C = 1, 2: newest staged planes of the A and B fields (tile cell, halo)
  double an = 0, ahn = 0, bn = 0, bhn = 0;
  // C = 0: squares of I4x(u) at planes i+1, i, i-1 (before the shift)
  // C = 1, 2: Q at planes i, i-1, i-2
  double r1 = 0, r0 = 0, rm1 = 0;
  if constexpr (C == 0) {
#pragma unroll
    for (int q = 1; q < 4; ++q) {
      const int64_t pb = G.plane(i0 - 2 + q);
      va[q] = FA[pb + ia.off];
      wa[q] = FB[pb + ib.off];
      vh[q] = ha.on ? FA[pb + ha.off] : 0.0;
      wh[q] = hb.on ? FB[pb + hb.off] : 0.0;
    }
    va[0] = vh[0] = wa[0] = wh[0] = 0.0;
    const int64_t pn = G.plane(i0 + 2);
    van = FA[pn + ia.off];
    wan = FB[pn + ib.off];
    vhn = ha.on ? FA[pn + ha.off] : 0.0;
    whn = hb.on ? FB[pn + hb.off] : 0.0;
    // I4x(u) centred at planes i0+1, i0, i0-1 from f[1..6] = planes i0-3 .. i0+2
    const double a1 = i4(f[6], f[5], f[4], f[3]);
    const double a0 = i4(f[5], f[4], f[3], f[2]);
    const double am1 = i4(f[4], f[3], f[2], f[1]);
    r1 = a1 * a1;
    r0 = a0 * a0;
    rm1 = am1 * am1;
  } else {
    // Q(p) = I4x(c)(planes p+2 .. p-1) * I4(u) across the tile cell at plane p,
    // for p = i0, i0-1, i0-2; the u interpolant straight from global memory
    auto iu = [&](int p) {
      const int64_t pb = G.plane(p);
      if constexpr (C == 1)  // along y: rows +2, +1, 0, -1
        return i4(u[pb + G.inplane(tj + 2, tk)], u[pb + G.inplane(tj + 1, tk)],
                  u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj - 1, tk)]);
      else  // along z: columns +2, +1, 0, -1
        return i4(u[pb + G.inplane(tj, tk + 2)], u[pb + G.inplane(tj, tk + 1)],
                  u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj, tk - 1)]);
    };
    r1 = i4(f[6], f[5], f[4], f[3]) * iu(i0);       // Q(i0): planes i0+2 .. i0-1
    r0 = i4(f[5], f[4], f[3], f[2]) * iu(i0 - 1);   // Q(i0-1)
    rm1 = i4(f[4], f[3], f[2], f[1]) * iu(i0 - 2);  // Q(i0-2)
    const int64_t pa = G.plane(i0 + 1), pbp = G.plane(i0);
    an = FA[pa + ia.off];
    ahn = ha.on ? FA[pa + ha.off] : 0.0;
    bn = FB[pbp + ib.off];
    bhn = hb.on ? FB[pbp + hb.off] : 0.0;
  }

  for (int i = i0; i < i1; ++i) {
    const int rel = i - i0, buf = rel & 1;
    // shift the windows to planes i-3 .. i+3 (own) and i-1 .. i+2 (C = 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) f[q] = f[q + 1];
    f[6] = fnew;
    const double cur_coo = SAVE ? f[3] : cnew, cur_vis = vnew;
    double* __restrict__ S = Ls[rel & 3];
    double* __restrict__ A = La[buf];
    double* __restrict__ B = Lb[buf];
    S[own.lds] = f[3];
    // halo of plane i+2 into the slot last read at plane i-2
    if (hs.on) Ls[(rel + 2) & 3][hs.lds] = fhalo;
    if constexpr (C == 0) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        va[q] = va[q + 1];
        vh[q] = vh[q + 1];
        wa[q] = wa[q + 1];
        wh[q] = wh[q + 1];
      }
      va[3] = van;
      vh[3] = vhn;
      wa[3] = wan;
      wh[3] = whn;
      // x-interpolants at plane i+1: (planes i+2, i+1, i, i-1)
      A[ia.lds] = i4(va[3], va[2], va[1], va[0]);
      if (ha.on) A[ha.lds] = i4(vh[3], vh[2], vh[1], vh[0]);
      B[ib.lds] = i4(wa[3], wa[2], wa[1], wa[0]);
      if (hb.on) B[hb.lds] = i4(wh[3], wh[2], wh[1], wh[0]);
    } else {
      A[ia.lds] = an;  // u, plane i+1
      if (ha.on) A[ha.lds] = ahn;
      B[ib.lds] = bn;  // w (C = 1) or v (C = 2), plane i
      if (hb.on) B[hb.lds] = bhn;
    }
    // loads for plane i+1
    if (i + 1 < i1) {
      fnew = F[G.plane(i + 4) + own.off];
      if (hs.on) fhalo = F[G.plane(i + 3) + hs.off];
      if constexpr (!SAVE) cnew = coo[G.plane(i + 1) + own.off];
      vnew = vis[G.plane(i + 1) + visoff];
      if constexpr (C == 0) {
        const int64_t pn = G.plane(i + 3);
        van = FA[pn + ia.off];
        wan = FB[pn + ib.off];
        if (ha.on) vhn = FA[pn + ha.off];
        if (hb.on) whn = FB[pn + hb.off];
      } else {
        const int64_t pa = G.plane(i + 2), pbp = G.plane(i + 1);
        an = FA[pa + ia.off];
        if (ha.on) ahn = FA[pa + ha.off];
        bn = FB[pbp + ib.off];
        if (hb.on) bhn = FB[pbp + hb.off];
      }
    }
    __syncthreads();

    // own component at (0, dj, dk) of plane i
    auto Fo = [&](int dj, int dk) { return S[OwnR::idx(tj + dj, tk + dk)]; };
    const double c0 = f[3];
    double nsum;
    if constexpr (C == 0) {
      // x: conv_self(u) along x, rolled squares of I4x(u)
      const double a2 = i4(f[6], f[5], f[4], f[3]);
      const double s2 = a2 * a2;
      const double tx = d4(s2, r1, r0, rm1, K.r24x);
      rm1 = r0;
      r0 = r1;
      r1 = s2;
      // y: I4y(u) at rows m+1 times I4x(v) at (plane i+1, row m)
      const double uy3 = Fo(3, 0), uy2 = Fo(2, 0), uy1 = Fo(1, 0);
      const double uym1 = Fo(-1, 0), uym2 = Fo(-2, 0), uym3 = Fo(-3, 0);
      const double ty =
          d4(i4(uy3, uy2, uy1, c0) * A[IvR::idx(tj + 1, tk)],
             i4(uy2, uy1, c0, uym1) * A[IvR::idx(tj, tk)],
             i4(uy1, c0, uym1, uym2) * A[IvR::idx(tj - 1, tk)],
             i4(c0, uym1, uym2, uym3) * A[IvR::idx(tj - 2, tk)], K.r24y);
      // z: I4z(u) at columns m+1 times I4x(w) at (plane i+1, column m)
      const double uz3 = Fo(0, 3), uz2 = Fo(0, 2), uz1 = Fo(0, 1);
      const double uzm1 = Fo(0, -1), uzm2 = Fo(0, -2), uzm3 = Fo(0, -3);
      const double tz =
          d4(i4(uz3, uz2, uz1, c0) * B[IwR::idx(tj, tk + 1)],
             i4(uz2, uz1, c0, uzm1) * B[IwR::idx(tj, tk)],
             i4(uz1, c0, uzm1, uzm2) * B[IwR::idx(tj, tk - 1)],
             i4(c0, uzm1, uzm2, uzm3) * B[IwR::idx(tj, tk - 2)], K.r24z);
      nsum = tx + ty + tz;
      if (live) {
        const double lap = l4(f[5], f[4], c0, f[2], f[1], K.ir12x) +
                           l4(uy2, uy1, c0, uym1, uym2, K.ir12y) +
                           l4(uz2, uz1, c0, uzm1, uzm2, K.ir12z);
        cstar[G.plane(i) + own.off] = rk_update(cur_coo, nsum, cur_vis, lap, K.dta);
        if constexpr (SAVE) coo_save[G.plane(i) + own.off] = cur_coo;
      }
    } else {
      // x: D4 over Q(i+1), Q(i), Q(i-1), Q(i-2); Q(i+1) from u at plane i+1
      double iu;
      if constexpr (C == 1)
        iu = i4(A[U1R::idx(tj + 2, tk)], A[U1R::idx(tj + 1, tk)], A[U1R::idx(tj, tk)],
                A[U1R::idx(tj - 1, tk)]);
      else
        iu = i4(A[U2R::idx(tj, tk + 2)], A[U2R::idx(tj, tk + 1)], A[U2R::idx(tj, tk)],
                A[U2R::idx(tj, tk - 1)]);
      const double q1 = i4(f[6], f[5], f[4], f[3]) * iu;
      const double tx = d4(q1, r1, r0, rm1, K.r24x);
      rm1 = r0;
      r0 = r1;
      r1 = q1;
      const double y3 = Fo(3, 0), y2 = Fo(2, 0), y1 = Fo(1, 0);
      const double ym1 = Fo(-1, 0), ym2 = Fo(-2, 0), ym3 = Fo(-3, 0);
      const double z3 = Fo(0, 3), z2 = Fo(0, 2), z1 = Fo(0, 1);
      const double zm1 = Fo(0, -1), zm2 = Fo(0, -2), zm3 = Fo(0, -3);
      double ty, tz;
      if constexpr (C == 1) {
        // y: conv_self(v) along y
        const double p2 = i4(y3, y2, y1, c0), p1 = i4(y2, y1, c0, ym1);
        const double p0 = i4(y1, c0, ym1, ym2), m1 = i4(c0, ym1, ym2, ym3);
        ty = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24y);
        // z: I4z(v) at columns m+1 times I4y(w) across rows (+2 .. -1) at column m
        auto wq = [&](int m) {
          return i4(B[W1R::idx(tj + 2, tk + m)], B[W1R::idx(tj + 1, tk + m)],
                    B[W1R::idx(tj, tk + m)], B[W1R::idx(tj - 1, tk + m)]);
        };
        tz = d4(i4(z3, z2, z1, c0) * wq(1), i4(z2, z1, c0, zm1) * wq(0),
                i4(z1, c0, zm1, zm2) * wq(-1), i4(c0, zm1, zm2, zm3) * wq(-2), K.r24z);
      } else {
        // y: I4y(w) at rows m+1 times I4z(v) across columns (+2 .. -1) at row m
        auto vq = [&](int m) {
          return i4(B[V2R::idx(tj + m, tk + 2)], B[V2R::idx(tj + m, tk + 1)],
                    B[V2R::idx(tj + m, tk)], B[V2R::idx(tj + m, tk - 1)]);
        };
        ty = d4(i4(y3, y2, y1, c0) * vq(1), i4(y2, y1, c0, ym1) * vq(0),
                i4(y1, c0, ym1, ym2) * vq(-1), i4(c0, ym1, ym2, ym3) * vq(-2), K.r24y);
        // z: conv_self(w) along z
        const double p2 = i4(z3, z2, z1, c0), p1 = i4(z2, z1, c0, zm1);
        const double p0 = i4(z1, c0, zm1, zm2), m1 = i4(c0, zm1, zm2, zm3);
        tz = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24z);
      }
      nsum = tx + ty + tz;
      if (live) {
        const double lap = l4(f[5], f[4], c0, f[2], f[1], K.ir12x) +
                           l4(y2, y1, c0, ym1, ym2, K.ir12y) +
                           l4(z2, z1, c0, zm1, zm2, K.ir12z);
        cstar[G.plane(i) + own.off] = rk_update(cur_coo, nsum, cur_vis, lap, K.dta);
        if constexpr (SAVE) coo_save[G.plane(i) + own.off] = cur_coo;
      }
    }
  }
}

// Chunk q of the launch is chunk q * cstep of the box (cstep 1: every chunk;
// the edge pass of u runs chunks 0 and d0 - 1 of one plane each).
template <int C, bool SAVE>
__global__ void __launch_bounds__(256)
convdiff_kernel(const double* __restrict__ u, const double* __restrict__ v,
                const double* __restrict__ w, const double* __restrict__ coo,
                const double* __restrict__ vis, double* __restrict__ cstar, wf_grid g, Box3 b,
                int ich, int kt, int jt, Coefs K, double* __restrict__ coo_save, int cstep) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  convdiff_body<C, SAVE>(u, v, w, coo, vis, cstar, g, b, ich, L % kt, (L / kt) % jt,
                         L / (kt * jt) * cstep, K, coo_save);
}

// ---------------------------------------------------------------------------
// The first RK stage's conv+diff of u, v and w in one pass.
//
// At the first stage the three components' ghosts are as the end-of-step
// fills (or the initial fill) left them, and a repeated fill changes none of
// the cells one component's stencil reads of another (it changes only deep
// edge cells no cross stencil reaches).  So once u, v and w are filled, the
// three conv+diff operators can run together and give what the reference's
// fill/conv sequence gives (solver.py:254-265).  The pass reads u, v, w and
// vis once and writes ustar, vstar, wstar and the saved bases uoo, voo, woo:
// 10 fields of traffic per cell instead of the 18 of three separate passes.
// At the second stage the cross stencils must see some of the other
// components' ghosts as they were before their fills; the same pass serves it
// (convdiff3_body<true>) with those cells taken from a saved copy or left to a
// pass that runs before the fills.
//
// A TJ x TK tile of threads walks a chunk of i-planes over the union of the
// three operator boxes; every output expression is convdiff_body's, term for
// term, so each component's values are bit for bit those of its own pass.
//   registers: planes i-3 .. i+3 of u, v and w in the thread's column; for
//     the halo rows (-2, -1, TJ) of I4x(v) and the halo columns (-2, -1, TK)
//     of I4x(w), four-plane windows of v and w at one position per thread;
//   LDS: u at planes i and i+1 (four slots), v and w at plane i (parity) on
//     a 14 x 38 rectangle of which the 3-cell cross and the corner cells of
//     the other components' cross terms are staged, and I4x(v), I4x(w) at
//     plane i+1 (parity);
//   every global load for the next plane is issued, unconditionally (clamped
//   planes), before the barrier of the current one.
namespace c3 {
constexpr int RW = TK + 6;                          // rectangle rows -3 .. TJ+2,
__device__ inline int rix(int r, int c) {           // columns -3 .. TK+2
  return (r + 3) * RW + (c + 3);
}
constexpr int RN = (TJ + 6) * RW;
constexpr int IVW = TK, IVN = (TJ + 3) * IVW;       // I4x(v): rows -2 .. TJ
__device__ inline int ivx(int r, int c) { return (r + 2) * IVW + c; }
constexpr int IWW = TK + 3, IWN = TJ * IWW;         // I4x(w): columns -2 .. TK
__device__ inline int iwx(int r, int c) { return r * IWW + (c + 2); }
constexpr int NCROSS = 6 * TK + 6 * TJ;             // halo cells of the 3-cell cross
static_assert(NCROSS + 9 <= NT, "one halo cell of each field per thread");
// Halo cell h of a field's staged region: the cross arms, then (FIELD 1, v)
// rows {-2, -1, TJ} x columns {-1, TK, TK+1} or (FIELD 2, w) rows
// {-1, TJ, TJ+1} x columns {-2, -1, TK}; u (FIELD 0) stages the cross only.
template <int FIELD>
__device__ inline bool halo_cell(int h, int& r, int& c) {
  r = c = 0;
  if (h < 3 * TK) {
    r = -3 + h / TK;
    c = h % TK;
  } else if (h < 6 * TK) {
    h -= 3 * TK;
    r = TJ + h / TK;
    c = h % TK;
  } else if (h < 6 * TK + 3 * TJ) {
    h -= 6 * TK;
    r = h / 3;
    c = -3 + h % 3;
  } else if (h < NCROSS) {
    h -= 6 * TK + 3 * TJ;
    r = h / 3;
    c = TK + h % 3;
  } else if (FIELD != 0 && h < NCROSS + 9) {
    h -= NCROSS;
    const int a = h / 3, b = h % 3;
    if (FIELD == 1) {
      r = a == 0 ? -2 : (a == 1 ? -1 : TJ);
      c = b == 0 ? -1 : TK + b - 1;
    } else {
      r = a == 0 ? -1 : TJ + a - 1;
      c = b == 2 ? TK : b - 2;
    }
  } else {
    return false;
  }
  return true;
}
}  // namespace c3

// Halo cells of the u plane and of w as v's z-term sees it, for the second
// stage's passes (convdiff3_body<true>, convdiff_vw_body).
namespace c2 {
// halo cells of the u plane: rows -1, TJ, TJ+1 on tile columns, columns -1,
// TK, TK+1 on tile rows
constexpr int NUH = 3 * TK + 3 * TJ;
__device__ inline bool u_halo(int h, int& r, int& c) {
  r = c = 0;
  if (h < 3 * TK) {
    r = h < TK ? -1 : TJ + h / TK - 1;
    c = h % TK;
  } else if (h < NUH) {
    h -= 3 * TK;
    r = h % TJ;
    const int q = h / TJ;
    c = q == 0 ? -1 : TK + q - 1;
  } else {
    return false;
  }
  return true;
}
// halo cells of the w plane v's z-term reads (rows -1 .. TJ+1, columns -2 ..
// TK): rows -1, TJ, TJ+1 on tile columns, columns -2, -1, TK on tile rows,
// and the nine corner cells
constexpr int NWH = 3 * TK + 3 * TJ + 9;
static_assert(NUH <= NT && NWH <= NT, "one halo cell per thread");
__device__ inline bool w1_halo(int h, int& r, int& c) {
  r = c = 0;
  auto row = [](int q) { return q == 0 ? -1 : TJ + q - 1; };  // -1, TJ, TJ+1
  auto col = [](int q) { return q == 2 ? TK : q - 2; };       // -2, -1, TK
  if (h < 3 * TK) {
    r = row(h / TK);
    c = h % TK;
  } else if (h < 3 * TK + 3 * TJ) {
    h -= 3 * TK;
    r = h % TJ;
    c = col(h / TJ);
  } else if (h < NWH) {
    h -= 3 * TK + 3 * TJ;
    r = row(h / 3);
    c = col(h % 3);
  } else {
    return false;
  }
  return true;
}
}  // namespace c2

// S2 = false: the first stage (convdiff3_save_kernel); uoo, voo, woo are
// written with the components' values.  S2 = true: the second stage
// (convdiff3_kernel), after the fills of u, v and w; uoo, voo, woo are read as
// the base values, v's z-term reads w as it was before w's fill (wpre) at the
// cells outside w's operator box, as convdiff_vw_body does, and u's planes 0
// and ni-2 are left to the caller's edge pass: their x-interpolants of v and
// w reach the x-ghost planes -1 and ni, which the fills of v and w change and
// which conv_u reads before those fills.  u's other planes read no cell either
// fill changes.
template <bool S2>
__device__ inline void convdiff3_body(
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
    const double* __restrict__ vis, const double* __restrict__ wpre, double* __restrict__ us,
    double* __restrict__ vs, double* __restrict__ ws, double* __restrict__ uoo,
    double* __restrict__ voo, double* __restrict__ woo, const wf_grid& g, int ich, int ktile,
    int jtile, int chunk, const Coefs& K) {
  using namespace c3;
  __shared__ double Us[3][RN];  // u at planes i, i+1 (a third slot for the next)
  __shared__ double Vs[2][RN];
  __shared__ double Ws[2][RN];
  __shared__ double IV[2][IVN];
  __shared__ double IW[2][IWN];
  __shared__ double W1s[S2 ? 2 : 1][S2 ? RN : 1];  // w as v's z-term reads it (S2)
  // v and w at the I4x halo positions, planes i-1 .. i+2 (slot (p - i0 + 4) & 3),
  // each position written and read by its own thread only
  __shared__ double XR[4][3 * TK + 3 * TJ];

  const int tid = threadIdx.x;
  const int tj = tid / TK, tk = tid % TK;
  const int j0 = jtile * TJ, k0 = ktile * TK;
  const int i0 = chunk * ich, i1 = min(i0 + ich, g.ni);
  if (i0 >= i1) return;  // uniform per workgroup

  Geo G;
  G.sj = g.sj;
  G.si = g.si;
  G.nI = g.ni + 2 * g.npl;
  G.nJ = g.nj + 2 * g.npl;
  G.nK = g.nk + 2 * g.npl;
  G.jb = g.npl + j0;
  G.kb = g.npl + k0;
  G.ib = g.npl;

  // operator boxes (the reference's _op_dim): u (ni-1, nj, nk), v (ni,
  // nj-1, nk), w (ni, nj, nk-1); the tile walks their union (ni, nj, nk)
  const int jj = j0 + tj, kk = k0 + tk;
  const bool in_jk = jj < g.nj && kk < g.nk;
  const bool lu_jk = in_jk, lv_jk = in_jk && jj < g.nj - 1, lw_jk = in_jk && kk < g.nk - 1;

  const int64_t own = G.inplane(tj, tk);
  const int ownl = rix(tj, tk);
  int hr, hc;
  const bool hu = halo_cell<0>(tid, hr, hc);
  const int64_t hu_off = G.inplane(hr, hc);
  const int hu_l = rix(hr, hc);
  const bool hv = halo_cell<1>(tid, hr, hc);
  const int64_t hv_off = G.inplane(hr, hc);
  const int hv_l = rix(hr, hc);
  const bool hw = halo_cell<2>(tid, hr, hc);
  const int64_t hw_off = G.inplane(hr, hc);
  const int hw_l = rix(hr, hc);
  // I4x(v) halo rows -2, -1, TJ on threads 0..95; I4x(w) halo columns -2,
  // -1, TK on threads 128..151
  const bool xv = tid < 3 * TK;
  const int xvr = tid < 2 * TK ? -2 + tid / TK : TJ, xvc = tid % TK;
  const int64_t xv_off = G.inplane(xv ? xvr : 0, xv ? xvc : 0);
  const int xv_l = ivx(xv ? xvr : 0, xv ? xvc : 0);
  const int tw = tid - 128;
  const bool xw = tw >= 0 && tw < 3 * TJ;
  const int xwc = tw < 2 * TJ ? -2 + tw / TJ : TK, xwr = tw % TJ;
  const int64_t xw_off = G.inplane(xw ? xwr : 0, xw ? xwc : 0);
  const int xw_l = iwx(xw ? xwr : 0, xw ? xwc : 0);
  const int xr = xv ? tid : 3 * TK + (xw ? tw : 0);  // ring position
  // vis at each component's face (kernels.py:198, :211, :224)
  const int64_t vis_u = own, vis_v = G.inplane(tj + 1, tk), vis_w = G.inplane(tj, tk + 1);
  // in the plane loop every access is a raw buffer access: a per-plane
  // descriptor (scalar) and a 32-bit byte offset per lane (the host keeps a
  // plane below 2^30 bytes); a store outside the component's box goes to
  // BUF_DEAD and is dropped
  auto prs = [&](const double* a, int p) { return buf_rsrc(a + G.plane(p), 0x3fffffffu); };
  const uint32_t b_own = (uint32_t)own * 8u, b_hu = (uint32_t)hu_off * 8u,
                 b_hv = (uint32_t)hv_off * 8u, b_hw = (uint32_t)hw_off * 8u,
                 b_xv = (uint32_t)xv_off * 8u, b_xw = (uint32_t)xw_off * 8u,
                 b_vv = (uint32_t)vis_v * 8u, b_vw = (uint32_t)vis_w * 8u;
  const uint32_t st_v = lv_jk ? b_own : BUF_DEAD, st_w = lw_jk ? b_own : BUF_DEAD;
  // S2: w as v's z-term reads it, from wpre outside w's operator box in (j, k)
  auto w_out = [&](int r, int c) {
    const int J = j0 + r, Kc = k0 + c;
    return J < 0 || J >= g.nj || Kc < 0 || Kc >= g.nk - 1;
  };
  int h1r = 0, h1c = 0;
  const bool h1 = S2 && c2::w1_halo(tid, h1r, h1c);
  const int64_t h1_off = G.inplane(h1r, h1c);
  const int h1_l = rix(h1r, h1c);
  const double* __restrict__ w1h_src = S2 && h1 && w_out(h1r, h1c) ? wpre : w;
  const double* __restrict__ w1t_src = S2 && w_out(tj, tk) ? wpre : w;

  // ---- prologue ----
  double fu[7], fv[7], fw[7];  // planes i-3 .. i+3 after the shift at the loop top
#pragma unroll
  for (int q = 1; q < 7; ++q) {
    const int64_t pb = G.plane(i0 - 4 + q);
    fu[q] = u[pb + own];
    fv[q] = v[pb + own];
    fw[q] = w[pb + own];
  }
  fu[0] = fv[0] = fw[0] = 0.0;
  double nu = u[G.plane(i0 + 3) + own], nv = v[G.plane(i0 + 3) + own],
         nw = w[G.plane(i0 + 3) + own];
  // u plane i0 into its slot (the loop stages plane i+1 in slot (i+1-i0) & 3)
  Us[0][ownl] = fu[4];
  if (hu) Us[0][hu_l] = u[G.plane(i0) + hu_off];
  double hun = u[G.plane(i0 + 1) + hu_off];  // halo of u plane i0+1
  double hvn = v[G.plane(i0) + hv_off];      // halo of v, w plane i0
  double hwn = w[G.plane(i0) + hw_off];
  // v, w at the I4x halo positions, planes i0-1 .. i0+1, into the ring
#pragma unroll
  for (int q = 1; q < 4; ++q) {
    const int64_t pb = G.plane(i0 - 2 + q);
    if (xv) XR[(q + 2) & 3][xr] = v[pb + xv_off];
    if (xw) XR[(q + 2) & 3][xr] = w[pb + xw_off];
  }
  double xvn = v[G.plane(i0 + 2) + xv_off], xwn = w[G.plane(i0 + 2) + xw_off];
  double snu = vis[G.plane(i0) + vis_u], snv = vis[G.plane(i0) + vis_v],
         snw = vis[G.plane(i0) + vis_w];
  double w1t = 0.0, w1h = 0.0;  // S2
  if constexpr (S2) {
    const int64_t p0 = G.plane(i0);
    w1t = w1t_src[p0 + own];
    w1h = w1h_src[p0 + h1_off];
  }
  // rolled x-terms: squares of I4x(u) (u), Q = I4x(c) * I4(u) (v, w), at
  // planes i+1, i, i-1 before the shift (f[1..6] = planes i0-3 .. i0+2)
  double ur1, ur0, urm1, vr1, vr0, vrm1, wr1, wr0, wrm1;
  {
    const double a1 = i4(fu[6], fu[5], fu[4], fu[3]);
    const double a0 = i4(fu[5], fu[4], fu[3], fu[2]);
    const double am1 = i4(fu[4], fu[3], fu[2], fu[1]);
    ur1 = a1 * a1;
    ur0 = a0 * a0;
    urm1 = am1 * am1;
    auto iuy = [&](int p) {  // I4y(u) across rows +2 .. -1 at plane p
      const int64_t pb = G.plane(p);
      return i4(u[pb + G.inplane(tj + 2, tk)], u[pb + G.inplane(tj + 1, tk)],
                u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj - 1, tk)]);
    };
    auto iuz = [&](int p) {  // I4z(u) across columns +2 .. -1 at plane p
      const int64_t pb = G.plane(p);
      return i4(u[pb + G.inplane(tj, tk + 2)], u[pb + G.inplane(tj, tk + 1)],
                u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj, tk - 1)]);
    };
    vr1 = i4(fv[6], fv[5], fv[4], fv[3]) * iuy(i0);
    vr0 = i4(fv[5], fv[4], fv[3], fv[2]) * iuy(i0 - 1);
    vrm1 = i4(fv[4], fv[3], fv[2], fv[1]) * iuy(i0 - 2);
    wr1 = i4(fw[6], fw[5], fw[4], fw[3]) * iuz(i0);
    wr0 = i4(fw[5], fw[4], fw[3], fw[2]) * iuz(i0 - 1);
    wrm1 = i4(fw[4], fw[3], fw[2], fw[1]) * iuz(i0 - 2);
  }

  for (int i = i0; i < i1; ++i) {
    const int rel = i - i0, buf = rel & 1;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      fu[q] = fu[q + 1];
      fv[q] = fv[q + 1];
      fw[q] = fw[q + 1];
    }
    fu[6] = nu;
    fv[6] = nv;
    fw[6] = nw;
    const double vis_pu = snu, vis_pv = snv, vis_pw = snw;  // first stage's form
    double* __restrict__ U0 = Us[rel % 3];        // u, plane i
    double* __restrict__ U1 = Us[(rel + 1) % 3];  // u, plane i+1
    double* __restrict__ V = Vs[buf];
    double* __restrict__ W = Ws[buf];
    double* __restrict__ AV = IV[buf];
    double* __restrict__ AW = IW[buf];
    U1[ownl] = fu[4];
    if (hu) U1[hu_l] = hun;
    V[ownl] = fv[3];
    if (hv) V[hv_l] = hvn;
    W[ownl] = fw[3];
    if (hw) W[hw_l] = hwn;
    // x-interpolants at plane i+1 (planes i+2, i+1, i, i-1)
    AV[ivx(tj, tk)] = i4(fv[5], fv[4], fv[3], fv[2]);
    const int r2 = (rel + 2) & 3, r1 = (rel + 1) & 3, r0 = rel & 3, rm = (rel + 3) & 3;
    if (xv) {
      XR[r2][xr] = xvn;
      AV[xv_l] = i4(xvn, XR[r1][xr], XR[r0][xr], XR[rm][xr]);
    }
    AW[iwx(tj, tk)] = i4(fw[5], fw[4], fw[3], fw[2]);
    if (xw) {
      XR[r2][xr] = xwn;
      AW[xw_l] = i4(xwn, XR[r1][xr], XR[r0][xr], XR[rm][xr]);
    }
    double* __restrict__ Wv = W;  // w as v's z-term reads it
    if constexpr (S2) {
      Wv = W1s[buf];
      Wv[ownl] = w1t;
      if (h1) Wv[h1_l] = w1h;
    }
    // loads for plane i+1 (clamped past the chunk: a load under a branch
    // would make the compiler wait for every load in flight where the paths
    // join)
    {
      const auto u4 = prs(u, i + 4), v4 = prs(v, i + 4), w4 = prs(w, i + 4);
      nu = buf_ld<double>(u4, b_own, 0);
      nv = buf_ld<double>(v4, b_own, 0);
      nw = buf_ld<double>(w4, b_own, 0);
      hun = buf_ld<double>(prs(u, i + 2), b_hu, 0);
      const auto v1 = prs(v, i + 1), w1 = prs(w, i + 1), s1 = prs(vis, i + 1);
      hvn = buf_ld<double>(v1, b_hv, 0);
      hwn = buf_ld<double>(w1, b_hw, 0);
      xvn = buf_ld<double>(prs(v, i + 3), b_xv, 0);
      xwn = buf_ld<double>(prs(w, i + 3), b_xw, 0);
      if constexpr (!S2) {  // (the second stage's form loads them per component)
        snu = buf_ld<double>(s1, b_own, 0);
        snv = buf_ld<double>(s1, b_vv, 0);
        snw = buf_ld<double>(s1, b_vw, 0);
      }
      if constexpr (S2) {
        const int64_t p1 = G.plane(i + 1);
        w1t = w1t_src[p1 + own];
        w1h = w1h_src[p1 + h1_off];
      }
    }
    __syncthreads();

    const uint32_t st_u =
        lu_jk && i < g.ni - 1 && !(S2 && (i == 0 || i == g.ni - 2)) ? b_own : BUF_DEAD;
    // ---- u (convdiff_body<0, true>) ----
    {
      double coo_u = 0.0, vis_cu = vis_pu;
      if constexpr (S2) {
        coo_u = buf_ld<double>(prs(uoo, i), b_own, 0);
        vis_cu = buf_ld<double>(prs(vis, i), b_own, 0);
      }
      auto Fo = [&](int dj, int dk) { return U0[rix(tj + dj, tk + dk)]; };
      const double c0 = fu[3];
      const double a2 = i4(fu[6], fu[5], fu[4], fu[3]);
      const double s2 = a2 * a2;
      const double tx = d4(s2, ur1, ur0, urm1, K.r24x);
      urm1 = ur0;
      ur0 = ur1;
      ur1 = s2;
      const double uy3 = Fo(3, 0), uy2 = Fo(2, 0), uy1 = Fo(1, 0);
      const double uym1 = Fo(-1, 0), uym2 = Fo(-2, 0), uym3 = Fo(-3, 0);
      const double ty = d4(i4(uy3, uy2, uy1, c0) * AV[ivx(tj + 1, tk)],
                           i4(uy2, uy1, c0, uym1) * AV[ivx(tj, tk)],
                           i4(uy1, c0, uym1, uym2) * AV[ivx(tj - 1, tk)],
                           i4(c0, uym1, uym2, uym3) * AV[ivx(tj - 2, tk)], K.r24y);
      const double uz3 = Fo(0, 3), uz2 = Fo(0, 2), uz1 = Fo(0, 1);
      const double uzm1 = Fo(0, -1), uzm2 = Fo(0, -2), uzm3 = Fo(0, -3);
      const double tz = d4(i4(uz3, uz2, uz1, c0) * AW[iwx(tj, tk + 1)],
                           i4(uz2, uz1, c0, uzm1) * AW[iwx(tj, tk)],
                           i4(uz1, c0, uzm1, uzm2) * AW[iwx(tj, tk - 1)],
                           i4(c0, uzm1, uzm2, uzm3) * AW[iwx(tj, tk - 2)], K.r24z);
      const double nsum = tx + ty + tz;
      const double lap = l4(fu[5], fu[4], c0, fu[2], fu[1], K.ir12x) +
                         l4(uy2, uy1, c0, uym1, uym2, K.ir12y) +
                         l4(uz2, uz1, c0, uzm1, uzm2, K.ir12z);
      buf_st<double, BUF_NT>(prs(us, i), st_u, 0,
                             rk_update(S2 ? coo_u : c0, nsum, vis_cu, lap, K.dta));
      if constexpr (!S2) buf_st<double, BUF_NT>(prs(uoo, i), st_u, 0, c0);
    }
    // the three components one after the other: the scheduler may not mix
    // them, which would hold all three stencils' operands live at once
    __builtin_amdgcn_sched_barrier(0);
    // ---- v (convdiff_body<1, true>) ----
    {
      double coo_v = 0.0, vis_cv = vis_pv;
      if constexpr (S2) {
        coo_v = buf_ld<double>(prs(voo, i), b_own, 0);
        vis_cv = buf_ld<double>(prs(vis, i), b_vv, 0);
      }
      auto Fo = [&](int dj, int dk) { return V[rix(tj + dj, tk + dk)]; };
      const double c0 = fv[3];
      const double iu = i4(U1[rix(tj + 2, tk)], U1[rix(tj + 1, tk)], U1[rix(tj, tk)],
                           U1[rix(tj - 1, tk)]);
      const double q1 = i4(fv[6], fv[5], fv[4], fv[3]) * iu;
      const double tx = d4(q1, vr1, vr0, vrm1, K.r24x);
      vrm1 = vr0;
      vr0 = vr1;
      vr1 = q1;
      const double y3 = Fo(3, 0), y2 = Fo(2, 0), y1 = Fo(1, 0);
      const double ym1 = Fo(-1, 0), ym2 = Fo(-2, 0), ym3 = Fo(-3, 0);
      const double z3 = Fo(0, 3), z2 = Fo(0, 2), z1 = Fo(0, 1);
      const double zm1 = Fo(0, -1), zm2 = Fo(0, -2), zm3 = Fo(0, -3);
      const double p2 = i4(y3, y2, y1, c0), p1 = i4(y2, y1, c0, ym1);
      const double p0 = i4(y1, c0, ym1, ym2), m1 = i4(c0, ym1, ym2, ym3);
      const double ty = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24y);
      auto wq = [&](int m) {
        return i4(Wv[rix(tj + 2, tk + m)], Wv[rix(tj + 1, tk + m)], Wv[rix(tj, tk + m)],
                  Wv[rix(tj - 1, tk + m)]);
      };
      const double tz = d4(i4(z3, z2, z1, c0) * wq(1), i4(z2, z1, c0, zm1) * wq(0),
                           i4(z1, c0, zm1, zm2) * wq(-1), i4(c0, zm1, zm2, zm3) * wq(-2),
                           K.r24z);
      const double nsum = tx + ty + tz;
      const double lap = l4(fv[5], fv[4], c0, fv[2], fv[1], K.ir12x) +
                         l4(y2, y1, c0, ym1, ym2, K.ir12y) +
                         l4(z2, z1, c0, zm1, zm2, K.ir12z);
      buf_st<double, BUF_NT>(prs(vs, i), st_v, 0,
                             rk_update(S2 ? coo_v : c0, nsum, vis_cv, lap, K.dta));
      if constexpr (!S2) buf_st<double, BUF_NT>(prs(voo, i), st_v, 0, c0);
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- w (convdiff_body<2, true>) ----
    {
      double coo_w = 0.0, vis_cw = vis_pw;
      if constexpr (S2) {
        coo_w = buf_ld<double>(prs(woo, i), b_own, 0);
        vis_cw = buf_ld<double>(prs(vis, i), b_vw, 0);
      }
      auto Fo = [&](int dj, int dk) { return W[rix(tj + dj, tk + dk)]; };
      const double c0 = fw[3];
      const double iu = i4(U1[rix(tj, tk + 2)], U1[rix(tj, tk + 1)], U1[rix(tj, tk)],
                           U1[rix(tj, tk - 1)]);
      const double q1 = i4(fw[6], fw[5], fw[4], fw[3]) * iu;
      const double tx = d4(q1, wr1, wr0, wrm1, K.r24x);
      wrm1 = wr0;
      wr0 = wr1;
      wr1 = q1;
      const double y3 = Fo(3, 0), y2 = Fo(2, 0), y1 = Fo(1, 0);
      const double ym1 = Fo(-1, 0), ym2 = Fo(-2, 0), ym3 = Fo(-3, 0);
      const double z3 = Fo(0, 3), z2 = Fo(0, 2), z1 = Fo(0, 1);
      const double zm1 = Fo(0, -1), zm2 = Fo(0, -2), zm3 = Fo(0, -3);
      auto vq = [&](int m) {
        return i4(V[rix(tj + m, tk + 2)], V[rix(tj + m, tk + 1)], V[rix(tj + m, tk)],
                  V[rix(tj + m, tk - 1)]);
      };
      const double ty = d4(i4(y3, y2, y1, c0) * vq(1), i4(y2, y1, c0, ym1) * vq(0),
                           i4(y1, c0, ym1, ym2) * vq(-1), i4(c0, ym1, ym2, ym3) * vq(-2),
                           K.r24y);
      const double p2 = i4(z3, z2, z1, c0), p1 = i4(z2, z1, c0, zm1);
      const double p0 = i4(z1, c0, zm1, zm2), m1 = i4(c0, zm1, zm2, zm3);
      const double tz = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24z);
      const double nsum = tx + ty + tz;
      const double lap = l4(fw[5], fw[4], c0, fw[2], fw[1], K.ir12x) +
                         l4(y2, y1, c0, ym1, ym2, K.ir12y) +
                         l4(z2, z1, c0, zm1, zm2, K.ir12z);
      buf_st<double, BUF_NT>(prs(ws, i), st_w, 0,
                             rk_update(S2 ? coo_w : c0, nsum, vis_cw, lap, K.dta));
      if constexpr (!S2) buf_st<double, BUF_NT>(prs(woo, i), st_w, 0, c0);
    }
  }
}

// Three workgroups per CU: 44 KB of LDS each, and the kernel fits 168 VGPRs
// without spilling (with the components evaluated one after the other).
__global__ void __launch_bounds__(256, 3)
convdiff3_save_kernel(const double* __restrict__ u, const double* __restrict__ v,
                      const double* __restrict__ w, const double* __restrict__ vis,
                      double* __restrict__ us, double* __restrict__ vs, double* __restrict__ ws,
                      double* __restrict__ uoo, double* __restrict__ voo,
                      double* __restrict__ woo, wf_grid g, int ich, int kt, int jt, Coefs K) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  convdiff3_body<false>(u, v, w, vis, nullptr, us, vs, ws, uoo, voo, woo, g, ich, L % kt,
                        (L / kt) % jt, L / (kt * jt), K);
}

// The second stage's form (convdiff3_body<true>): three workgroups per CU
// (52 KB of LDS each), which its 168 VGPRs allow because it loads each
// component's base value and vis at the start of that component instead of
// one plane ahead.
__global__ void __launch_bounds__(256, 3)
convdiff3_kernel(const double* __restrict__ u, const double* __restrict__ v,
                 const double* __restrict__ w, const double* __restrict__ vis,
                 const double* __restrict__ wpre, double* __restrict__ us,
                 double* __restrict__ vs, double* __restrict__ ws, double* __restrict__ uoo,
                 double* __restrict__ voo, double* __restrict__ woo, wf_grid g, int ich, int kt,
                 int jt, Coefs K) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  convdiff3_body<true>(u, v, w, vis, wpre, us, vs, ws, uoo, voo, woo, g, ich, L % kt,
                       (L / kt) % jt, L / (kt * jt), K);
}

// ---------------------------------------------------------------------------
// The second RK stage's conv+diff of v and w in one pass.
//
// At the second stage calvel has changed u, v and w but not their ghosts, and
// the reference fills and convects one component at a time (solver.py:254-
// 265): conv_v reads u and v after their fills and w before its fill; conv_w
// reads all three after their fills.  u's pass runs on its own right after
// u's fill (where this pass is used instead of convdiff3_kernel, which also
// takes u); this pass runs after the fills of v and w and reads, for v's
// z-term, w's values from before w's fill where w's fill writes: the cells
// outside w's operator box that the term reaches (rows -1 and nj, columns
// -2, -1, nk-1, nk of each plane), which save_cross_ghosts copied to w_pre
// just before that fill.  Every other read is the fresh field, as in the
// reference.  It reads u, v, w, voo, woo and vis once (the two single passes
// read u, v, w, vis twice) -- 8 fields of traffic instead of 12.  Output
// expressions are convdiff_body<1>'s and <2>'s term for term.
//   registers: planes i-3 .. i+3 of v and w in the thread's column;
//   LDS (parity): u at plane i+1 (rows -1 .. TJ+1 and columns -1 .. TK+1
//     around the tile), v and w at plane i (3-cell cross; v also the corner
//     cells w's y-term reaches) and w at plane i as v's z-term sees it.

__device__ inline void convdiff_vw_body(
    const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
    const double* __restrict__ voo, const double* __restrict__ woo,
    const double* __restrict__ vis, const double* __restrict__ wpre, double* __restrict__ vs,
    double* __restrict__ ws, const wf_grid& g, int ich, int ktile, int jtile, int chunk,
    const Coefs& K) {
  using c3::rix;
  using c3::RN;
  __shared__ double Ux[2][RN];
  __shared__ double Vs[2][RN];
  __shared__ double Ws[2][RN];
  __shared__ double W1[2][RN];

  const int tid = threadIdx.x;
  const int tj = tid / TK, tk = tid % TK;
  const int j0 = jtile * TJ, k0 = ktile * TK;
  const int i0 = chunk * ich, i1 = min(i0 + ich, g.ni);
  if (i0 >= i1) return;  // uniform per workgroup

  Geo G;
  G.sj = g.sj;
  G.si = g.si;
  G.nI = g.ni + 2 * g.npl;
  G.nJ = g.nj + 2 * g.npl;
  G.nK = g.nk + 2 * g.npl;
  G.jb = g.npl + j0;
  G.kb = g.npl + k0;
  G.ib = g.npl;

  // operator boxes: v (ni, nj-1, nk), w (ni, nj, nk-1)
  const int jj = j0 + tj, kk = k0 + tk;
  const bool lv_jk = jj < g.nj - 1 && kk < g.nk, lw_jk = jj < g.nj && kk < g.nk - 1;
  // outside w's operator box in (j, k): there v's z-term reads w_pre
  auto w_out = [&](int r, int c) {
    const int J = j0 + r, Kc = k0 + c;
    return J < 0 || J >= g.nj || Kc < 0 || Kc >= g.nk - 1;
  };

  const int64_t own = G.inplane(tj, tk);
  const int ownl = rix(tj, tk);
  int hr, hc;
  const bool hu = c2::u_halo(tid, hr, hc);
  const int64_t hu_off = G.inplane(hr, hc);
  const int hu_l = rix(hr, hc);
  const bool hv = c3::halo_cell<1>(tid, hr, hc);
  const int64_t hv_off = G.inplane(hr, hc);
  const int hv_l = rix(hr, hc);
  const bool hw = c3::halo_cell<0>(tid, hr, hc);
  const int64_t hw_off = G.inplane(hr, hc);
  const int hw_l = rix(hr, hc);
  const bool h1 = c2::w1_halo(tid, hr, hc);
  const int64_t h1_off = G.inplane(hr, hc);
  const int h1_l = rix(hr, hc);
  const double* __restrict__ w1h_src = h1 && w_out(hr, hc) ? wpre : w;
  const double* __restrict__ w1t_src = w_out(tj, tk) ? wpre : w;
  const int64_t vis_v = G.inplane(tj + 1, tk), vis_w = G.inplane(tj, tk + 1);
  auto prs = [&](const double* a, int p) { return buf_rsrc(a + G.plane(p), 0x3fffffffu); };
  const uint32_t b_own = (uint32_t)own * 8u, b_hu = (uint32_t)hu_off * 8u,
                 b_hv = (uint32_t)hv_off * 8u, b_hw = (uint32_t)hw_off * 8u,
                 b_vv = (uint32_t)vis_v * 8u, b_vw = (uint32_t)vis_w * 8u;
  const uint32_t st_v = lv_jk ? b_own : BUF_DEAD, st_w = lw_jk ? b_own : BUF_DEAD;

  // ---- prologue ----
  double fv[7], fw[7];  // planes i-3 .. i+3 after the shift at the loop top
#pragma unroll
  for (int q = 1; q < 7; ++q) {
    const int64_t pb = G.plane(i0 - 4 + q);
    fv[q] = v[pb + own];
    fw[q] = w[pb + own];
  }
  fv[0] = fw[0] = 0.0;
  double nv = v[G.plane(i0 + 3) + own], nw = w[G.plane(i0 + 3) + own];
  double utn = u[G.plane(i0 + 1) + own], uhn = u[G.plane(i0 + 1) + hu_off];
  double hvn = v[G.plane(i0) + hv_off], hwn = w[G.plane(i0) + hw_off];
  double w1t = w1t_src[G.plane(i0) + own], w1h = w1h_src[G.plane(i0) + h1_off];
  double cvn = voo[G.plane(i0) + own], cwn = woo[G.plane(i0) + own];
  double snv = vis[G.plane(i0) + vis_v], snw = vis[G.plane(i0) + vis_w];
  double vr1, vr0, vrm1, wr1, wr0, wrm1;
  {
    auto iuy = [&](int p) {  // I4y(u) across rows +2 .. -1 at plane p
      const int64_t pb = G.plane(p);
      return i4(u[pb + G.inplane(tj + 2, tk)], u[pb + G.inplane(tj + 1, tk)],
                u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj - 1, tk)]);
    };
    auto iuz = [&](int p) {  // I4z(u) across columns +2 .. -1 at plane p
      const int64_t pb = G.plane(p);
      return i4(u[pb + G.inplane(tj, tk + 2)], u[pb + G.inplane(tj, tk + 1)],
                u[pb + G.inplane(tj, tk)], u[pb + G.inplane(tj, tk - 1)]);
    };
    vr1 = i4(fv[6], fv[5], fv[4], fv[3]) * iuy(i0);
    vr0 = i4(fv[5], fv[4], fv[3], fv[2]) * iuy(i0 - 1);
    vrm1 = i4(fv[4], fv[3], fv[2], fv[1]) * iuy(i0 - 2);
    wr1 = i4(fw[6], fw[5], fw[4], fw[3]) * iuz(i0);
    wr0 = i4(fw[5], fw[4], fw[3], fw[2]) * iuz(i0 - 1);
    wrm1 = i4(fw[4], fw[3], fw[2], fw[1]) * iuz(i0 - 2);
  }

  for (int i = i0; i < i1; ++i) {
    const int buf = (i - i0) & 1;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      fv[q] = fv[q + 1];
      fw[q] = fw[q + 1];
    }
    fv[6] = nv;
    fw[6] = nw;
    const double coo_v = cvn, coo_w = cwn, vis_cv = snv, vis_cw = snw;
    double* __restrict__ U1 = Ux[buf];  // u, plane i+1
    double* __restrict__ V = Vs[buf];
    double* __restrict__ W = Ws[buf];
    double* __restrict__ Wv = W1[buf];  // w as v's z-term reads it
    U1[ownl] = utn;
    if (hu) U1[hu_l] = uhn;
    V[ownl] = fv[3];
    if (hv) V[hv_l] = hvn;
    W[ownl] = fw[3];
    if (hw) W[hw_l] = hwn;
    Wv[ownl] = w1t;
    if (h1) Wv[h1_l] = w1h;
    // loads for plane i+1, unconditional (clamped planes)
    {
      const auto v4 = prs(v, i + 4), w4 = prs(w, i + 4), u2 = prs(u, i + 2);
      nv = buf_ld<double>(v4, b_own, 0);
      nw = buf_ld<double>(w4, b_own, 0);
      utn = buf_ld<double>(u2, b_own, 0);
      uhn = buf_ld<double>(u2, b_hu, 0);
      const auto v1 = prs(v, i + 1), w1 = prs(w, i + 1), s1 = prs(vis, i + 1);
      hvn = buf_ld<double>(v1, b_hv, 0);
      hwn = buf_ld<double>(w1, b_hw, 0);
      const int64_t p1 = G.plane(i + 1);
      w1t = w1t_src[p1 + own];
      w1h = w1h_src[p1 + h1_off];
      cvn = buf_ld<double>(prs(voo, i + 1), b_own, 0);
      cwn = buf_ld<double>(prs(woo, i + 1), b_own, 0);
      snv = buf_ld<double>(s1, b_vv, 0);
      snw = buf_ld<double>(s1, b_vw, 0);
    }
    __syncthreads();

    // ---- v (convdiff_body<1, false>) ----
    {
      auto Fo = [&](int dj, int dk) { return V[rix(tj + dj, tk + dk)]; };
      const double c0 = fv[3];
      const double iu = i4(U1[rix(tj + 2, tk)], U1[rix(tj + 1, tk)], U1[rix(tj, tk)],
                           U1[rix(tj - 1, tk)]);
      const double q1 = i4(fv[6], fv[5], fv[4], fv[3]) * iu;
      const double tx = d4(q1, vr1, vr0, vrm1, K.r24x);
      vrm1 = vr0;
      vr0 = vr1;
      vr1 = q1;
      const double y3 = Fo(3, 0), y2 = Fo(2, 0), y1 = Fo(1, 0);
      const double ym1 = Fo(-1, 0), ym2 = Fo(-2, 0), ym3 = Fo(-3, 0);
      const double z3 = Fo(0, 3), z2 = Fo(0, 2), z1 = Fo(0, 1);
      const double zm1 = Fo(0, -1), zm2 = Fo(0, -2), zm3 = Fo(0, -3);
      const double p2 = i4(y3, y2, y1, c0), p1 = i4(y2, y1, c0, ym1);
      const double p0 = i4(y1, c0, ym1, ym2), m1 = i4(c0, ym1, ym2, ym3);
      const double ty = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24y);
      auto wq = [&](int m) {
        return i4(Wv[rix(tj + 2, tk + m)], Wv[rix(tj + 1, tk + m)], Wv[rix(tj, tk + m)],
                  Wv[rix(tj - 1, tk + m)]);
      };
      const double tz = d4(i4(z3, z2, z1, c0) * wq(1), i4(z2, z1, c0, zm1) * wq(0),
                           i4(z1, c0, zm1, zm2) * wq(-1), i4(c0, zm1, zm2, zm3) * wq(-2),
                           K.r24z);
      const double nsum = tx + ty + tz;
      const double lap = l4(fv[5], fv[4], c0, fv[2], fv[1], K.ir12x) +
                         l4(y2, y1, c0, ym1, ym2, K.ir12y) +
                         l4(z2, z1, c0, zm1, zm2, K.ir12z);
      buf_st<double, BUF_NT>(prs(vs, i), st_v, 0, rk_update(coo_v, nsum, vis_cv, lap, K.dta));
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- w (convdiff_body<2, false>) ----
    {
      auto Fo = [&](int dj, int dk) { return W[rix(tj + dj, tk + dk)]; };
      const double c0 = fw[3];
      const double iu = i4(U1[rix(tj, tk + 2)], U1[rix(tj, tk + 1)], U1[rix(tj, tk)],
                           U1[rix(tj, tk - 1)]);
      const double q1 = i4(fw[6], fw[5], fw[4], fw[3]) * iu;
      const double tx = d4(q1, wr1, wr0, wrm1, K.r24x);
      wrm1 = wr0;
      wr0 = wr1;
      wr1 = q1;
      const double y3 = Fo(3, 0), y2 = Fo(2, 0), y1 = Fo(1, 0);
      const double ym1 = Fo(-1, 0), ym2 = Fo(-2, 0), ym3 = Fo(-3, 0);
      const double z3 = Fo(0, 3), z2 = Fo(0, 2), z1 = Fo(0, 1);
      const double zm1 = Fo(0, -1), zm2 = Fo(0, -2), zm3 = Fo(0, -3);
      auto vq = [&](int m) {
        return i4(V[rix(tj + m, tk + 2)], V[rix(tj + m, tk + 1)], V[rix(tj + m, tk)],
                  V[rix(tj + m, tk - 1)]);
      };
      const double ty = d4(i4(y3, y2, y1, c0) * vq(1), i4(y2, y1, c0, ym1) * vq(0),
                           i4(y1, c0, ym1, ym2) * vq(-1), i4(c0, ym1, ym2, ym3) * vq(-2),
                           K.r24y);
      const double p2 = i4(z3, z2, z1, c0), p1 = i4(z2, z1, c0, zm1);
      const double p0 = i4(z1, c0, zm1, zm2), m1 = i4(c0, zm1, zm2, zm3);
      const double tz = d4(p2 * p2, p1 * p1, p0 * p0, m1 * m1, K.r24z);
      const double nsum = tx + ty + tz;
      const double lap = l4(fw[5], fw[4], c0, fw[2], fw[1], K.ir12x) +
                         l4(y2, y1, c0, ym1, ym2, K.ir12y) +
                         l4(z2, z1, c0, zm1, zm2, K.ir12z);
      buf_st<double, BUF_NT>(prs(ws, i), st_w, 0, rk_update(coo_w, nsum, vis_cw, lap, K.dta));
    }
  }
}

__global__ void __launch_bounds__(256)
convdiff_vw_kernel(const double* __restrict__ u, const double* __restrict__ v,
                   const double* __restrict__ w, const double* __restrict__ voo,
                   const double* __restrict__ woo, const double* __restrict__ vis,
                   const double* __restrict__ wpre, double* __restrict__ vs,
                   double* __restrict__ ws, wf_grid g, int ich, int kt, int jt, Coefs K) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  convdiff_vw_body(u, v, w, voo, woo, vis, wpre, vs, ws, g, ich, L % kt, (L / kt) % jt,
                   L / (kt * jt), K);
}

// w_pre <- w at the cells convdiff_vw reads from w_pre: on each real plane,
// rows -1 and nj (columns -2 .. nk) and columns -2, -1, nk-1, nk (rows 0 ..
// nj-1), in real-cell coordinates.
__device__ inline void save_cross_body(const double* __restrict__ w, double* __restrict__ wpre,
                                       const wf_grid& g, int t, int plane) {
  const int nrow = g.nk + 3, n = 2 * nrow + 4 * g.nj;
  if (t >= n) return;
  int J, Kc;
  if (t < 2 * nrow) {
    J = t < nrow ? -1 : g.nj;
    Kc = t % nrow - 2;
  } else {
    const int h = t - 2 * nrow, q = h / g.nj;
    J = h % g.nj;
    Kc = q < 2 ? q - 2 : g.nk - 1 + (q - 2);
  }
  const int64_t x = (int64_t)(g.npl + plane) * g.si + (int64_t)(g.npl + J) * g.sj +
                    (g.npl + Kc);
  wpre[x] = w[x];
}
__global__ void __launch_bounds__(256)
save_cross_ghosts_kernel(const double* __restrict__ w, double* __restrict__ wpre, wf_grid g) {
  save_cross_body(w, wpre, g, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// Ensemble form: a persistent grid over the (member, tile, chunk) items of
// component C, chunks of BATCH_CD_ICH planes.
template <int C, bool SAVE>
__global__ void __launch_bounds__(256)
batch_convdiff_kernel(BView V, Plan P, const double* __restrict__ u,
                      const double* __restrict__ v, const double* __restrict__ w,
                      const double* __restrict__ coo, const double* __restrict__ vis,
                      double* __restrict__ cstar, double alpha, double* __restrict__ coo_save) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    const int q = it - P.start[m];
    const MemDev& M = V.mem[m];
    const wf_grid g = M.g;
    int d0, d1, d2;
    op_dims(C, g.ni, g.nj, g.nk, d0, d1, d2);
    const Box3 b{g.npl, g.npl, g.npl, d0, d1, d2};
    const int kt = (d2 + TK - 1) / TK, jt = (d1 + TJ - 1) / TJ;
    Coefs K;  // as convdiff_launch forms them on the host
    K.r24x = 1.0 / (24.0 * M.dx);
    K.r24y = 1.0 / (24.0 * M.dy);
    K.r24z = 1.0 / (24.0 * M.dz);
    K.ir12x = 1.0 / (12.0 * M.dx * M.dx);
    K.ir12y = 1.0 / (12.0 * M.dy * M.dy);
    K.ir12z = 1.0 / (12.0 * M.dz * M.dz);
    K.dta = M.dt * alpha;
    convdiff_body<C, SAVE>(u + M.off, v + M.off, w + M.off, SAVE ? nullptr : coo + M.off,
                           vis + M.off, cstar + M.off, g, b, BATCH_CD_ICH, q % kt,
                           (q / kt) % jt, q / (kt * jt), K, SAVE ? coo_save + M.off : nullptr);
    __syncthreads();  // LDS of this item is free for the next
  }
}

// Coefficients of member M (as convdiff_launch forms them on the host).
__device__ inline Coefs member_coefs(const MemDev& M, double alpha) {
  Coefs K;
  K.r24x = 1.0 / (24.0 * M.dx);
  K.r24y = 1.0 / (24.0 * M.dy);
  K.r24z = 1.0 / (24.0 * M.dz);
  K.ir12x = 1.0 / (12.0 * M.dx * M.dx);
  K.ir12y = 1.0 / (12.0 * M.dy * M.dy);
  K.ir12z = 1.0 / (12.0 * M.dz * M.dz);
  K.dta = M.dt * alpha;
  return K;
}

// Ensemble forms of the three-component passes: a persistent grid over the
// (member, tile, chunk) items of PK_WALE (8 x 32 tiles over the real cells,
// chunks of BATCH_CD_ICH planes), two workgroups per CU (the item loop's
// member lookup pushes the first stage's form past 168 VGPRs as well).
template <bool S2>
__device__ inline void batch_convdiff3_items(BView V, Plan P, const double* u, const double* v,
                                             const double* w, const double* vis,
                                             const double* wpre, double* us, double* vs,
                                             double* ws, double* uoo, double* voo, double* woo,
                                             double alpha) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    const int q = it - P.start[m];
    const MemDev& M = V.mem[m];
    const wf_grid g = M.g;
    const int kt = (g.nk + TK - 1) / TK, jt = (g.nj + TJ - 1) / TJ;
    convdiff3_body<S2>(u + M.off, v + M.off, w + M.off, vis + M.off, S2 ? wpre + M.off : nullptr,
                       us + M.off, vs + M.off, ws + M.off, uoo + M.off, voo + M.off, woo + M.off,
                       g, BATCH_CD_ICH, q % kt, (q / kt) % jt, q / (kt * jt),
                       member_coefs(M, alpha));
    __syncthreads();  // LDS of this item is free for the next
  }
}
__global__ void __launch_bounds__(256, 2)
batch_convdiff3_save_kernel(BView V, Plan P, const double* u, const double* v, const double* w,
                            const double* vis, double* us, double* vs, double* ws, double* uoo,
                            double* voo, double* woo, double alpha) {
  batch_convdiff3_items<false>(V, P, u, v, w, vis, nullptr, us, vs, ws, uoo, voo, woo, alpha);
}
__global__ void __launch_bounds__(256, 2)
batch_convdiff3_kernel(BView V, Plan P, const double* u, const double* v, const double* w,
                       const double* vis, const double* wpre, double* us, double* vs,
                       double* ws, double* uoo, double* voo, double* woo, double alpha) {
  batch_convdiff3_items<true>(V, P, u, v, w, vis, wpre, us, vs, ws, uoo, voo, woo, alpha);
}

// u's planes 0 and ni-2 of every member (PK_CDE items: a tile and one of the
// two planes, one-plane chunks).
__global__ void __launch_bounds__(256)
batch_convdiff_u_edges_kernel(BView V, Plan P, const double* __restrict__ u,
                              const double* __restrict__ v, const double* __restrict__ w,
                              const double* __restrict__ uoo, const double* __restrict__ vis,
                              double* __restrict__ us, double alpha) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    const int q = it - P.start[m];
    const MemDev& M = V.mem[m];
    const wf_grid g = M.g;
    const int d0 = g.ni - 1;
    const Box3 b{g.npl, g.npl, g.npl, d0, g.nj, g.nk};
    const int kt = (g.nk + TK - 1) / TK, jt = (g.nj + TJ - 1) / TJ;
    const int tile = q % (kt * jt), which = q / (kt * jt);
    convdiff_body<0, false>(u + M.off, v + M.off, w + M.off, uoo + M.off, vis + M.off,
                            us + M.off, g, b, 1, tile % kt, tile / kt, which ? d0 - 1 : 0,
                            member_coefs(M, alpha));
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
batch_save_cross_ghosts_kernel(BView V, Plan P, const double* __restrict__ w,
                               double* __restrict__ wpre) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    const int q = it - P.start[m];
    const MemDev& M = V.mem[m];
    const wf_grid g = M.g;
    const int nb = (2 * (g.nk + 3) + 4 * g.nj + 255) / 256;
    save_cross_body(w + M.off, wpre + M.off, g, (q % nb) * 256 + (int)threadIdx.x, q / nb);
  }
}

}  // namespace cdv

int batch_convdiff3(const Batch& B, const double* u, const double* v, const double* w,
                    const double* vis, const double* wpre, double* us, double* vs, double* ws,
                    double* uoo, double* voo, double* woo, double alpha, hipStream_t s) {
  using namespace cdv;
  const Plan& P = B.plan[PK_WALE];
  if (wpre)
    hipLaunchKernelGGL(batch_convdiff3_kernel, dim3(P.grid), dim3(NT), 0, s, B.view(), P, u, v,
                       w, vis, wpre, us, vs, ws, uoo, voo, woo, alpha);
  else
    hipLaunchKernelGGL(batch_convdiff3_save_kernel, dim3(P.grid), dim3(NT), 0, s, B.view(), P, u,
                       v, w, vis, us, vs, ws, uoo, voo, woo, alpha);
  WF_LAUNCHED();
  return 0;
}

int batch_convdiff_u_edges(const Batch& B, const double* u, const double* v, const double* w,
                           const double* uoo, const double* vis, double* us, double alpha,
                           hipStream_t s) {
  using namespace cdv;
  const Plan& P = B.plan[PK_CDE];
  if (P.total == 0) return 0;
  hipLaunchKernelGGL(batch_convdiff_u_edges_kernel, dim3(P.grid), dim3(NT), 0, s, B.view(), P,
                     u, v, w, uoo, vis, us, alpha);
  WF_LAUNCHED();
  return 0;
}

int batch_save_cross_ghosts(const Batch& B, const double* w, double* wpre, hipStream_t s) {
  using namespace cdv;
  const Plan& P = B.plan[PK_XG];
  hipLaunchKernelGGL(batch_save_cross_ghosts_kernel, dim3(P.grid), dim3(256), 0, s, B.view(), P,
                     w, wpre);
  WF_LAUNCHED();
  return 0;
}

int batch_convdiff(const Batch& B, int comp, const double* u, const double* v,
                   const double* w, const double* coo, const double* vis, double* cstar,
                   double alpha, hipStream_t s, double* coo_save) {
  using namespace cdv;
  const Plan& P = B.plan[PK_CD0 + comp];
#define BCD_LAUNCH(C_, S_)                                                                    \
  hipLaunchKernelGGL((batch_convdiff_kernel<C_, S_>), dim3(P.grid), dim3(NT), 0, s, B.view(), \
                     P, u, v, w, coo, vis, cstar, alpha, coo_save)
  if (coo_save) {
    if (comp == 0)
      BCD_LAUNCH(0, true);
    else if (comp == 1)
      BCD_LAUNCH(1, true);
    else
      BCD_LAUNCH(2, true);
  } else {
    if (comp == 0)
      BCD_LAUNCH(0, false);
    else if (comp == 1)
      BCD_LAUNCH(1, false);
    else
      BCD_LAUNCH(2, false);
  }
#undef BCD_LAUNCH
  WF_LAUNCHED();
  return 0;
}

// Launch of the fused operator on box b (checked by the caller).
int convdiff_launch(int comp, const double* u, const double* v, const double* w,
                    const double* coo, const double* vis, double* cstar, const wf_grid& g,
                    const Box3& b, double dx, double dy, double dz, double dta, int ich,
                    hipStream_t s, double* coo_save, int nch, int cstep) {
  using namespace cdv;
  const int kt = div_up(b.d2, TK), jt = div_up(b.d1, TJ);
  if (ich <= 0) ich = tune_get(WF_TUNE_CD_ICH, g.ni, g.nj, g.nk);
  if (ich <= 0) {  // aim for >= 2048 workgroups with at least 16 planes each
    const int64_t tiles = (int64_t)kt * jt;
    ich = (int)std::max<int64_t>(16, (tiles * b.d0 + 2047) / 2048);
  }
  if (nch <= 0) nch = div_up(b.d0, ich);
  if (cstep < 1 || (int64_t)(nch - 1) * cstep * ich >= b.d0) return WF_EINVAL;
  const int64_t nb = (int64_t)kt * jt * nch;
  if (nb >= (int64_t)1 << 31) return WF_EINVAL;
  Coefs K;
  K.r24x = 1.0 / (24.0 * dx);
  K.r24y = 1.0 / (24.0 * dy);
  K.r24z = 1.0 / (24.0 * dz);
  K.ir12x = 1.0 / (12.0 * dx * dx);
  K.ir12y = 1.0 / (12.0 * dy * dy);
  K.ir12z = 1.0 / (12.0 * dz * dz);
  K.dta = dta;
  // 1-D id over (k tile, j tile) per grid.x row, one row per i-chunk
  const dim3 grid((unsigned)(kt * jt), (unsigned)nch);
#define CD_LAUNCH(C_, S_)                                                                   \
  hipLaunchKernelGGL((convdiff_kernel<C_, S_>), grid, dim3(NT), 0, s, u, v, w, coo, vis, cstar, \
                     g, b, ich, kt, jt, K, coo_save, cstep)
  if (coo_save) {
    if (comp == 0)
      CD_LAUNCH(0, true);
    else if (comp == 1)
      CD_LAUNCH(1, true);
    else
      CD_LAUNCH(2, true);
  } else {
    if (comp == 0)
      CD_LAUNCH(0, false);
    else if (comp == 1)
      CD_LAUNCH(1, false);
    else
      CD_LAUNCH(2, false);
  }
#undef CD_LAUNCH
  WF_LAUNCHED();
  return 0;
}

int convdiff_vw_launch(const double* u, const double* v, const double* w, const double* voo,
                       const double* woo, const double* vis, const double* wpre, double* vs,
                       double* ws, const wf_grid& g, double dx, double dy, double dz,
                       double dta, int ich, hipStream_t s) {
  using namespace cdv;
  const int kt = div_up(g.nk, TK), jt = div_up(g.nj, TJ);
  if (ich <= 0) ich = tune_get(WF_TUNE_CDVW_ICH, g.ni, g.nj, g.nk);
  if (ich <= 0) {  // aim for >= 2048 workgroups with at least 16 planes each
    const int64_t tiles = (int64_t)kt * jt;
    ich = (int)std::max<int64_t>(16, (tiles * g.ni + 2047) / 2048);
  }
  const int nch = div_up(g.ni, ich);
  if ((int64_t)kt * jt * nch >= (int64_t)1 << 31) return WF_EINVAL;
  if (g.si * (int64_t)sizeof(double) >= ((int64_t)1 << 30)) return WF_EINVAL;  // buffer range
  Coefs K;
  K.r24x = 1.0 / (24.0 * dx);
  K.r24y = 1.0 / (24.0 * dy);
  K.r24z = 1.0 / (24.0 * dz);
  K.ir12x = 1.0 / (12.0 * dx * dx);
  K.ir12y = 1.0 / (12.0 * dy * dy);
  K.ir12z = 1.0 / (12.0 * dz * dz);
  K.dta = dta;
  hipLaunchKernelGGL(convdiff_vw_kernel, dim3((unsigned)(kt * jt), (unsigned)nch), dim3(NT), 0,
                     s, u, v, w, voo, woo, vis, wpre, vs, ws, g, ich, kt, jt, K);
  WF_LAUNCHED();
  return 0;
}

int save_cross_ghosts_launch(const double* w, double* wpre, const wf_grid& g, hipStream_t s) {
  using namespace cdv;
  const int n = 2 * (g.nk + 3) + 4 * g.nj;
  hipLaunchKernelGGL(save_cross_ghosts_kernel, dim3((unsigned)div_up(n, 256), (unsigned)g.ni),
                     dim3(256), 0, s, w, wpre, g);
  WF_LAUNCHED();
  return 0;
}

// The three-component pass: the first stage's form (wpre == nullptr, uoo, voo,
// woo written) or the second stage's (uoo, voo, woo read, wpre as in
// convdiff_vw_launch).
static int convdiff3_any_launch(const double* u, const double* v, const double* w,
                                const double* vis, const double* wpre, double* us, double* vs,
                                double* ws, double* uoo, double* voo, double* woo,
                                const wf_grid& g, double dx, double dy, double dz, double dta,
                                int ich, hipStream_t s) {
  using namespace cdv;
  const int kt = div_up(g.nk, TK), jt = div_up(g.nj, TJ);
  if (ich <= 0) ich = tune_get(wpre ? WF_TUNE_CD3S2_ICH : WF_TUNE_CD3_ICH, g.ni, g.nj, g.nk);
  if (ich <= 0) {  // one round of workgroups (three per CU), >= 16 planes each
    const int64_t tiles = (int64_t)kt * jt, slots = 768;
    ich = (int)std::max<int64_t>(16, (tiles * g.ni + slots - 1) / slots);
  }
  const int nch = div_up(g.ni, ich);
  if ((int64_t)kt * jt * nch >= (int64_t)1 << 31) return WF_EINVAL;
  if (g.si * (int64_t)sizeof(double) >= ((int64_t)1 << 30)) return WF_EINVAL;  // buffer range
  Coefs K;
  K.r24x = 1.0 / (24.0 * dx);
  K.r24y = 1.0 / (24.0 * dy);
  K.r24z = 1.0 / (24.0 * dz);
  K.ir12x = 1.0 / (12.0 * dx * dx);
  K.ir12y = 1.0 / (12.0 * dy * dy);
  K.ir12z = 1.0 / (12.0 * dz * dz);
  K.dta = dta;
  const dim3 grid((unsigned)(kt * jt), (unsigned)nch);
  if (wpre)
    hipLaunchKernelGGL(convdiff3_kernel, grid, dim3(NT), 0, s, u, v, w, vis, wpre, us, vs, ws,
                       uoo, voo, woo, g, ich, kt, jt, K);
  else
    hipLaunchKernelGGL(convdiff3_save_kernel, grid, dim3(NT), 0, s, u, v, w, vis, us, vs, ws,
                       uoo, voo, woo, g, ich, kt, jt, K);
  WF_LAUNCHED();
  return 0;
}

int convdiff3_save_launch(const double* u, const double* v, const double* w,
                          const double* vis, double* us, double* vs, double* ws, double* uoo,
                          double* voo, double* woo, const wf_grid& g, double dx, double dy,
                          double dz, double dta, int ich, hipStream_t s) {
  return convdiff3_any_launch(u, v, w, vis, nullptr, us, vs, ws, uoo, voo, woo, g, dx, dy, dz,
                              dta, ich, s);
}

int convdiff3_launch(const double* u, const double* v, const double* w, const double* uoo,
                     const double* voo, const double* woo, const double* vis,
                     const double* wpre, double* us, double* vs, double* ws, const wf_grid& g,
                     double dx, double dy, double dz, double dta, int ich, hipStream_t s) {
  if (!wpre) return WF_EINVAL;
  // read-only in the second stage's form
  return convdiff3_any_launch(u, v, w, vis, wpre, us, vs, ws, const_cast<double*>(uoo),
                              const_cast<double*>(voo), const_cast<double*>(woo), g, dx, dy, dz,
                              dta, ich, s);
}

// u's conv+diff on planes 0 and ni-2 of its operator box (the planes
// convdiff3_kernel leaves to this pass): one launch of one-plane chunks 0 and
// d0 - 1.
int convdiff_u_edges_launch(const double* u, const double* v, const double* w,
                            const double* uoo, const double* vis, double* us, const wf_grid& g,
                            double dx, double dy, double dz, double dta, hipStream_t s) {
  const int d0 = g.ni - 1;
  if (d0 <= 0) return 0;
  const Box3 b{g.npl, g.npl, g.npl, d0, g.nj, g.nk};
  return convdiff_launch(0, u, v, w, uoo, vis, us, g, b, dx, dy, dz, dta, 1, s, nullptr,
                         d0 > 1 ? 2 : 1, d0 > 1 ? d0 - 1 : 1);
}

}  // namespace wf
