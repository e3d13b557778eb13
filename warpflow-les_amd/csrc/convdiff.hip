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
// (exact).
//
// Workgroup order.  Consecutive workgroup ids are dispatched round-robin over
// the 8 XCDs, each with its own L2; the ids are remapped so that each XCD
// walks a contiguous run of tiles (k fastest, then j, then the i-chunk) and
// neighbouring tiles share their halo rows in the same L2.
#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "wf_common.h"

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
  return (-p1 + 9.0 * c0 + 9.0 * m1 - m2) * 0.0625;
}
// (-s[+] + 27 s[0] - 27 s[-] + s[--]) / (24 h)  (kernels.py:69)
__device__ inline double d4(double sp, double sn, double sm, double smm, double r24h) {
  return (-sp + 27.0 * sn - 27.0 * sm + smm) * r24h;
}
// (-f[+2] + 16 f[+1] - 30 f[0] + 16 f[-1] - f[-2]) / (12 h^2)  (kernels.py:177-186)
__device__ inline double l4(double p2, double p1, double c, double m1, double m2,
                            double ir12) {
  return (-p2 + 16.0 * p1 - 30.0 * c + 16.0 * m1 - m2) * ir12;
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
        double t = cur_coo - K.dta * nsum;
        t = t + K.dta * (cur_vis * lap);
        cstar[G.plane(i) + own.off] = t;
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
        double t = cur_coo - K.dta * nsum;
        t = t + K.dta * (cur_vis * lap);
        cstar[G.plane(i) + own.off] = t;
        if constexpr (SAVE) coo_save[G.plane(i) + own.off] = cur_coo;
      }
    }
  }
}

template <int C, bool SAVE>
__global__ void __launch_bounds__(256)
convdiff_kernel(const double* __restrict__ u, const double* __restrict__ v,
                const double* __restrict__ w, const double* __restrict__ coo,
                const double* __restrict__ vis, double* __restrict__ cstar, wf_grid g, Box3 b,
                int ich, int kt, int jt, Coefs K, double* __restrict__ coo_save) {
  const int nb = kt * jt * gridDim.y;
  const int L = xcd_linear(blockIdx.x + blockIdx.y * gridDim.x, nb);
  convdiff_body<C, SAVE>(u, v, w, coo, vis, cstar, g, b, ich, L % kt, (L / kt) % jt,
                         L / (kt * jt), K, coo_save);
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

}  // namespace cdv

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
                    hipStream_t s, double* coo_save) {
  using namespace cdv;
  const int kt = div_up(b.d2, TK), jt = div_up(b.d1, TJ);
  if (ich <= 0) ich = tune_get(WF_TUNE_CD_ICH, g.ni, g.nj, g.nk);
  if (ich <= 0) {  // aim for >= 2048 workgroups with at least 16 planes each
    const int64_t tiles = (int64_t)kt * jt;
    ich = (int)std::max<int64_t>(16, (tiles * b.d0 + 2047) / 2048);
  }
  const int nch = div_up(b.d0, ich);
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
                     g, b, ich, kt, jt, K, coo_save)
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

}  // namespace wf
