// Geometric multigrid for the cell-centred 7-point Neumann Poisson problem
// (the reference's warpflow/multigrid.py), matrix-free, templated on the
// level precision T (double for the plain FP64 form, float for the V-cycle
// of the mixed form).
//
// Level arrays are (n0+2, n1+2, n2+2) with a one-cell ring.  Face
// coefficients are generated from the cell index: 1/h^2, or 0 across a
// physical face (all six faces of the single cavity domain), and
// ap = -(aw+ae+as+an+ab+at) -- the values _mgcoeff stores
// (multigrid.py:23-67), without the seven coefficient arrays.
//
// Ring state.  The reference refreshes the ring faces (Neumann mirror,
// faces only) after every full sweep and after every prolongation, zeroes the
// whole coarse array after restriction, and never writes edge/corner ring
// cells.  Smoothing and restriction multiply ring values by zero
// coefficients; only prolongation and the final copy-out read them.  Here
// the kernel that last updates a boundary cell writes its face-ring images
// (relax, prolong), restriction zeroes them, and edges/corners stay at the
// zero they were allocated with -- the same ring state at every point the
// reference could observe it.
#include <algorithm>
#include <type_traits>

#include "batch.h"
#include "mg_common.h"

namespace wf {

// Thread -> (i, j, k) of colour `color` cells: k = 2m + ((i+j+color+1) & 1),
// i.e. (i+j+k+color) odd (multigrid.py:80).
__device__ inline bool colour_cell(const MgDev& d, int color, int& i, int& j, int& k) {
  const int m = blockIdx.x * 64 + threadIdx.x;
  j = blockIdx.y * 4 + threadIdx.y;
  i = blockIdx.z;
  if (j >= d.n1 || i >= d.n0) return false;
  k = 2 * m + ((i + j + color + 1) & 1);
  return k < d.n2;
}

template <class T>
__global__ void __launch_bounds__(256)
rbgs_kernel(T* __restrict__ phi, const T* __restrict__ rhs, MgDev d, T c0, T c1, T c2,
            int color, int write_ring) {
  int i, j, k;
  if (!colour_cell(d, color, i, j, k)) return;
  const int64_t x = mg_at(d, i, j, k);
  const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
  const T v = (rhs[x] - nb_sum(phi, x, d, c)) * (T(1) / c.ap);
  phi[x] = v;
  if (write_ring) ring_images(phi, d, i, j, k, x, v);
}

// Faces-only Neumann mirror of all six sides (bcs.py:326-360 via
// multigrid.py:264-287); one thread per (face, a, b).
template <class T>
__global__ void __launch_bounds__(256)
refill_kernel(T* __restrict__ phi, MgDev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int fx = d.n1 * d.n2, fy = d.n0 * d.n2, fz = d.n0 * d.n1;
  int q = t;
  if (q < 2 * fx) {
    const int hi = q >= fx;
    q -= hi * fx;
    const int j = q / d.n2, k = q % d.n2;
    const int i = hi ? d.n0 - 1 : 0;
    const int64_t x = mg_at(d, i, j, k);
    phi[x + (hi ? d.s0 : -d.s0)] = phi[x];
    return;
  }
  q -= 2 * fx;
  if (q < 2 * fy) {
    const int hi = q >= fy;
    q -= hi * fy;
    const int i = q / d.n2, k = q % d.n2;
    const int j = hi ? d.n1 - 1 : 0;
    const int64_t x = mg_at(d, i, j, k);
    phi[x + (hi ? d.s1 : -d.s1)] = phi[x];
    return;
  }
  q -= 2 * fy;
  if (q < 2 * fz) {
    const int hi = q >= fz;
    q -= hi * fz;
    const int i = q / d.n1, j = q % d.n1;
    const int k = hi ? d.n2 - 1 : 0;
    const int64_t x = mg_at(d, i, j, k);
    phi[x + (hi ? 1 : -1)] = phi[x];
  }
}

// Residual rhs - A phi at a fine real cell, ap*phi added last
// (multigrid.py:110-117).
template <class T>
__device__ inline T residual_at(const T* __restrict__ phi, const T* __restrict__ rhs,
                                const MgDev& d, int i, int j, int k, T c0, T c1, T c2) {
  const int64_t x = mg_at(d, i, j, k);
  const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
  return rhs[x] - (nb_sum(phi, x, d, c) + c.ap * phi[x]);
}

// _restrict (multigrid.py:93-119) + coarse phi = 0 (multigrid.py:312).
template <class T>
__device__ inline void restrict_body(const T* __restrict__ phif, const T* __restrict__ rhsf,
                                     const MgDev& f, T c0, T c1, T c2, T* __restrict__ phic,
                                     T* __restrict__ rhsc, const MgDev& c, Blk B) {
  const int ck = B.x * 64 + threadIdx.x;
  const int cj = B.y * 4 + threadIdx.y;
  const int ci = B.z;
  if (ck >= c.n2 || cj >= c.n1 || ci >= c.n0) return;
  T acc = T(0);
  for (int di = 0; di < 2; ++di)
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk)
        acc = acc + residual_at(phif, rhsf, f, 2 * ci + di, 2 * cj + dj, 2 * ck + dk, c0, c1,
                                c2);
  const int64_t x = mg_at(c, ci, cj, ck);
  rhsc[x] = acc * T(0.125);
  phic[x] = T(0);
  ring_images(phic, c, ci, cj, ck, x, T(0));
}
template <class T>
__global__ void __launch_bounds__(256)
restrict_kernel(const T* __restrict__ phif, const T* __restrict__ rhsf, MgDev f, T c0, T c1,
                T c2, T* __restrict__ phic, T* __restrict__ rhsc, MgDev c, int ich) {
  for_planes(c.n0, ich,
             [&](Blk B) { restrict_body(phif, rhsf, f, c0, c1, c2, phic, rhsc, c, B); });
}

// _prolong (multigrid.py:122-146) + fine refill (multigrid.py:321).
template <class T>
__device__ inline void prolong_body(const T* __restrict__ phic, const MgDev& c,
                                    T* __restrict__ phif, const MgDev& f, Blk B) {
  const int k = B.x * 64 + threadIdx.x;
  const int j = B.y * 4 + threadIdx.y;
  const int i = B.z;
  if (k >= f.n2 || j >= f.n1 || i >= f.n0) return;
  const int c0x = (i % 2 == 0) ? i / 2 - 1 : i / 2;
  const int c0y = (j % 2 == 0) ? j / 2 - 1 : j / 2;
  const int c0z = (k % 2 == 0) ? k / 2 - 1 : k / 2;
  const T wx0 = (i % 2 == 0) ? T(0.25) : T(0.75);
  const T wy0 = (j % 2 == 0) ? T(0.25) : T(0.75);
  const T wz0 = (k % 2 == 0) ? T(0.25) : T(0.75);
  T acc = T(0);
  for (int di = 0; di < 2; ++di) {
    const T wxi = di == 0 ? wx0 : T(1) - wx0;
    for (int dj = 0; dj < 2; ++dj) {
      const T wyj = dj == 0 ? wy0 : T(1) - wy0;
      for (int dk = 0; dk < 2; ++dk) {
        const T wzk = dk == 0 ? wz0 : T(1) - wz0;
        // coarse array index of real cell (c0+d) is (c0+d+1): the ring at -1
        const int64_t xc = (int64_t)(c0x + di + 1) * c.s0 + (int64_t)(c0y + dj + 1) * c.s1 +
                           (c0z + dk + 1);
        acc = acc + wxi * wyj * wzk * phic[xc];
      }
    }
  }
  const int64_t x = mg_at(f, i, j, k);
  const T v = phif[x] + acc;
  phif[x] = v;
  ring_images(phif, f, i, j, k, x, v);
}
template <class T>
__global__ void __launch_bounds__(256)
prolong_kernel(const T* __restrict__ phic, MgDev c, T* __restrict__ phif, MgDev f, int ich) {
  for_planes(f.n0, ich, [&](Blk B) { prolong_body(phic, c, phif, f, B); });
}

// The same prolongation by 2 x 2 x 2 fine blocks (fine levels of even
// extents, each twice the coarse one): a thread owns the fine cells
// (2a+ei, 2b+ej, 2c+ek) of coarse cell (a, b, c), whose eight interpolations
// read the 27 coarse cells a-1..a+1 x b-1..b+1 x c-1..c+1 (loaded once
// instead of 64 times), and the weights of each fine cell are constants.
// Every fine value is prolong_body's: the same products, summed in the same
// order, each term one FMA.
template <class T>
__device__ inline void prolong8_body(const T* __restrict__ phic, const MgDev& c,
                                     T* __restrict__ phif, const MgDev& f, Blk B) {
  const int ck = B.x * 64 + threadIdx.x;
  const int cj = B.y * 4 + threadIdx.y;
  const int ci = B.z;
  if (ck >= c.n2 || cj >= c.n1 || ci >= c.n0) return;
  // coarse real cell (ci-1+di, cj-1+dj, ck-1+dk) sits at array index
  // (ci+di, cj+dj, ck+dk): the ring at -1
  T cc[3][3][3];
#pragma unroll
  for (int di = 0; di < 3; ++di)
#pragma unroll
    for (int dj = 0; dj < 3; ++dj)
#pragma unroll
      for (int dk = 0; dk < 3; ++dk)
        cc[di][dj][dk] = phic[(int64_t)(ci + di) * c.s0 + (int64_t)(cj + dj) * c.s1 + (ck + dk)];
#pragma unroll
  for (int ei = 0; ei < 2; ++ei) {
    const int i = 2 * ci + ei;
    const T wx0 = ei == 0 ? T(0.25) : T(0.75);
#pragma unroll
    for (int ej = 0; ej < 2; ++ej) {
      const int j = 2 * cj + ej;
      const T wy0 = ej == 0 ? T(0.25) : T(0.75);
#pragma unroll
      for (int ek = 0; ek < 2; ++ek) {
        const int k = 2 * ck + ek;
        const T wz0 = ek == 0 ? T(0.25) : T(0.75);
        // fine parity e: the first coarse cell of the pair is a-1+e
        T acc = T(0);
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          const T wxi = di == 0 ? wx0 : T(1) - wx0;
#pragma unroll
          for (int dj = 0; dj < 2; ++dj) {
            const T wyj = dj == 0 ? wy0 : T(1) - wy0;
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
              const T wzk = dk == 0 ? wz0 : T(1) - wz0;
              // explicit FMA: prolong_body's terms contract to this, and a
              // product shared by several fine cells must not be rounded
              // once and reused instead
              acc = fma(wxi * wyj * wzk, cc[ei + di][ej + dj][ek + dk], acc);
            }
          }
        }
        const int64_t x = mg_at(f, i, j, k);
        const T v = phif[x] + acc;
        phif[x] = v;
        ring_images(phif, f, i, j, k, x, v);
      }
    }
  }
}
template <class T>
__global__ void __launch_bounds__(256)
prolong8_kernel(const T* __restrict__ phic, MgDev c, T* __restrict__ phif, MgDev f, int ich) {
  for_planes(c.n0, ich, [&](Blk B) { prolong8_body(phic, c, phif, f, B); });
}

// Buffer access to a level array (buf_rsrc, buf_ld, buf_st in wf_common.h):
// the base, chosen per workgroup, sits in a scalar buffer descriptor, the
// plane offset in the scalar soffset operand, and each lane supplies a 32-bit
// byte offset within the plane, so a load or store costs no vector address
// arithmetic.  A lane offset of SWEEP_DEAD (or SWEEP_DEAD plus or minus one
// row) lies past every descriptor's range -- the host keeps a chunk's plane
// offsets below 2^30 -- so its load returns zero and its store is dropped,
// which is how cells outside the domain are masked.
constexpr uint32_t SWEEP_DEAD = BUF_DEAD;

// One full RBGS sweep (colour 0, then colour 1, then the ring refill) out of
// place: pout = S(pin).  Colour-0 (odd i+j+k) cells read only even cells of
// pin; colour-1 (even) cells read only the new odd values.  A workgroup owns a
// TJ x TK (j, k) tile (TJ*TK/2 (even, odd) pairs, NPAIR per thread) and
// streams along i over a chunk of planes.  LDS holds, for three planes, the
// even cells of pin over the tile + 2-cell halo and the new odd values over
// the tile + 1-cell halo, each packed checkerboard-wise (a row keeps only its
// cells of the one parity, so neighbouring lanes touch neighbouring words).
// At plane q the workgroup computes the odd values of q, then the even values
// of q-1, and writes plane q-1 of pout (both colours + face-ring images).
// Global loads run two planes ahead of their use.
//
// Coefficients.  A neighbour outside the domain carries a zero coefficient in
// _mgcoeff; here its value enters LDS as zero instead (pin cells outside the
// domain, odd cells outside the domain: their reciprocal diagonal is zero),
// so every cell uses the uniform coefficients c0, c1, c2 in the same order of
// terms, and only the diagonal depends on the cell.  A zero product adds
// nothing either way, so the sums are the coefficient-array sums bit for bit.
// The reciprocal diagonal of each cell the thread owns is computed once for
// the interior planes; the planes at the i faces recompute it.
//
// Plane parity.  Which k of a row holds a colour alternates with the plane;
// the loop is unrolled over six planes (3 LDS slots x 2 parities) so every
// per-parity offset is a compile-time choice between two registers, and
// every iteration issues the same memory instructions (loads that are not
// needed read a valid plane and are ignored), so the waits on the prefetched
// planes stay exact.  HBM traffic per sweep: one read of pin and rhs, one
// write of pout.
template <class T, int TJ, int TK, int NB = 0>
__device__ inline void sweep_body(const T* __restrict__ pin, T* __restrict__ pout,
                                  const T* __restrict__ rhs, const MgDev& d, T c0, T c1, T c2,
                                  int ich, Blk B) {
  constexpr int NPAIR = TJ * TK / 512;  // (even, odd) pairs per thread
  static_assert(TJ * TK == 512 * NPAIR && (NPAIR == 1 || NPAIR == 2), "256 threads, 1-2 pairs");
  static_assert(TK % 2 == 0, "pairs along k");
  constexpr int PJ = TJ + 4, PH = (TK + 4) / 2;  // pin halo rows, even cells per row
  constexpr int OJ = TJ + 2, OH = (TK + 2) / 2;  // odd halo rows, odd cells per row
  constexpr int HK = TK / 2;                     // pairs per tile row
  constexpr int NPE = (PJ * PH + 255) / 256;     // pin cells per thread per plane
  constexpr int NO = (OJ * OH + 255) / 256;      // odd cells per thread per plane
  // plane slots: three (a ring) when pipelined, every plane of the chunk in
  // the one-pass chunks (NB > 0)
  constexpr int QP = NB ? NB + 4 : 3, QO = NB ? NB + 2 : 3;
  __shared__ T P[QP][NPE * 256];
  __shared__ T O[QO][NO * 256];
  const int tid = threadIdx.x;
  const int j0 = B.y * TJ, k0 = B.x * TK;
  const int i0 = B.z * ich, i1 = min(i0 + ich, d.n0);
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int s1 = (int)d.s1;
  const int sb = (i0 - 1 + j0 + k0) & 1;  // parity class of plane i0-1; index r: sb ^ r
  const int imid = n0 >= 3 ? 1 : 0;       // an interior plane (unused when n0 < 3)
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
  // byte offset of (j, k) inside one level plane
  auto plane_boff = [&](int j, int k) {
    return (uint32_t)(((j + 1) * s1 + (k + 1)) * (int)sizeof(T));
  };
  auto in_jk = [&](int j, int k) { return j >= 0 && j < n1 && k >= 0 && k < n2; };
  auto rdiag = [&](int i, int j, int k) {
    return T(1) / mg_coef(d, i, clampi(j, 0, n1 - 1), clampi(k, 0, n2 - 1), c0, c1, c2).ap;
  };
  // odd cell l = tid + 256 e of the 1-halo tile: row jr, packed column m;
  // on parity class s its k is k0 - 1 + 2m + ph, ph = (s + jr + 1) & 1
  auto odd_cell = [&](int t, int e, int r, int& j, int& k) {
    const int l = min(t + 256 * e, OJ * OH - 1);
    const int jr = l / OH, m = l % OH;
    j = j0 - 1 + jr;
    k = k0 - 1 + 2 * m + ((sb ^ r) + jr + 1) % 2;
    return t + 256 * e < OJ * OH && in_jk(j, k);
  };
  // pair p = tid + 256 e: row jr, k_a = k0 + 2 mm; the even cell is k_a + kx
  auto pair_cell = [&](int t, int e, int r, int& j, int& ke, int& ko) {
    const int p = t + 256 * e;
    const int jr = p / HK, mm = p % HK;
    const int kx = ((sb ^ r) + jr) % 2;
    j = j0 + jr;
    ke = k0 + 2 * mm + kx;
    ko = k0 + 2 * mm + 1 - kx;
  };
  // Every tile cell is a real cell off the j and k faces: no j/k ring images.
  const bool wfast = j0 >= 1 && j0 + TJ <= n1 - 1 && k0 >= 1 && k0 + TK <= n2 - 1;

  // pin cells: LDS word w = tid + 256 e = row jr, packed column h; on class s
  // the cell is (j0 - 2 + jr, k0 - 2 + 2h + ((s + jr) & 1))
  uint32_t gofs[2][NPE];
#pragma unroll
  for (int e = 0; e < NPE; ++e) {
    const int w = tid + 256 * e;
    const int wc = min(w, PJ * PH - 1);
    const int jr = wc / PH, h = wc % PH;
    const int j = j0 - 2 + jr;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int k = k0 - 2 + 2 * h + ((sb ^ r) + jr) % 2;
      gofs[r][e] = w < PJ * PH && in_jk(j, k) ? plane_boff(j, k) : SWEEP_DEAD;
    }
  }
  // odd cells: rhs offsets, LDS columns (k neighbours at a0, a0+1 of pin row
  // jr+1; j and i neighbours at column aj), interior-plane 1/ap (0 if dead)
  uint32_t oofs[2][NO];
  int a0[NO], aj[2][NO];
  T rapo[2][NO];
#pragma unroll
  for (int e = 0; e < NO; ++e) {
    const int l = min(tid + 256 * e, OJ * OH - 1);
    const int jr = l / OH, m = l % OH;
    a0[e] = (jr + 1) * PH + m;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int j, k;
      const bool live = odd_cell(tid, e, r, j, k);
      oofs[r][e] = live ? plane_boff(j, k) : SWEEP_DEAD;
      aj[r][e] = a0[e] + ((sb ^ r) + jr + 1) % 2;
      rapo[r][e] = live ? rdiag(imid, j, k) : T(0);
    }
  }
  // pairs: offsets of the even and the odd cell (rhs and output), LDS columns
  // (k neighbours of the even cell at b0, b0+1 of odd row jr+1, j and i
  // neighbours at column bj, the odd partner at bo), interior-plane 1/ap
  // Face flags of the pair's cells a = k_a and b = k_a + 1 (set for live
  // cells only): FL_JLO/FL_JHI j == 0 / n1-1, FL_ALO/FL_AHI k_a == 0 / n2-1,
  // FL_BHI k_a + 1 == n2-1; FL_X: the even cell is b on class 0.
  constexpr int FL_JLO = 1, FL_JHI = 2, FL_ALO = 4, FL_AHI = 8, FL_BHI = 16, FL_AL = 32,
                FL_BL = 64, FL_X = 128;
  uint32_t eofs[2][NPAIR], pofs[2][NPAIR];
  int b0[NPAIR], bj[2][NPAIR], bo[2][NPAIR], fl[NPAIR];
  T rape[2][NPAIR];
#pragma unroll
  for (int e = 0; e < NPAIR; ++e) {
    const int p = tid + 256 * e;
    const int jr = p / HK, mm = p % HK;
    b0[e] = (jr + 1) * OH + mm;
    {
      const int j = j0 + jr, ka = k0 + 2 * mm;
      const bool la = j < n1 && ka < n2, lb = j < n1 && ka + 1 < n2;
      fl[e] = (la || lb ? (j == 0 ? FL_JLO : 0) | (j == n1 - 1 ? FL_JHI : 0) : 0) |
              (la ? (ka == 0 ? FL_ALO : 0) | (ka == n2 - 1 ? FL_AHI : 0) | FL_AL : 0) |
              (lb ? (ka + 1 == n2 - 1 ? FL_BHI : 0) | FL_BL : 0) |
              ((sb + jr) % 2 ? FL_X : 0);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int j, ke, ko;
      pair_cell(tid, e, r, j, ke, ko);
      const int kx = ke - (k0 + 2 * mm);
      eofs[r][e] = j < n1 && ke < n2 ? plane_boff(j, ke) : SWEEP_DEAD;
      pofs[r][e] = j < n1 && ko < n2 ? plane_boff(j, ko) : SWEEP_DEAD;
      bj[r][e] = b0[e] + kx;
      bo[r][e] = b0[e] + 1 - kx;
      rape[r][e] = rdiag(imid, j, min(ke, n2 - 1));
    }
  }

  // The rare i-face paths recompute their cells from an opaque copy of the
  // thread index, so the compiler keeps none of that work live across the
  // loop.
  auto opaque_tid = [&]() {
    int t = tid;
    asm volatile("" : "+v"(t));
    return t;
  };
  // j and k face-ring images of a pair's cells a and b (class rn)
  auto jk_images = [&](const auto& dst, int e, int rn, int so, T ve, T vo, uint32_t oe,
                       uint32_t oo) {
    const int f = fl[e];
    const bool xbf = ((f & FL_X) != 0) != (rn != 0);  // the even cell is b
    const uint32_t oa = xbf ? oo : oe, ob = xbf ? oe : oo;
    const T va = xbf ? vo : ve, vb = xbf ? ve : vo;
    const uint32_t sj = (uint32_t)(s1 * (int)sizeof(T)), sk = (uint32_t)sizeof(T);
    const uint32_t jd = (f & FL_JLO) ? 0u - sj : sj;
    const bool jf = (f & (FL_JLO | FL_JHI)) != 0;
    buf_st(dst, jf && (f & FL_AL) ? oa + jd : SWEEP_DEAD, so, va);
    buf_st(dst, jf && (f & FL_BL) ? ob + jd : SWEEP_DEAD, so, vb);
    buf_st(dst, (f & FL_ALO) ? oa - sk : (f & FL_AHI) ? oa + sk : SWEEP_DEAD, so, va);
    buf_st(dst, (f & FL_BHI) ? ob + sk : SWEEP_DEAD, so, vb);
    if (n1 == 1) {  // both j faces: the high images too
      buf_st(dst, (f & FL_AL) ? oa + sj : SWEEP_DEAD, so, va);
      buf_st(dst, (f & FL_BL) ? ob + sj : SWEEP_DEAD, so, vb);
    }
    if (n2 == 1) buf_st(dst, (f & FL_AL) ? oa + sk : SWEEP_DEAD, so, va);
  };
  // One-pass chunks (NB > 0), for the small levels, whose pipelined
  // launches are a few dependent plane steps long: every load the chunk
  // needs is issued at once -- pin's even cells of planes i0-2 .. i1+1 (tile
  // + 2-cell halo), the odd cells' rhs of planes i0-1 .. i1 and the even
  // cells' rhs of planes i0 .. i1-1.  Then the odd values of planes
  // i0-1 .. i1 (tile + 1-cell halo; the neighbouring chunks compute the two
  // outer planes too), a barrier, and the even values of planes i0 .. i1-1
  // with the write.  Cells, terms and their order are the pipelined path's,
  // so the result is its bit for bit; a launch waits for one memory round
  // trip instead of one per plane step.
  if constexpr (NB > 0) {
    // Descriptors start at the array plane of real plane xb = max(i0-2, -1);
    // a plane x is read at soffset (clamp(x, -1, n0) - xb) * pb (planes past
    // the array are read clamped and their values ignored or zeroed).
    const int pb = (int)(d.s0 * (int64_t)sizeof(T));
    const int xb = max(i0 - 2, -1);
    auto chunk_rsrc = [&](const T* a) {
      return buf_rsrc(reinterpret_cast<const char*>(a) + (int64_t)(xb + 1) * pb, 0x7fffffffu);
    };
    const auto dpin = pin ? chunk_rsrc(pin) : buf_rsrc(rhs, 0u);
    const auto drhs = chunk_rsrc(rhs), dout = chunk_rsrc(pout);
    auto soff = [&](int x) { return (clampi(x, -1, n0) - xb) * pb; };

    // every load of the chunk
    T pre[QP][NPE], rodd[QO][NO], reven[NB][NPAIR];
#pragma unroll
    for (int p = 0; p < QP; ++p) {  // pin plane x = i0 - 2 + p, class (p + 1) & 1
      const int so = soff(i0 - 2 + p);
#pragma unroll
      for (int e = 0; e < NPE; ++e) pre[p][e] = buf_ld<T>(dpin, gofs[(p + 1) & 1][e], so);
    }
#pragma unroll
    for (int p = 0; p < QO; ++p) {  // odd plane x = i0 - 1 + p, class p & 1
      const int so = soff(i0 - 1 + p);
#pragma unroll
      for (int e = 0; e < NO; ++e) rodd[p][e] = buf_ld<T>(drhs, oofs[p & 1][e], so);
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {  // even plane x = i0 + p, class (p + 1) & 1
      const int so = soff(i0 + p);
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) reven[p][e] = buf_ld<T>(drhs, eofs[(p + 1) & 1][e], so);
    }
    // pin to LDS; the ring planes -1 and n0 (and beyond) enter as zero
#pragma unroll
    for (int p = 0; p < QP; ++p) {
      const int x = i0 - 2 + p;
      const bool z = x < 0 || x >= n0;
#pragma unroll
      for (int e = 0; e < NPE; ++e) P[p][tid + 256 * e] = z ? T(0) : pre[p][e];
    }
    __syncthreads();
    // odd values of planes i0-1 .. i0+NB (those past i1 are not read)
#pragma unroll
    for (int p = 0; p < QO; ++p) {
      const int q = i0 - 1 + p;
      const int r = p & 1;
      T rp[NO];
#pragma unroll
      for (int e = 0; e < NO; ++e) rp[e] = rapo[r][e];
      if (q <= 0 || q >= n0 - 1) {  // i faces (and the ring planes: all dead)
        const int t = opaque_tid();
        const bool qin = q >= 0 && q < n0;
#pragma unroll
        for (int e = 0; e < NO; ++e) {
          int j, k;
          const bool live = odd_cell(t, e, r, j, k) && qin;
          rp[e] = live ? rdiag(qin ? q : 0, j, k) : T(0);
        }
      }
      const T* SM = P[p];      // pin plane q-1
      const T* S0 = P[p + 1];  // pin plane q
      const T* SP = P[p + 2];  // pin plane q+1
#pragma unroll
      for (int e = 0; e < NO; ++e) {
        const int a = a0[e], c = aj[r][e];
        const T sum = c0 * SM[c] + c0 * SP[c] + c1 * S0[c - PH] + c1 * S0[c + PH] + c2 * S0[a] +
                      c2 * S0[a + 1];
        O[p][tid + 256 * e] = (rodd[p][e] - sum) * rp[e];
      }
    }
    __syncthreads();
    // even values of planes i0 .. i1-1 and their write
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int pl = i0 + p;
      if (pl >= i1) break;
      const int rn = (p + 1) & 1;  // class of plane pl
      T re[NPAIR];
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) re[e] = rape[rn][e];
      if (pl == 0 || pl == n0 - 1) {
        const int t = opaque_tid();
#pragma unroll
        for (int e = 0; e < NPAIR; ++e) {
          int j, ke, ko;
          pair_cell(t, e, rn, j, ke, ko);
          re[e] = rdiag(pl, j, min(ke, n2 - 1));
        }
      }
      const T* Om = O[p];      // odd values of plane pl-1
      const T* O0 = O[p + 1];  // plane pl
      const T* Op = O[p + 2];  // plane pl+1
      const int so = soff(pl);
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) {
        const int b = b0[e], c = bj[rn][e];
        const T sum = c0 * Om[c] + c0 * Op[c] + c1 * O0[c - OH] + c1 * O0[c + OH] + c2 * O0[b] +
                      c2 * O0[b + 1];
        const T ve = (reven[p][e] - sum) * re[e];
        const T vo = O0[bo[rn][e]];
        const uint32_t oe = eofs[rn][e], oo = pofs[rn][e];
        buf_st(dout, oe, so, ve);
        buf_st(dout, oo, so, vo);
        if (pl == 0) {  // i-face ring images
          buf_st(dout, oe, soff(-1), ve);
          buf_st(dout, oo, soff(-1), vo);
        }
        if (pl == n0 - 1) {
          buf_st(dout, oe, soff(n0), ve);
          buf_st(dout, oo, soff(n0), vo);
        }
        if (!wfast) jk_images(dout, e, rn, so, ve, vo, oe, oo);
      }
    }
    return;
  }

  // Descriptors start at array plane i0 (real plane i0-1); real plane x >= i0-1
  // sits at soffset (x + 1 - i0) * pb.
  const int pb = (int)(d.s0 * (int64_t)sizeof(T));
  auto chunk_rsrc = [&](const T* a) {
    return buf_rsrc(reinterpret_cast<const char*>(a) + (int64_t)i0 * pb, 0x7fffffffu);
  };
  // pin == nullptr: an iterate of zeros (a V-cycle's first sweep), read
  // through an empty descriptor -- every load returns 0, nothing is fetched
  const auto dpin = pin ? chunk_rsrc(pin) : buf_rsrc(rhs, 0u);
  const auto drhs = chunk_rsrc(rhs), dout = chunk_rsrc(pout);
  const auto dnone = buf_rsrc(pout, 0u);  // every store dropped
  auto soff = [&](int x) { return (x + 1 - i0) * pb; };

  // Loads run three planes ahead of their use: pre[x] holds the pin plane
  // to be stored into LDS slot x; rodd[u] and reven[u] the odd-cell rhs of
  // the plane iteration t = u (mod 6) relaxes and the even-cell rhs of the
  // plane it writes.
  T pre[3][NPE], rodd[6][NO], reven[6][NPAIR];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
#pragma unroll
    for (int e = 0; e < NO; ++e) rodd[u][e] = T(0);
#pragma unroll
    for (int e = 0; e < NPAIR; ++e) reven[u][e] = T(0);
  }
  // pin plane x (class r) into set X
  auto load_pin = [&](auto Rc, auto Xc, int x) {
    constexpr int r = decltype(Rc)::value, X = decltype(Xc)::value;
#pragma unroll
    for (int e = 0; e < NPE; ++e) pre[X][e] = buf_ld<T>(dpin, gofs[r][e], soff(x));
  };
  // set X (plane x) to LDS; the ring planes -1 and n0 enter as zero
  auto store_pin = [&](auto Xc, T* dst, int x) {
    constexpr int X = decltype(Xc)::value;
    const bool z = x < 0 || x >= n0;
#pragma unroll
    for (int e = 0; e < NPE; ++e) dst[tid + 256 * e] = z ? T(0) : pre[X][e];
  };
  auto load_rhs_odd = [&](auto Rc, auto Uc, int x) {  // plane x of class r into set U
    constexpr int r = decltype(Rc)::value, U = decltype(Uc)::value;
#pragma unroll
    for (int e = 0; e < NO; ++e) rodd[U][e] = buf_ld<T>(drhs, oofs[r][e], soff(x));
  };
  auto load_rhs_even = [&](auto Rc, auto Uc, int x) {
    constexpr int r = decltype(Rc)::value, U = decltype(Uc)::value;
#pragma unroll
    for (int e = 0; e < NPAIR; ++e) reven[U][e] = buf_ld<T>(drhs, eofs[r][e], soff(x));
  };

  // Iteration t handles q = i0-1+t.  Plane x sits in LDS slot
  // (x - i0 + 2) mod 3 and has parity class sb ^ ((x - i0 + 1) mod 2); with
  // u = t mod 6 these are compile-time constants here.  Planes past what the
  // chunk needs are clamped to one it may read and their values ignored.
  const int xpin = min(i1 + 1, n0);  // last pin plane read (n0: the ring plane)
  auto step = [&](auto Uc, int q) {
    constexpr int U = decltype(Uc)::value;
    constexpr int R = U % 3, r = U & 1, rn = r ^ 1;  // classes of planes q, q +- 1
    constexpr int SM = R, S0 = (R + 1) % 3, SP = (R + 2) % 3;  // planes q-1, q, q+1
    constexpr int U3 = (U + 3) % 6;
    store_pin(std::integral_constant<int, SP>{}, P[SP], q + 1);
    load_pin(std::integral_constant<int, r>{}, std::integral_constant<int, SP>{},
             min(q + 4, xpin));
    load_rhs_odd(std::integral_constant<int, rn>{}, std::integral_constant<int, U3>{},
                 min(q + 3, i1));
    load_rhs_even(std::integral_constant<int, r>{}, std::integral_constant<int, U3>{},
                  min(q + 2, i1));
    __syncthreads();
    {  // odd values of plane q over the halo tile
      T rp[NO];
#pragma unroll
      for (int e = 0; e < NO; ++e) rp[e] = rapo[r][e];
      if (q <= 0 || q >= n0 - 1) {  // i faces (and the ring planes: all dead)
        const int t = opaque_tid();
        const bool qin = q >= 0 && q < n0;
#pragma unroll
        for (int e = 0; e < NO; ++e) {
          int j, k;
          const bool live = odd_cell(t, e, r, j, k) && qin;
          rp[e] = live ? rdiag(qin ? q : 0, j, k) : T(0);
        }
      }
#pragma unroll
      for (int e = 0; e < NO; ++e) {
        const int a = a0[e], c = aj[r][e];
        const T sum = c0 * P[SM][c] + c0 * P[SP][c] + c1 * P[S0][c - PH] +
                      c1 * P[S0][c + PH] + c2 * P[S0][a] + c2 * P[S0][a + 1];
        O[S0][tid + 256 * e] = (rodd[U][e] - sum) * rp[e];
      }
    }
    __syncthreads();
    // Even values of plane pl (class rn) and the write of plane pl.  The
    // first two iterations (pl < i0) compute too, into dropped stores.
    const int pl = q - 1;
    const bool own = pl >= i0;
    T re[NPAIR];
#pragma unroll
    for (int e = 0; e < NPAIR; ++e) re[e] = rape[rn][e];
    if (pl == 0 || pl == n0 - 1) {
      const int t = opaque_tid();
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) {
        int j, ke, ko;
        pair_cell(t, e, rn, j, ke, ko);
        re[e] = rdiag(pl, j, min(ke, n2 - 1));
      }
    }
    const auto dst = own ? dout : dnone;
    const int so = own ? soff(pl) : 0;
#pragma unroll
    for (int e = 0; e < NPAIR; ++e) {
      const int b = b0[e], c = bj[rn][e];
      const T sum = c0 * O[SP][c] + c0 * O[S0][c] + c1 * O[SM][c - OH] + c1 * O[SM][c + OH] +
                    c2 * O[SM][b] + c2 * O[SM][b + 1];
      const T ve = (reven[U][e] - sum) * re[e];
      const T vo = O[SM][bo[rn][e]];
      const uint32_t oe = eofs[rn][e], oo = pofs[rn][e];
      buf_st(dst, oe, so, ve);
      buf_st(dst, oo, so, vo);
      if (own && (pl == 0 || pl == n0 - 1)) {  // i-face ring images
        if (pl == 0) {
          buf_st(dout, oe, soff(-1), ve);
          buf_st(dout, oo, soff(-1), vo);
        }
        if (pl == n0 - 1) {
          buf_st(dout, oe, soff(n0), ve);
          buf_st(dout, oo, soff(n0), vo);
        }
      }
      if (!wfast) jk_images(dst, e, rn, so, ve, vo, oe, oo);  // (dst: dropped if !own)
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I5 = std::integral_constant<int, 5>;
  if (i0 >= 2) {  // pin plane i0-2 lies before the descriptors' base
    const auto dp =
        pin ? buf_rsrc(reinterpret_cast<const char*>(pin) + (int64_t)(i0 - 1) * pb, 0x7fffffffu)
            : buf_rsrc(rhs, 0u);
#pragma unroll
    for (int e = 0; e < NPE; ++e) pre[0][e] = buf_ld<T>(dp, gofs[1][e], 0);
  }
  store_pin(I0{}, P[0], i0 >= 2 ? i0 - 2 : -1);
  load_pin(I0{}, I1{}, i0 - 1);
  store_pin(I1{}, P[1], i0 - 1);
  // planes i0, i0+1, i0+2 go to slots 2, 0, 1 at iterations 0, 1, 2
  load_pin(I1{}, I2{}, i0);
  load_pin(I0{}, I0{}, min(i0 + 1, xpin));
  load_pin(I1{}, I1{}, min(i0 + 2, xpin));
  load_rhs_odd(I0{}, I0{}, i0 - 1);
  load_rhs_odd(I1{}, I1{}, i0);
  load_rhs_odd(I0{}, I2{}, min(i0 + 1, i1));
  load_rhs_even(I1{}, I2{}, i0);  // written at iteration 2 (plane i0)
  const int cnt = i1 - i0 + 2;  // q = i0-1 .. i1, cnt >= 3
  for (int t = 0;; t += 6) {
    step(I0{}, i0 - 1 + t);
    if (t + 1 >= cnt) break;
    step(I1{}, i0 + t);
    if (t + 2 >= cnt) break;
    step(I2{}, i0 + 1 + t);
    if (t + 3 >= cnt) break;
    step(I3{}, i0 + 2 + t);
    if (t + 4 >= cnt) break;
    step(I4{}, i0 + 3 + t);
    if (t + 5 >= cnt) break;
    step(I5{}, i0 + 4 + t);
    if (t + 6 >= cnt) break;
  }
}
template <class T, int TJ, int TK, int NB = 0>
__global__ void __launch_bounds__(256)
sweep_kernel(const T* __restrict__ pin, T* __restrict__ pout, const T* __restrict__ rhs,
             MgDev d, T c0, T c1, T c2, int ich) {
  sweep_body<T, TJ, TK, NB>(pin, pout, rhs, d, c0, c1, c2, ich, this_blk());
}

// Two full RBGS sweeps fused in one pass: pout = S(S(pin)).  The workgroup
// streams its (j, k) tile along i through four lagged stages, one plane
// apart in pairs: at plane q it computes
//   A-odd  (first sweep, odd cells)  at q      over the tile + 3-cell halo,
//   A-even (first sweep, even cells) at q - 2  over the tile + 2-cell halo,
//   B-odd  (second sweep)            at q - 4  over the tile + 1-cell halo,
//   B-even (second sweep)            at q - 6  over the tile, and writes it
// with the B-odd values of that plane to pout.  Each stage reads only values
// the previous stage produced in earlier iterations (pin's even cells over
// the tile + 4-cell halo for A-odd), so one barrier per plane separates
// every producer from its consumers; each LDS array keeps four planes.  A
// stage is the single sweep's packed-checkerboard update: destination cell
// (row jr, packed column m) of the halo-h layout reads source row jr + 1 of
// the halo-(h+1) layout at columns m and m + 1 (k neighbours) and m + ph (j
// and i neighbours), ph the destination colour's phase in its row.  Halo and
// out-of-domain cells are masked as in the single sweep (zero values, zero
// reciprocal diagonal), so every value, halo copies included, is the single
// sweep's bit for bit.  HBM traffic per two sweeps: one read of pin and rhs,
// one write of pout; the halo recomputation costs arithmetic only.
template <class T, int TJ, int TK>
__device__ inline void sweep2_body(const T* __restrict__ pin, T* __restrict__ pout,
                                   const T* __restrict__ rhs, const MgDev& d, T c0, T c1, T c2,
                                   int ich, Blk B) {
  constexpr int NPAIR = TJ * TK / 512;  // output pairs per thread
  static_assert(TJ * TK == 512 * NPAIR && (NPAIR == 1 || NPAIR == 2), "256 threads, 1-2 pairs");
  static_assert(TK % 2 == 0, "pairs along k");
  // packed halo-h layout: TJ + 2h rows of W_h = (TK + 2h) / 2 cells of one colour
  constexpr int W0 = TK / 2, W1 = W0 + 1, W2 = W0 + 2, W3 = W0 + 3, W4 = W0 + 4;
  constexpr int C1 = (TJ + 2) * W1, C2 = (TJ + 4) * W2, C3 = (TJ + 6) * W3, C4 = (TJ + 8) * W4;
  constexpr int N1 = (C1 + 255) / 256, N2 = (C2 + 255) / 256, N3 = (C3 + 255) / 256,
                N4 = (C4 + 255) / 256;  // cells per thread
  // a plane slot holds the cells and one spare word that dead lanes write
  constexpr int Z1 = (C1 + 2) & ~1, Z2 = (C2 + 2) & ~1, Z3 = (C3 + 2) & ~1, Z4 = (C4 + 2) & ~1;
  __shared__ T P[4][Z4];   // pin, even cells, halo 4
  __shared__ T OA[4][Z3];  // first sweep, odd cells, halo 3
  __shared__ T EA[4][Z2];  // first sweep, even cells, halo 2
  __shared__ T OB[4][Z1];  // second sweep, odd cells, halo 1
  // rhs of the first sweep's stages, kept for the second sweep's (planes x
  // in slot (x - i0 + 3) % 5: A-odd plane x at iteration x, B-odd at x + 4)
  __shared__ T RAO[5][Z3];  // odd cells, halo 3
  __shared__ T RAE[5][Z2];  // even cells, halo 2
  const int tid = threadIdx.x;
  const int j0 = B.y * TJ, k0 = B.x * TK;
  const int i0 = B.z * ich, i1 = min(i0 + ich, d.n0);
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int s1 = (int)d.s1;
  const int sb = (i0 - 3 + j0 + k0) & 1;  // parity class of plane i0-3; index r: sb ^ r
  const int imid = n0 >= 3 ? 1 : 0;
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
  auto plane_boff = [&](int j, int k) {
    return (uint32_t)(((j + 1) * s1 + (k + 1)) * (int)sizeof(T));
  };
  auto in_jk = [&](int j, int k) { return j >= 0 && j < n1 && k >= 0 && k < n2; };
  auto rdiag = [&](int i, int j, int k) {
    return T(1) / mg_coef(d, i, clampi(j, 0, n1 - 1), clampi(k, 0, n2 - 1), c0, c1, c2).ap;
  };
  // cell e of a halo-h stage on class r: (j, k) and liveness; odd: the
  // colour of i + j + k odd
  auto stage_cell = [&](int t, int e, int r, int h, int W, int C, bool odd, int& j, int& k,
                        int& ph) {
    const int l = min(t + 256 * e, C - 1);
    const int jr = l / W, m = l % W;
    ph = ((sb ^ r) + jr + (odd ? 1 : 0)) & 1;
    j = j0 - h + jr;
    k = k0 - h + 2 * m + ph;
    return t + 256 * e < C && in_jk(j, k);
  };
  const bool wfast = j0 >= 1 && j0 + TJ <= n1 - 1 && k0 >= 1 && k0 + TK <= n2 - 1;

  // pin even cells (halo 4), offsets per class (SWEEP_DEAD outside the domain)
  uint32_t gofs[2][N4];
#pragma unroll
  for (int e = 0; e < N4; ++e)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int j, k, ph;
      const bool live = stage_cell(tid, e, r, 4, W4, C4, false, j, k, ph);
      gofs[r][e] = live ? plane_boff(j, k) : SWEEP_DEAD;
    }
  // stage tables: source columns (a0: k neighbours a0, a0+1 of source row jr+1;
  // aj[r]: j and i neighbours), rhs offsets and interior-plane 1/ap per class
#define WF_STAGE_TABLES(NAME, N, H, W, C, WS, ODD)                                   \
  int NAME##_a0[N], NAME##_aj[2][N];                                                 \
  uint32_t NAME##_ro[2][N];                                                          \
  T NAME##_rp[2][N];                                                                 \
  _Pragma("unroll") for (int e = 0; e < N; ++e) {                                    \
    const int l = min(tid + 256 * e, C - 1);                                         \
    NAME##_a0[e] = (l / W + 1) * WS + l % W;                                         \
    _Pragma("unroll") for (int r = 0; r < 2; ++r) {                                  \
      int j, k, ph;                                                                  \
      const bool live = stage_cell(tid, e, r, H, W, C, ODD, j, k, ph);               \
      NAME##_aj[r][e] = NAME##_a0[e] + ph;                                           \
      NAME##_ro[r][e] = live ? plane_boff(j, k) : SWEEP_DEAD;                        \
      NAME##_rp[r][e] = live ? rdiag(imid, j, k) : T(0);                             \
    }                                                                                \
  }
  WF_STAGE_TABLES(ao, N3, 3, W3, C3, W4, true)
  WF_STAGE_TABLES(ae, N2, 2, W2, C2, W3, false)
  WF_STAGE_TABLES(bo, N1, 1, W1, C1, W2, true)
  WF_STAGE_TABLES(be, NPAIR, 0, W0, TJ * W0, W1, false)
#undef WF_STAGE_TABLES
  // LDS index, in the first sweep's layouts (odd halo 3, even halo 2), of
  // the second sweep's cells (odd halo 1, even halo 0): row + 2, column + 1
  // (the colour phase of a row is the same two rows on)
  int bo_rix[N1], be_rix[NPAIR];
#pragma unroll
  for (int e = 0; e < N1; ++e) {
    const int l = min(tid + 256 * e, C1 - 1);
    bo_rix[e] = (l / W1 + 2) * W3 + l % W1 + 1;
  }
#pragma unroll
  for (int e = 0; e < NPAIR; ++e) {
    const int l = tid + 256 * e;
    be_rix[e] = (l / W0 + 2) * W2 + l % W0 + 1;
  }
  auto rslot = [&](int x) { return (x - i0 + 3) % 5; };
  // output: the even cell's offset is be_ro; the odd partner's offset and
  // OB column, and face flags of the pair's cells a = k_a, b = k_a + 1
  constexpr int FL_JLO = 1, FL_JHI = 2, FL_ALO = 4, FL_AHI = 8, FL_BHI = 16, FL_AL = 32,
                FL_BL = 64, FL_X = 128;
  uint32_t pofs[2][NPAIR];
  int pcol[2][NPAIR], fl[NPAIR];
#pragma unroll
  for (int e = 0; e < NPAIR; ++e) {
    const int p = tid + 256 * e;
    const int jr = p / W0, mm = p % W0;
    const int j = j0 + jr, ka = k0 + 2 * mm;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int kx = ((sb ^ r) + jr) & 1;  // the even cell is k_a + kx
      const int ko = ka + 1 - kx;
      pofs[r][e] = j < n1 && ko < n2 ? plane_boff(j, ko) : SWEEP_DEAD;
      pcol[r][e] = (jr + 1) * W1 + mm + 1 - kx;
    }
    const bool la = j < n1 && ka < n2, lb = j < n1 && ka + 1 < n2;
    fl[e] = (la || lb ? (j == 0 ? FL_JLO : 0) | (j == n1 - 1 ? FL_JHI : 0) : 0) |
            (la ? (ka == 0 ? FL_ALO : 0) | (ka == n2 - 1 ? FL_AHI : 0) | FL_AL : 0) |
            (lb ? (ka + 1 == n2 - 1 ? FL_BHI : 0) | FL_BL : 0) | (((sb + jr) & 1) ? FL_X : 0);
  }

  auto opaque_tid = [&]() {
    int t = tid;
    asm volatile("" : "+v"(t));
    return t;
  };
  // Descriptors start at the array plane of real plane max(i0-4, -1); every
  // plane a chunk touches lies within (ich + 8) planes of it.
  const int pb = (int)(d.s0 * (int64_t)sizeof(T));
  const int xb = max(i0 - 4, -1);
  auto chunk_rsrc = [&](const T* a) {
    return buf_rsrc(reinterpret_cast<const char*>(a) + (int64_t)(xb + 1) * pb, 0x7fffffffu);
  };
  // pin == nullptr: an iterate of zeros, read through an empty descriptor
  const auto dpin = pin ? chunk_rsrc(pin) : buf_rsrc(rhs, 0u);
  const auto drhs = chunk_rsrc(rhs), dout = chunk_rsrc(pout);
  // soffset of real plane x, clamped to the planes [lo, hi] a stage needs
  // (values of other planes are never used) and to the array
  auto soff = [&](int x, int lo, int hi) {
    return (clampi(clampi(x, lo, hi), -1, n0) - xb) * pb;
  };

  // zero every slot once, so no value a stage reads is ever uninitialised
  for (int w = tid; w < 4 * Z4; w += 256) (&P[0][0])[w] = T(0);
  for (int w = tid; w < 4 * Z3; w += 256) (&OA[0][0])[w] = T(0);
  for (int w = tid; w < 4 * Z2; w += 256) (&EA[0][0])[w] = T(0);
  for (int w = tid; w < 4 * Z1; w += 256) (&OB[0][0])[w] = T(0);
  __syncthreads();

  // Prefetch: pin two planes ahead (pre[c]: the next plane of class c); the
  // first sweep's rhs two planes ahead (4 sets), the second's one plane
  // ahead (2 sets; those planes were read a few planes earlier).
  T pre[2][N4], rao[4][N3], rae[4][N2];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int e = 0; e < N3; ++e) rao[x][e] = T(0);
#pragma unroll
    for (int e = 0; e < N2; ++e) rae[x][e] = T(0);
  }
  const int pin_hi = min(i1 + 3, n0);  // last pin plane needed (n0: ring, zero)
  auto load_pin = [&](auto Rc, int x) {
    constexpr int r = decltype(Rc)::value;
    const int so = soff(x, i0 - 4, pin_hi);
#pragma unroll
    for (int e = 0; e < N4; ++e) pre[r][e] = buf_ld<T>(dpin, gofs[r][e], so);
  };
  auto store_pin = [&](auto Rc, T* dst, int x) {  // ring planes and beyond enter as zero
    constexpr int r = decltype(Rc)::value;
    const bool z = x < 0 || x >= n0;
#pragma unroll
    for (int e = 0; e < N4; ++e) dst[tid + 256 * e < C4 ? tid + 256 * e : C4] = z ? T(0) : pre[r][e];
  };

  // one stage's cells: v = (rhs - sum) * 1/ap into dst (dead lanes: spare word)
  auto relax = [&](auto Nc, auto WSc, auto Cc, const T* sm, const T* s0, const T* sp,
                   const int* a0, const int* aj, const T* rv, const T* rp, T* dst) {
    constexpr int N = decltype(Nc)::value, WS = decltype(WSc)::value, C = decltype(Cc)::value;
#pragma unroll
    for (int e = 0; e < N; ++e) {
      const int a = a0[e], c = aj[e];
      const T sum = c0 * sm[c] + c0 * sp[c] + c1 * s0[c - WS] + c1 * s0[c + WS] +
                    c2 * s0[a] + c2 * s0[a + 1];
      const int l = tid + 256 * e;
      dst[(e + 1) * 256 <= C ? l : (l < C ? l : C)] = (rv[e] - sum) * rp[e];
    }
  };
  // a first-sweep stage's rhs into its ring slot (dead lanes: spare word)
  auto keep_rhs = [&](auto Nc, auto Cc, const T* rv, T* dst) {
    constexpr int N = decltype(Nc)::value, C = decltype(Cc)::value;
#pragma unroll
    for (int e = 0; e < N; ++e) {
      const int l = tid + 256 * e;
      dst[(e + 1) * 256 <= C ? l : (l < C ? l : C)] = rv[e];
    }
  };
  // 1/ap of a stage's cells on an i-face (or ring) plane x
  auto face_rp = [&](auto Nc, int x, int r, int h, int W, int C, bool odd, T* rp) {
    constexpr int N = decltype(Nc)::value;
    const int t = opaque_tid();
    const bool xin = x >= 0 && x < n0;
#pragma unroll
    for (int e = 0; e < N; ++e) {
      int j, k, ph;
      const bool live = stage_cell(t, e, r, h, W, C, odd, j, k, ph) && xin;
      rp[e] = live ? rdiag(xin ? x : 0, j, k) : T(0);
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  // j and k face-ring images of an output pair (cells a = k_a, b = k_a + 1)
  auto jk_images = [&](int e, int r, int so, T ve, T vo, uint32_t oe, uint32_t oo) {
    const int f = fl[e];
    const bool xbf = ((f & FL_X) != 0) != (r != 0);  // the even cell is b
    const uint32_t oa = xbf ? oo : oe, ob = xbf ? oe : oo;
    const T va = xbf ? vo : ve, vb = xbf ? ve : vo;
    const uint32_t sj = (uint32_t)(s1 * (int)sizeof(T)), sk = (uint32_t)sizeof(T);
    const uint32_t jd = (f & FL_JLO) ? 0u - sj : sj;
    const bool jf = (f & (FL_JLO | FL_JHI)) != 0;
    buf_st(dout, jf && (f & FL_AL) ? oa + jd : SWEEP_DEAD, so, va);
    buf_st(dout, jf && (f & FL_BL) ? ob + jd : SWEEP_DEAD, so, vb);
    buf_st(dout, (f & FL_ALO) ? oa - sk : (f & FL_AHI) ? oa + sk : SWEEP_DEAD, so, va);
    buf_st(dout, (f & FL_BHI) ? ob + sk : SWEEP_DEAD, so, vb);
    if (n1 == 1) {  // both j faces: the high images too
      buf_st(dout, (f & FL_AL) ? oa + sj : SWEEP_DEAD, so, va);
      buf_st(dout, (f & FL_BL) ? ob + sj : SWEEP_DEAD, so, vb);
    }
    if (n2 == 1) buf_st(dout, (f & FL_AL) ? oa + sk : SWEEP_DEAD, so, va);
  };
  // The four stages of a steady iteration.  Each reads LDS slots that no
  // stage of the same iteration writes; the restrict qualifiers say so, which
  // lets the compiler issue a later stage's reads before an earlier stage's
  // writes.
  auto steady_stages = [&](auto Rc, auto XAc, const T* __restrict__ pm,
                           const T* __restrict__ p0, const T* __restrict__ pp, T* __restrict__ oa,
                           const T* __restrict__ oam, const T* __restrict__ oa0,
                           const T* __restrict__ oap, T* __restrict__ ea,
                           const T* __restrict__ eam, const T* __restrict__ ea0,
                           const T* __restrict__ eap, T* __restrict__ ob,
                           const T* __restrict__ obm, const T* __restrict__ ob0,
                           const T* __restrict__ obp, T* __restrict__ kao,
                           T* __restrict__ kae, const T* __restrict__ rbo,
                           const T* __restrict__ rbe, int so) {
    constexpr int r = decltype(Rc)::value, XA = decltype(XAc)::value;
    relax(std::integral_constant<int, N3>{}, std::integral_constant<int, W4>{},
          std::integral_constant<int, C3>{}, pm, p0, pp, ao_a0, ao_aj[r], rao[XA], ao_rp[r], oa);
    keep_rhs(std::integral_constant<int, N3>{}, std::integral_constant<int, C3>{}, rao[XA], kao);
    relax(std::integral_constant<int, N2>{}, std::integral_constant<int, W3>{},
          std::integral_constant<int, C2>{}, oam, oa0, oap, ae_a0, ae_aj[r], rae[XA], ae_rp[r],
          ea);
    keep_rhs(std::integral_constant<int, N2>{}, std::integral_constant<int, C2>{}, rae[XA], kae);
    T rv[N1];
#pragma unroll
    for (int e = 0; e < N1; ++e) rv[e] = rbo[bo_rix[e]];
    relax(std::integral_constant<int, N1>{}, std::integral_constant<int, W2>{},
          std::integral_constant<int, C1>{}, eam, ea0, eap, bo_a0, bo_aj[r], rv, bo_rp[r], ob);
#pragma unroll
    for (int e = 0; e < NPAIR; ++e) {
      const int a = be_a0[e], c = be_aj[r][e];
      const T sum = c0 * obm[c] + c0 * obp[c] + c1 * ob0[c - W1] + c1 * ob0[c + W1] +
                    c2 * ob0[a] + c2 * ob0[a + 1];
      const T ve = (rbe[be_rix[e]] - sum) * be_rp[r][e];
      const T vo = ob0[pcol[r][e]];
      const uint32_t oe = be_ro[r][e], oo = pofs[r][e];
      buf_st(dout, oe, so, ve);
      buf_st(dout, oo, so, vo);
      if (!wfast) jk_images(e, r, so, ve, vo, oe, oo);
    }
  };
  // Steady iterations q in [qs, qe]: every stage runs and no stage's plane is
  // an i face or ring plane (A-odd's q <= n0 - 2, B-even's q - 6 >= 1), so
  // they skip the stage tests, the face diagonals and the i-face images, the
  // pin plane they store (q + 1 < n0) enters as loaded, and their load planes
  // need only the upper clamps (the plane offsets are those soff gives).
  const int qs = max(i0 + 6, 7), qe = min(i1 + 2, n0 - 2);
  const int hx_ao = min(i1 + 2, n0), hx_ae = min(i1 + 1, n0);
  // Iteration t handles q = i0-3+t; plane x sits in LDS slot (x - i0 + 3) & 3
  // and has class index (x - i0 + 3) & 1.
  auto step = [&](auto Uc, auto Sc, int q) {
    constexpr int U = decltype(Uc)::value;
    constexpr bool ST = decltype(Sc)::value;
    constexpr int r = U & 1, rn = r ^ 1;
    constexpr int Sm1 = (U + 3) & 3, S0 = U, Sp1 = (U + 1) & 3, Sm2 = (U + 2) & 3;
    using Cr = std::integral_constant<int, r>;
    using Crn = std::integral_constant<int, rn>;
    // Loads, two iterations ahead: pin and the A stages' rhs (the B stages
    // read theirs from the ring the A stages fill).
    if constexpr (ST) {
#pragma unroll
      for (int e = 0; e < N4; ++e) P[Sp1][tid + 256 * e < C4 ? tid + 256 * e : C4] = pre[rn][e];
      const int so = (min(q + 3, pin_hi) - xb) * pb;
#pragma unroll
      for (int e = 0; e < N4; ++e) pre[rn][e] = buf_ld<T>(dpin, gofs[rn][e], so);
    } else {
      store_pin(Crn{}, P[Sp1], q + 1);
      load_pin(Crn{}, q + 3);
    }
    {  // rhs of the A stages of iteration t+2 (class r)
      constexpr int XA = (U + 2) & 3;
      const int sao = ST ? (min(q + 2, hx_ao) - xb) * pb : soff(q + 2, i0 - 3, i1 + 2);
      const int sae = ST ? (min(q, hx_ae) - xb) * pb : soff(q, i0 - 2, i1 + 1);
#pragma unroll
      for (int e = 0; e < N3; ++e) rao[XA][e] = buf_ld<T>(drhs, ao_ro[r][e], sao);
#pragma unroll
      for (int e = 0; e < N2; ++e) rae[XA][e] = buf_ld<T>(drhs, ae_ro[r][e], sae);
    }
    __syncthreads();
    constexpr int XA = U;
    if constexpr (ST) {
      steady_stages(Cr{}, std::integral_constant<int, XA>{}, P[Sm1], P[S0], P[Sp1], OA[S0],
                    OA[(U + 1) & 3], OA[Sm2], OA[Sm1], EA[Sm2], EA[Sm1], EA[S0], EA[Sp1],
                    OB[S0], OB[Sp1], OB[Sm2], OB[Sm1], RAO[rslot(q)], RAE[rslot(q - 2)],
                    RAO[rslot(q - 4)], RAE[rslot(q - 6)], (q - 6 - xb) * pb);
      return;
    }
    // Each stage runs only for the planes a later stage or the output needs
    // (A-odd i0-3 .. i1+2, A-even i0-2 .. i1+1, B-odd i0-1 .. i1, B-even
    // i0 .. i1-1); the pipeline's fill and drain iterations skip the rest.
    // The conditions are uniform over the workgroup.
    if (q <= i1 + 2) {  // A-odd at q: pin planes q-1, q, q+1 -> OA[q]
      T rp[N3];
#pragma unroll
      for (int e = 0; e < N3; ++e) rp[e] = ao_rp[r][e];
      if (q <= 0 || q >= n0 - 1) face_rp(std::integral_constant<int, N3>{}, q, r, 3, W3, C3, true, rp);
      relax(std::integral_constant<int, N3>{}, std::integral_constant<int, W4>{},
            std::integral_constant<int, C3>{}, P[Sm1], P[S0], P[Sp1], ao_a0, ao_aj[r],
            rao[XA], rp, OA[S0]);
      keep_rhs(std::integral_constant<int, N3>{}, std::integral_constant<int, C3>{}, rao[XA],
               RAO[rslot(q)]);
    }
    if (q - 2 >= i0 - 2 && q - 2 <= i1 + 1) {  // A-even at q-2: OA planes q-3 .. q-1 -> EA[q-2]
      const int x = q - 2;
      T rp[N2];
#pragma unroll
      for (int e = 0; e < N2; ++e) rp[e] = ae_rp[r][e];
      if (x <= 0 || x >= n0 - 1) face_rp(std::integral_constant<int, N2>{}, x, r, 2, W2, C2, false, rp);
      relax(std::integral_constant<int, N2>{}, std::integral_constant<int, W3>{},
            std::integral_constant<int, C2>{}, OA[(U + 1) & 3], OA[Sm2], OA[Sm1], ae_a0,
            ae_aj[r], rae[XA], rp, EA[Sm2]);
      keep_rhs(std::integral_constant<int, N2>{}, std::integral_constant<int, C2>{}, rae[XA],
               RAE[rslot(x)]);
    }
    if (q - 4 >= i0 - 1 && q - 4 <= i1) {  // B-odd at q-4: EA planes q-5 .. q-3 -> OB[q-4]
      const int x = q - 4;
      T rp[N1];
#pragma unroll
      for (int e = 0; e < N1; ++e) rp[e] = bo_rp[r][e];
      if (x <= 0 || x >= n0 - 1) face_rp(std::integral_constant<int, N1>{}, x, r, 1, W1, C1, true, rp);
      T rv[N1];
#pragma unroll
      for (int e = 0; e < N1; ++e) rv[e] = RAO[rslot(x)][bo_rix[e]];
      relax(std::integral_constant<int, N1>{}, std::integral_constant<int, W2>{},
            std::integral_constant<int, C1>{}, EA[Sm1], EA[S0], EA[Sp1], bo_a0, bo_aj[r], rv,
            rp, OB[S0]);
    }
    if (q - 6 >= i0 && q - 6 < i1) {  // B-even at q-6: OB planes q-7 .. q-5 -> pout plane q-6
      const int x = q - 6;
      T rp[NPAIR];
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) rp[e] = be_rp[r][e];
      if (x <= 0 || x >= n0 - 1) face_rp(std::integral_constant<int, NPAIR>{}, x, r, 0, W0, TJ * W0, false, rp);
      const T* om = OB[Sp1];  // plane q-7
      const T* o0 = OB[Sm2];  // plane q-6
      const T* op = OB[Sm1];  // plane q-5
      const auto dst = dout;
      const int so = (x - xb) * pb;
#pragma unroll
      for (int e = 0; e < NPAIR; ++e) {
        const int a = be_a0[e], c = be_aj[r][e];
        const T sum = c0 * om[c] + c0 * op[c] + c1 * o0[c - W1] + c1 * o0[c + W1] +
                      c2 * o0[a] + c2 * o0[a + 1];
        const T ve = (RAE[rslot(x)][be_rix[e]] - sum) * rp[e];
        const T vo = o0[pcol[r][e]];
        const uint32_t oe = be_ro[r][e], oo = pofs[r][e];
        buf_st(dst, oe, so, ve);
        buf_st(dst, oo, so, vo);
        if (x == 0 || x == n0 - 1) {  // i-face ring images
          if (x == 0) {
            buf_st(dout, oe, (-1 - xb) * pb, ve);
            buf_st(dout, oo, (-1 - xb) * pb, vo);
          }
          if (x == n0 - 1) {
            buf_st(dout, oe, (n0 - xb) * pb, ve);
            buf_st(dout, oo, (n0 - xb) * pb, vo);
          }
        }
        if (!wfast) jk_images(e, r, so, ve, vo, oe, oo);  // j and k face-ring images
      }
    }
  };

  // prologue: pin planes i0-4 and i0-3 to LDS, i0-2 and i0-1 to registers;
  // the first sweep's odd rhs of planes i0-3 and i0-2
  load_pin(I1{}, i0 - 4);
  store_pin(I1{}, P[3], i0 - 4);
  load_pin(I0{}, i0 - 3);
  store_pin(I0{}, P[0], i0 - 3);
  load_pin(I1{}, i0 - 2);
  load_pin(I0{}, i0 - 1);
  {
    const int s0x = soff(i0 - 3, i0 - 3, i1 + 2), s1x = soff(i0 - 2, i0 - 3, i1 + 2);
#pragma unroll
    for (int e = 0; e < N3; ++e) {
      rao[0][e] = buf_ld<T>(drhs, ao_ro[0][e], s0x);
      rao[1][e] = buf_ld<T>(drhs, ao_ro[1][e], s1x);
    }
  }
  // Let the prologue's loads land before the loop, so that the compiler's
  // memory-counter waits inside the loop count only the loop's own loads.
  __builtin_amdgcn_s_waitcnt(0);
  const int cnt = i1 - i0 + 9;  // q = i0-3 .. i1+5
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  auto iter = [&](auto Uc, int q) {
    if (q >= qs && q <= qe)
      step(Uc, std::true_type{}, q);
    else
      step(Uc, std::false_type{}, q);
  };
  for (int t = 0;; t += 4) {
    iter(I0{}, i0 - 3 + t);
    if (t + 1 >= cnt) break;
    iter(I1{}, i0 - 2 + t);
    if (t + 2 >= cnt) break;
    iter(I2{}, i0 - 1 + t);
    if (t + 3 >= cnt) break;
    iter(I3{}, i0 + t);
    if (t + 4 >= cnt) break;
  }
}
// The FP32 two-sweep kernel is held to 128 VGPRs so that four workgroups
// share a CU (it fits without spilling); the FP64 one keeps the compiler's
// allocation.
// The FP32 tiles with two pairs per thread (16 x 64, 32 x 32) carry twice the
// stage tables and are held to 256 VGPRs (two workgroups per CU).
template <class T, int TJ, int TK>
constexpr int sweep2_min_blocks() {
  return sizeof(T) == 4 ? (TJ * TK == 512 ? 3 : 2) : 1;
}
template <class T, int TJ, int TK>
__global__ void __launch_bounds__(256, (sweep2_min_blocks<T, TJ, TK>()))
sweep2_kernel(const T* __restrict__ pin, T* __restrict__ pout, const T* __restrict__ rhs,
              MgDev d, T c0, T c1, T c2, int ich) {
  sweep2_body<T, TJ, TK>(pin, pout, rhs, d, c0, c1, c2, ich, this_blk());
}

// n sweeps of RBGS on one small level held in LDS by one workgroup
// (multigrid.py:289-299 with the refill folded in).  The real cells of phi
// live in LDS; ring terms carry zero coefficients and are skipped.  rhs is
// in LDS too when RHS_LDS, else read from global memory (L2-resident).
template <class T, bool RHS_LDS>
__device__ inline void relax_lds_body(T* __restrict__ phi, const T* __restrict__ rhs,
                                      const MgDev& d, T c0, T c1, T c2, int nsweeps, T* sp) {
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int ncell = n0 * n1 * n2;
  T* sr = sp + ncell;
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int l = tid; l < ncell; l += nt) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    const int64_t x = mg_at(d, i, j, k);
    sp[l] = phi[x];
    if (RHS_LDS) sr[l] = rhs[x];
  }
  __syncthreads();
  const int np = (n2 + 1) / 2;
  const int nrow = n0 * n1;
  const int S0 = n1 * n2, S1 = n2;
  for (int s = 0; s < nsweeps; ++s) {
    for (int color = 0; color < 2; ++color) {
      for (int q = tid; q < nrow * np; q += nt) {
        const int m = q % np, r = q / np;
        const int j = r % n1, i = r / n1;
        const int k = 2 * m + ((i + j + color + 1) & 1);
        if (k >= n2) continue;
        const int l = (i * n1 + j) * n2 + k;
        const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
        const T W = i > 0 ? sp[l - S0] : T(0);
        const T E = i < n0 - 1 ? sp[l + S0] : T(0);
        const T S = j > 0 ? sp[l - S1] : T(0);
        const T N = j < n1 - 1 ? sp[l + S1] : T(0);
        const T B = k > 0 ? sp[l - 1] : T(0);
        const T Tt = k < n2 - 1 ? sp[l + 1] : T(0);
        const T rv = RHS_LDS ? sr[l] : rhs[mg_at(d, i, j, k)];
        const T sum = c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt;
        sp[l] = (rv - sum) * (T(1) / c.ap);
      }
      __syncthreads();
    }
  }
  for (int l = tid; l < ncell; l += nt) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    const int64_t x = mg_at(d, i, j, k);
    const T v = sp[l];
    phi[x] = v;
    if (nsweeps > 0) ring_images(phi, d, i, j, k, x, v);
  }
}
template <class T, bool RHS_LDS>
__global__ void __launch_bounds__(1024)
relax_lds_kernel(T* __restrict__ phi, const T* __restrict__ rhs, MgDev d, T c0, T c1, T c2,
                 int nsweeps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  relax_lds_body<T, RHS_LDS>(phi, rhs, d, c0, c1, c2, nsweeps, reinterpret_cast<T*>(smem_raw));
}

// n sweeps of RBGS on one level held in LDS by one workgroup of 1024 threads,
// for levels whose zero-ringed copy fits LDS and whose cells of one colour
// number at most 4 per thread (e.g. 16^3).
// LDS holds the (n0+2) x (n1+2) x (n2+2) padded level split by colour: the
// cells of one colour of a padded row packed into W = ceil((n2+2)/2) words,
// so cell (I, J, K) (padded coordinates) of colour a = (I+J+K) & 1 sits at
// word ((I*(n1+2) + J)*W + K/2) of array a, its i and j neighbours at that
// word -/+ (n1+2)*W and -/+ W of the other array, and its k neighbours at
// km and km + 1 with km = word - 1 + (K & 1): neighbouring lanes read
// neighbouring words.  The ring is zero, so every cell uses the uniform
// coefficients c0, c1, c2 in _mgcoeff's order of terms and only 1/ap is per
// cell -- the arithmetic of the sweep kernels, bit for bit.  Each thread keeps
// its cells' words, 1/ap and rhs in registers (CPT cells per colour); dead
// slots point at a ring (or spare) word of their own array with 1/ap = 0,
// which keeps the ring zero.
template <class T, int CPT>
__device__ inline void relax_lds2_body(T* __restrict__ phi, const T* __restrict__ rhs,
                                       const MgDev& d, T c0, T c1, T c2, int nsweeps, T* sm) {
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int W = (n2 + 3) / 2;          // packed words per padded row and colour
  const int R1 = n1 + 2;               // padded rows per padded plane
  const int PW = (n0 + 2) * R1 * W;    // words per colour array
  const int SI = R1 * W;               // i stride in words
  T* A[2] = {sm, sm + PW};
  const int nt = blockDim.x, tid = threadIdx.x;
  const int ncell = n0 * n1 * n2;
  auto word = [&](int i, int j, int k) {  // real cell (i, j, k)
    return ((i + 1) * R1 + (j + 1)) * W + ((k + 1) >> 1);
  };
  for (int w = tid; w < 2 * PW; w += nt) sm[w] = T(0);
  __syncthreads();
  for (int l = tid; l < ncell; l += nt) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    A[(i + j + k + 3) & 1][word(i, j, k)] = phi[mg_at(d, i, j, k)];
  }
  // this thread's cells of colour c: k = 2m + ((i + j + c + 1) & 1) in row
  // (i, j), slot q = (i*n1 + j)*np + m; word*2 + (K & 1) packed in pk
  const int np = (n2 + 1) / 2, nslot = n0 * n1 * np;
  int pk[2][CPT];
  T rp[2][CPT], rv[2][CPT];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int e = 0; e < CPT; ++e) {
      const int q = tid + nt * e;
      const int qc = q < nslot ? q : nslot - 1;
      const int m = qc % np, r = qc / np;
      const int j = r % n1, i = r / n1;
      const int k = 2 * m + ((i + j + c + 1) & 1);
      const bool live = q < nslot && k < n2;
      if (live) {
        pk[c][e] = word(i, j, k) * 2 + ((k + 1) & 1);
      } else {  // a ring (or spare) word of array c in row (i+1, j+1): K = 0 or n2+1
        const int sp = (c + i + j) & 1;  // array c holds K = 2t + sp in this row
        const int t = sp == 0 ? 0 : W - 1;
        pk[c][e] = (((i + 1) * R1 + (j + 1)) * W + t) * 2 + sp;
      }
      rp[c][e] = live ? T(1) / mg_coef(d, i, j, k, c0, c1, c2).ap : T(0);
      rv[c][e] = live ? rhs[mg_at(d, i, j, k)] : T(0);
    }
  __syncthreads();
  for (int s = 0; s < nsweeps; ++s) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T* src = A[c ^ 1];
      T* dst = A[c];
#pragma unroll
      for (int e = 0; e < CPT; ++e) {
        const int wd = pk[c][e] >> 1, km = wd - 1 + (pk[c][e] & 1);
        const T sum = c0 * src[wd - SI] + c0 * src[wd + SI] + c1 * src[wd - W] +
                      c1 * src[wd + W] + c2 * src[km] + c2 * src[km + 1];
        dst[wd] = (rv[c][e] - sum) * rp[c][e];
      }
      __syncthreads();
    }
  }
  for (int l = tid; l < ncell; l += nt) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    const int64_t x = mg_at(d, i, j, k);
    const T v = A[(i + j + k + 3) & 1][word(i, j, k)];
    phi[x] = v;
    if (nsweeps > 0) ring_images(phi, d, i, j, k, x, v);
  }
}
template <class T, int CPT>
__global__ void __launch_bounds__(1024)
relax_lds2_kernel(T* __restrict__ phi, const T* __restrict__ rhs, MgDev d, T c0, T c1, T c2,
                  int nsweeps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  relax_lds2_body<T, CPT>(phi, rhs, d, c0, c1, c2, nsweeps, reinterpret_cast<T*>(smem_raw));
}

// Column `col` (= workgroup) of a small level's coarse operator: nsweeps RBGS
// sweeps in LDS from phi = 0 with rhs = e_col, the relax_lds_kernel
// arithmetic; M is column-major (M[col * ncell + row]).
template <class T>
__device__ inline void coarse_op_build_body(T* __restrict__ M, const MgDev& d, T c0, T c1, T c2,
                                            int nsweeps, int col, T* sp) {
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int ncell = n0 * n1 * n2;
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int l = tid; l < ncell; l += nt) sp[l] = T(0);
  __syncthreads();
  const int np = (n2 + 1) / 2, nrow = n0 * n1, S0 = n1 * n2, S1 = n2;
  for (int s = 0; s < nsweeps; ++s) {
    for (int color = 0; color < 2; ++color) {
      for (int q = tid; q < nrow * np; q += nt) {
        const int m = q % np, r = q / np;
        const int j = r % n1, i = r / n1;
        const int k = 2 * m + ((i + j + color + 1) & 1);
        if (k >= n2) continue;
        const int l = (i * n1 + j) * n2 + k;
        const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
        const T W = i > 0 ? sp[l - S0] : T(0);
        const T E = i < n0 - 1 ? sp[l + S0] : T(0);
        const T S = j > 0 ? sp[l - S1] : T(0);
        const T N = j < n1 - 1 ? sp[l + S1] : T(0);
        const T B = k > 0 ? sp[l - 1] : T(0);
        const T Tt = k < n2 - 1 ? sp[l + 1] : T(0);
        const T rv = l == col ? T(1) : T(0);
        const T sum = c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt;
        sp[l] = (rv - sum) * (T(1) / c.ap);
      }
      __syncthreads();
    }
  }
  for (int l = tid; l < ncell; l += nt) M[(size_t)col * ncell + l] = sp[l];
}
template <class T>
__global__ void __launch_bounds__(256)
coarse_op_build_kernel(T* __restrict__ M, MgDev d, T c0, T c1, T c2, int nsweeps) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  coarse_op_build_body<T>(M, d, c0, c1, c2, nsweeps, blockIdx.x, reinterpret_cast<T*>(smem_raw));
}

// phi = M rhs on the real cells of a small level, plus face-ring images.
// acc + m * r with one rounding, as both coarse-operator kernels accumulate
__device__ inline float cop_fma(float m, float r, float acc) { return __builtin_fmaf(m, r, acc); }
__device__ inline double cop_fma(double m, double r, double acc) {
  return __builtin_fma(m, r, acc);
}
template <class T>
__device__ inline void coarse_op_apply_body(const T* __restrict__ M, T* __restrict__ phi,
                                            const T* __restrict__ rhs, const MgDev& d, T* r) {
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int ncell = n0 * n1 * n2;
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int l = tid; l < ncell; l += nt) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    r[l] = rhs[mg_at(d, i, j, k)];
  }
  __syncthreads();
  for (int row = tid; row < ncell; row += nt) {
    // serial sum in column order; the loads of 16 columns are issued
    // before their products so that one workgroup keeps them in flight
    T acc = T(0);
    int col = 0;
    for (; col + 16 <= ncell; col += 16) {
      T m[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) m[q] = M[(size_t)(col + q) * ncell + row];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc = cop_fma(m[q], r[col + q], acc);
    }
    for (; col < ncell; ++col) acc = cop_fma(M[(size_t)col * ncell + row], r[col], acc);
    const int k = row % n2, j = (row / n2) % n1, i = row / (n1 * n2);
    const int64_t x = mg_at(d, i, j, k);
    phi[x] = acc;
    ring_images(phi, d, i, j, k, x, acc);
  }
}
// The rows of coarse_op_apply_body spread over one 64-thread workgroup per 64 rows, so that the
// matrix streams through many CUs instead of one; each row is still one
// thread's serial sum in column order (the bits of coarse_op_apply_body).
template <class T>
__global__ void __launch_bounds__(64)
coarse_op_apply_rows_kernel(const T* __restrict__ M, T* __restrict__ phi,
                            const T* __restrict__ rhs, MgDev d) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* r = reinterpret_cast<T*>(smem_raw);
  const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
  const int ncell = n0 * n1 * n2;
  const int tid = threadIdx.x;
  for (int l = tid; l < ncell; l += 64) {
    const int k = l % n2, j = (l / n2) % n1, i = l / (n1 * n2);
    r[l] = rhs[mg_at(d, i, j, k)];
  }
  __syncthreads();
  const int row = blockIdx.x * 64 + tid;
  if (row >= ncell) return;
  // CB columns' loads in flight before their products: the few workgroups
  // of a coarse level wait for one memory round trip per CB columns
  constexpr int CB = sizeof(T) == 4 ? 64 : 32;
  T acc = T(0);
  int col = 0;
  for (; col + CB <= ncell; col += CB) {
    T m[CB];
#pragma unroll
    for (int q = 0; q < CB; ++q) m[q] = M[(size_t)(col + q) * ncell + row];
#pragma unroll
    for (int q = 0; q < CB; ++q) acc = cop_fma(m[q], r[col + q], acc);
  }
  for (; col < ncell; ++col) acc = cop_fma(M[(size_t)col * ncell + row], r[col], acc);
  const int k = row % n2, j = (row / n2) % n1, i = row / (n1 * n2);
  const int64_t x = mg_at(d, i, j, k);
  phi[x] = acc;
  ring_images(phi, d, i, j, k, x, acc);
}

// Coarse operators of more than COARSE_OP_SERIAL_MAX_CELLS cells (a 20^3
// coarsest level is 8000 cells: M is 512 MB in FP64) are applied in two
// passes so that the matrix streams through the whole chip: workgroup
// (x, y) sums columns [y*clen, (y+1)*clen) of rows [x*256, x*256+256) in
// column order into part[y][row], and the fold adds a row's chunk sums in
// chunk order.  Both orders are fixed, so the result is deterministic; it
// differs from the serial row sum only by rounding.
constexpr int COP_SPLIT_ROWS = 256;
constexpr int COP_SPLIT_MAX_CLEN = 512;
template <class T>
__global__ void __launch_bounds__(256)
coarse_op_split_kernel(const T* __restrict__ M, const T* __restrict__ rhs, MgDev d, int clen,
                       T* __restrict__ part) {
  __shared__ T r[COP_SPLIT_MAX_CLEN];
  const int n1 = d.n1, n2 = d.n2;
  const int ncell = d.n0 * n1 * n2;
  const int c0 = blockIdx.y * clen;
  const int nc = min(clen, ncell - c0);
  for (int l = threadIdx.x; l < nc; l += COP_SPLIT_ROWS) {
    const int col = c0 + l;
    const int k = col % n2, j = (col / n2) % n1, i = col / (n1 * n2);
    r[l] = rhs[mg_at(d, i, j, k)];
  }
  __syncthreads();
  const int row = blockIdx.x * COP_SPLIT_ROWS + threadIdx.x;
  if (row >= ncell) return;
  const T* __restrict__ m0 = M + (size_t)c0 * ncell + row;
  T acc = T(0);
  int c = 0;
  for (; c + 16 <= nc; c += 16) {
    T m[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) m[q] = m0[(size_t)(c + q) * ncell];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc = cop_fma(m[q], r[c + q], acc);
  }
  for (; c < nc; ++c) acc = cop_fma(m0[(size_t)c * ncell], r[c], acc);
  part[(size_t)blockIdx.y * ncell + row] = acc;
}
template <class T>
__global__ void __launch_bounds__(256)
coarse_op_fold_kernel(const T* __restrict__ part, int nchunk, T* __restrict__ phi, MgDev d) {
  const int n1 = d.n1, n2 = d.n2;
  const int ncell = d.n0 * n1 * n2;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= ncell) return;
  T acc = part[row];
  for (int c = 1; c < nchunk; ++c) acc += part[(size_t)c * ncell + row];
  const int k = row % n2, j = (row / n2) % n1, i = row / (n1 * n2);
  const int64_t x = mg_at(d, i, j, k);
  phi[x] = acc;
  ring_images(phi, d, i, j, k, x, acc);
}

__global__ void __launch_bounds__(256)
copy_in_kernel(const double* __restrict__ su, wf_grid g, double* __restrict__ rhs, MgDev d) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.n2 || j >= d.n1 || i >= d.n0) return;
  rhs[mg_at(d, i, j, k)] = su[at(g, i + g.npl, j + g.npl, k + g.npl)];
}

__global__ void __launch_bounds__(256)
copy_out_kernel(const double* __restrict__ phi, MgDev d, double* __restrict__ dst, wf_grid g) {
  // over the whole (n+2)^3 level array, ring included (multigrid.py:157-163)
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.n2 + 2 || j >= d.n1 + 2 || i >= d.n0 + 2) return;
  dst[at(g, i + g.npl - 1, j + g.npl - 1, k + g.npl - 1)] =
      phi[(int64_t)i * d.s0 + (int64_t)j * d.s1 + k];
}

template <class T>
__global__ void __launch_bounds__(256) zero_kernel(T* __restrict__ p, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = T(0);
}

template <class T>
__global__ void __launch_bounds__(256)
coeff_kernel(MgDev d, T c0, T c1, T c2, T* aw, T* ae, T* as, T* an, T* ab, T* at, T* ap) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int i = blockIdx.z;
  if (k >= d.n2 || j >= d.n1 || i >= d.n0) return;
  const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
  const int64_t x = mg_at(d, i, j, k);
  aw[x] = c.aw;
  ae[x] = c.ae;
  as[x] = c.as;
  an[x] = c.an;
  ab[x] = c.ab;
  at[x] = c.at;
  ap[x] = c.ap;
}

// ---------------------------------------------------------------------------
// Batched (ensemble) multigrid kernels over FP32 level pools: level l of
// member m is the (n_l+2)^3 array at pool + loff[l], spacing h * 2^l.
// ---------------------------------------------------------------------------
__device__ inline MgDev level_dev(const MemDev& M, int l) {
  MgDev d;
  d.n0 = M.ln[l][0];
  d.n1 = M.ln[l][1];
  d.n2 = M.ln[l][2];
  d.s1 = d.n2 + 2;
  d.s0 = (int64_t)(d.n1 + 2) * d.s1;
  return d;
}
// 1/h^2 of level l as the single-domain path forms it: (float)(1.0 / (h*h))
// with h = spacing * 2^l.
__device__ inline void level_coef(const MemDev& M, int l, float& c0, float& c1, float& c2) {
  const double f = (double)(1 << l);
  const double h0 = M.dx * f, h1 = M.dy * f, h2 = M.dz * f;
  c0 = (float)(1.0 / (h0 * h0));
  c1 = (float)(1.0 / (h1 * h1));
  c2 = (float)(1.0 / (h2 * h2));
}

template <bool SYNC, class F>
__device__ inline void mg_items(const Plan& P, const BView& V, const int* active, F&& f) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    if (active && !active[m]) continue;
    f(it - P.start[m], V.mem[m]);
    if (SYNC) __syncthreads();
  }
}

// restrict level lc-1 -> lc
__global__ void __launch_bounds__(256)
batch_restrict_kernel(BView V, Plan P, int lc, const float* phif, const float* rhsf,
                      float* phic, float* rhsc, const int* active) {
  mg_items<false>(P, V, active, [&](int q, const MemDev& M) {
    const MgDev f = level_dev(M, lc - 1), c = level_dev(M, lc);
    float c0, c1, c2;
    level_coef(M, lc - 1, c0, c1, c2);
    restrict_body<float>(phif + M.loff[lc - 1], rhsf + M.loff[lc - 1], f, c0, c1, c2,
                         phic + M.loff[lc], rhsc + M.loff[lc], c, cells_blk(q, c.n1, c.n2));
  });
}

// prolong level lf+1 -> lf
__global__ void __launch_bounds__(256)
batch_prolong_kernel(BView V, Plan P, int lf, const float* phic, float* phif,
                     const int* active) {
  mg_items<false>(P, V, active, [&](int q, const MemDev& M) {
    const MgDev c = level_dev(M, lf + 1), f = level_dev(M, lf);
    prolong_body<float>(phic + M.loff[lf + 1], c, phif + M.loff[lf], f,
                        cells_blk(q, f.n1, f.n2));
  });
}
// the same by 2 x 2 x 2 fine blocks (every member's level lf twice level
// lf+1), items over the coarse cells
__global__ void __launch_bounds__(256)
batch_prolong8_kernel(BView V, Plan P, int lf, const float* phic, float* phif,
                      const int* active) {
  mg_items<false>(P, V, active, [&](int q, const MemDev& M) {
    const MgDev c = level_dev(M, lf + 1), f = level_dev(M, lf);
    prolong8_body<float>(phic + M.loff[lf + 1], c, phif + M.loff[lf], f,
                         cells_blk(q, c.n1, c.n2));
  });
}

// one fused full sweep pout = S(pin) on level l, tile shape per member
__global__ void __launch_bounds__(256)
batch_sweep_kernel(BView V, Plan P, int l, const float* pin, float* pout, const float* rhs,
                   const int* active) {
  mg_items<true>(P, V, active, [&](int q, const MemDev& M) {
    const MgDev d = level_dev(M, l);
    float c0, c1, c2;
    level_coef(M, l, c0, c1, c2);
    const int TJ = sweep_tj(d.n2), TK = sweep_tk(d.n2);
    const int kt = (d.n2 + TK - 1) / TK, jt = (d.n1 + TJ - 1) / TJ;
    const Blk B{q % kt, (q / kt) % jt, q / (kt * jt)};
    const int64_t o = M.loff[l];
    if (TK == 64)
      sweep_body<float, 8, 64>(pin + o, pout + o, rhs + o, d, c0, c1, c2, BATCH_SWEEP_ICH, B);
    else
      sweep_body<float, 16, 32>(pin + o, pout + o, rhs + o, d, c0, c1, c2, BATCH_SWEEP_ICH, B);
  });
}

// nsweeps RBGS sweeps of level l, one LDS-resident workgroup per member
__global__ void __launch_bounds__(1024)
batch_relax_lds_kernel(BView V, int l, float* phi, const float* rhs, int nsweeps,
                       const int* active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int m = blockIdx.x;
  if (active && !active[m]) return;  // uniform per workgroup
  const MemDev& M = V.mem[m];
  const MgDev d = level_dev(M, l);
  float c0, c1, c2;
  level_coef(M, l, c0, c1, c2);
  relax_lds_body<float, true>(phi + M.loff[l], rhs + M.loff[l], d, c0, c1, c2, nsweeps,
                              reinterpret_cast<float*>(smem_raw));
}

// phi = M rhs on the coarsest level l, one workgroup per member
__global__ void __launch_bounds__(1024)
batch_coarse_apply_kernel(BView V, int l, const float* cop, float* phi, const float* rhs,
                          const int* active) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int m = blockIdx.x;
  if (active && !active[m]) return;
  const MemDev& M = V.mem[m];
  coarse_op_apply_body<float>(cop + M.coff, phi + M.loff[l], rhs + M.loff[l], level_dev(M, l),
                              reinterpret_cast<float*>(smem_raw));
}

int batch_lds_bytes() {
  int dev = 0, v = 65536;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    v = 65536;
  return v;
}

// Dynamic-LDS limits of the batched single-workgroup kernels, set once when a
// batch is created (never inside a captured step).
int batch_mg_init(int* lds_out) {
  const int lds = batch_lds_bytes();
  WF_TRY(hipFuncSetAttribute((const void*)batch_relax_lds_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  WF_TRY(hipFuncSetAttribute((const void*)batch_coarse_apply_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  *lds_out = lds;
  return 0;
}

// One batched V-cycle of the FP32 levels (vcycle_any's order).
int batch_vcycle(const Batch& B, float* const* phi, float* const* rhs, float* const* tmp,
                 const float* cop, int npre, int npost, int ncoarse, const int* active,
                 hipStream_t s) {
  const int nlev = B.nlev;
  const BView V = B.view();
  const int lds = B.lds;
  auto relax = [&](int l, int n) -> int {
    if (n <= 0) return 0;
    const size_t cells = (size_t)B.max_lds_cells[l];
    if (cells <= 4096 && 2 * cells * sizeof(float) <= (size_t)lds) {
      hipLaunchKernelGGL(batch_relax_lds_kernel, dim3(B.nmem), dim3(1024),
                         2 * cells * sizeof(float), s, V, l, phi[l], (const float*)rhs[l], n,
                         active);
      WF_LAUNCHED();
      return 0;
    }
    if (n % 2) return WF_EINVAL;  // the fused form runs sweeps in pairs
    const Plan& P = B.plan[PK_SWEEP0 + l];
    for (int q = 0; q < n; q += 2) {
      hipLaunchKernelGGL(batch_sweep_kernel, dim3(P.grid), dim3(256), 0, s, V, P, l,
                         (const float*)phi[l], tmp[l], (const float*)rhs[l], active);
      WF_LAUNCHED();
      hipLaunchKernelGGL(batch_sweep_kernel, dim3(P.grid), dim3(256), 0, s, V, P, l,
                         (const float*)tmp[l], phi[l], (const float*)rhs[l], active);
      WF_LAUNCHED();
    }
    return 0;
  };
  for (int l = 0; l < nlev; ++l) {
    if (l > 0) {
      const Plan& P = B.plan[PK_LCELLS0 + l];
      hipLaunchKernelGGL(batch_restrict_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, l,
                         (const float*)phi[l - 1], (const float*)rhs[l - 1], phi[l], rhs[l],
                         active);
      WF_LAUNCHED();
    }
    if (l > 0 && l == nlev - 1 && cop) {
      hipLaunchKernelGGL(batch_coarse_apply_kernel, dim3(B.nmem), dim3(1024),
                         (size_t)B.max_coarse_cells * sizeof(float), s, V, l, cop, phi[l],
                         (const float*)rhs[l], active);
      WF_LAUNCHED();
      continue;
    }
    const int r = relax(l, l < nlev - 1 ? npre : ncoarse);
    if (r) return r;
  }
  for (int l = nlev - 2; l >= 0; --l) {
    bool halves = true;  // every member's level l twice its level l+1
    for (int m = 0; m < B.nmem; ++m)
      for (int a = 0; a < 3; ++a)
        halves = halves && B.mem_host[m].ln[l][a] == 2 * B.mem_host[m].ln[l + 1][a];
    if (halves) {
      const Plan& P = B.plan[PK_LCELLS0 + l + 1];
      hipLaunchKernelGGL(batch_prolong8_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, l,
                         (const float*)phi[l + 1], phi[l], active);
    } else {
      const Plan& P = B.plan[PK_LCELLS0 + l];
      hipLaunchKernelGGL(batch_prolong_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, l,
                         (const float*)phi[l + 1], phi[l], active);
    }
    WF_LAUNCHED();
    const int r = relax(l, npost);
    if (r) return r;
  }
  return 0;
}

}  // namespace wf

using namespace wf;

namespace {

bool level_ok(const wf_mg_level* L) {
  return L && L->phi && L->rhs && L->n0 > 0 && L->n1 > 0 && L->n2 > 0 &&
         (L->prec == 8 || L->prec == 4) && L->h0 > 0 && L->h1 > 0 && L->h2 > 0;
}

int64_t level_elems(const wf_mg_level& L) {
  return (int64_t)(L.n0 + 2) * (L.n1 + 2) * (L.n2 + 2);
}

// 1/h^2 exactly as _mgcoeff forms it: 1.0 / (h*h) in float64
// (multigrid.py:34-38); the float32 form rounds that value.
template <class T>
void inv_h2(const wf_mg_level& L, T& c0, T& c1, T& c2) {
  c0 = (T)(1.0 / (L.h0 * L.h0));
  c1 = (T)(1.0 / (L.h1 * L.h1));
  c2 = (T)(1.0 / (L.h2 * L.h2));
}

dim3 cell_grid(int n0, int n1, int n2) { return dim3(div_up(n2, 64), div_up(n1, 4), n0); }

int g_max_lds = -1;

int max_lds() {
  if (g_max_lds < 0) {
    int dev = 0, v = 65536;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
            hipSuccess)
      g_max_lds = v;
    else
      g_max_lds = 65536;
  }
  return g_max_lds;
}

template <class T>
int launch_rbgs(const wf_mg_level& L, int color, int write_ring, hipStream_t s) {
  const MgDev d = mg_dev(L);
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  hipLaunchKernelGGL(rbgs_kernel<T>, cell_grid(d.n0, d.n1, (d.n2 + 1) / 2), dim3(64, 4), 0, s,
                     (T*)L.phi, (const T*)L.rhs, d, c0, c1, c2, color, write_ring);
  WF_LAUNCHED();
  return 0;
}

// Packed single-workgroup relax: bytes of LDS and cells per thread per colour
// (0 when the level does not fit).
inline size_t relax_lds2_bytes(const wf_mg_level& L, size_t elem) {
  return 2 * (size_t)(L.n0 + 2) * (L.n1 + 2) * ((L.n2 + 3) / 2) * elem;
}
inline int relax_lds2_cpt(const wf_mg_level& L) {
  const int64_t nslot = (int64_t)L.n0 * L.n1 * ((L.n2 + 1) / 2);
  for (int cpt = 1; cpt <= 4; cpt *= 2)  // more cells per thread spill registers
    if (nslot <= (int64_t)cpt * 1024) return cpt;
  return 0;
}

template <class T, int CPT>
int launch_relax_lds2_t(const wf_mg_level& L, int nsweeps, size_t bytes, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    WF_TRY(hipFuncSetAttribute((const void*)relax_lds2_kernel<T, CPT>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, max_lds()));
    attr_set = true;
  }
  const MgDev d = mg_dev(L);
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  hipLaunchKernelGGL((relax_lds2_kernel<T, CPT>), dim3(1), dim3(1024), bytes, s, (T*)L.phi,
                     (const T*)L.rhs, d, c0, c1, c2, nsweeps);
  WF_LAUNCHED();
  return 0;
}

template <class T>
int launch_relax_lds2(const wf_mg_level& L, int nsweeps, hipStream_t s) {
  const size_t bytes = relax_lds2_bytes(L, sizeof(T));
  switch (relax_lds2_cpt(L)) {
    case 1: return launch_relax_lds2_t<T, 1>(L, nsweeps, bytes, s);
    case 2: return launch_relax_lds2_t<T, 2>(L, nsweeps, bytes, s);
    case 4: return launch_relax_lds2_t<T, 4>(L, nsweeps, bytes, s);
    default: return WF_EINVAL;
  }
}

template <class T, bool RHS_LDS>
int launch_relax_lds(const wf_mg_level& L, int nsweeps, size_t bytes, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    WF_TRY(hipFuncSetAttribute((const void*)relax_lds_kernel<T, RHS_LDS>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, max_lds()));
    attr_set = true;
  }
  const MgDev d = mg_dev(L);
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  const int ncell = d.n0 * d.n1 * d.n2;
  const int nt = ncell >= 4096 ? 1024 : (ncell >= 1024 ? 512 : 256);
  hipLaunchKernelGGL((relax_lds_kernel<T, RHS_LDS>), dim3(1), dim3(nt), bytes, s, (T*)L.phi,
                     (const T*)L.rhs, d, c0, c1, c2, nsweeps);
  WF_LAUNCHED();
  return 0;
}

// Levels up to LDS_RELAX_MAX_CELLS cells -- or any level that fits LDS when
// asked for many sweeps (the coarsest solve) -- relax in one LDS-resident
// workgroup; the rest in fused full-sweep launches (or half-sweep launches
// without tmp).
constexpr size_t LDS_RELAX_MAX_CELLS = 4096;
constexpr int LDS_RELAX_MIN_SWEEPS = 32;

template <class T, int TJ, int TK>
int launch_sweep(const wf_mg_level& L, const void* pin, void* pout, hipStream_t s) {
  const MgDev d = mg_dev(L);
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  const int kt = div_up(d.n2, TK), jt = div_up(d.n1, TJ);
  const int64_t tiles = (int64_t)kt * jt;
  const int tuned = tune_get(sizeof(T) == 8 ? WF_TUNE_SWEEP_ICH_F64 : WF_TUNE_SWEEP_ICH_F32,
                             d.n0, d.n1, d.n2);
  const int64_t pb = d.s0 * (int64_t)sizeof(T);
  if (tuned == -2 || tuned == -4) {  // chunks of 2 or 4 planes in one pass each
    const int nb = -tuned;
    // the kernel's plane offsets within a chunk (nb + 4 planes) stay below 2^30
    if ((int64_t)(nb + 4) * pb >= ((int64_t)1 << 30)) return WF_EINVAL;
    const dim3 grid(kt, jt, div_up(d.n0, nb));
    if (nb == 2)
      hipLaunchKernelGGL((sweep_kernel<T, TJ, TK, 2>), grid, dim3(256), 0, s, (const T*)pin,
                         (T*)pout, (const T*)L.rhs, d, c0, c1, c2, nb);
    else
      hipLaunchKernelGGL((sweep_kernel<T, TJ, TK, 4>), grid, dim3(256), 0, s, (const T*)pin,
                         (T*)pout, (const T*)L.rhs, d, c0, c1, c2, nb);
    WF_LAUNCHED();
    return 0;
  }
  int ich = tuned > 0 ? tuned
                      : (int)std::max<int64_t>(4, std::min<int64_t>(16, tiles * d.n0 / 2048));
  // the kernel's plane offsets within a chunk ((ich + 3) planes) stay below 2^30
  if ((int64_t)4 * pb >= ((int64_t)1 << 30)) return WF_EINVAL;
  ich = (int)std::min<int64_t>(ich, ((int64_t)1 << 30) / pb - 3);
  const dim3 grid(kt, jt, div_up(d.n0, ich));
  hipLaunchKernelGGL((sweep_kernel<T, TJ, TK>), grid, dim3(256), 0, s, (const T*)pin, (T*)pout,
                     (const T*)L.rhs, d, c0, c1, c2, ich);
  WF_LAUNCHED();
  return 0;
}

template <class T>
int sweep_t(const wf_mg_level& L, const void* pin, void* pout, hipStream_t s) {
  const int tile = tune_get(sizeof(T) == 8 ? WF_TUNE_SWEEP_TILE_F64 : WF_TUNE_SWEEP_TILE_F32,
                            L.n0, L.n1, L.n2);
  switch (tile) {  // (j x k): 1 = 8 x 64, 2 = 16 x 32, 3 = 16 x 64, 4 = 32 x 32
    case 1:
      return launch_sweep<T, 8, 64>(L, pin, pout, s);
    case 2:
      return launch_sweep<T, 16, 32>(L, pin, pout, s);
    case 3:
      return launch_sweep<T, 16, 64>(L, pin, pout, s);
    case 4:
      return launch_sweep<T, 32, 32>(L, pin, pout, s);
    default:
      return L.n2 > 32 ? launch_sweep<T, 8, 64>(L, pin, pout, s)
                       : launch_sweep<T, 16, 32>(L, pin, pout, s);
  }
}

template <class T, int TJ, int TK>
int launch_sweep2(const wf_mg_level& L, const void* pin, void* pout, hipStream_t s) {
  const MgDev d = mg_dev(L);
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  const int kt = div_up(d.n2, TK), jt = div_up(d.n1, TJ);
  const int64_t tiles = (int64_t)kt * jt;
  const int tuned = tune_get(sizeof(T) == 8 ? WF_TUNE_SWEEP2_ICH_F64 : WF_TUNE_SWEEP2_ICH_F32,
                             d.n0, d.n1, d.n2);
  // the kernel's plane offsets within a chunk ((ich + 8) planes) stay below 2^30
  const int64_t pb = d.s0 * (int64_t)sizeof(T);
  if ((int64_t)9 * pb >= ((int64_t)1 << 30)) return WF_EINVAL;
  int ich = tuned > 0 ? tuned
                      : (int)std::max<int64_t>(4, std::min<int64_t>(32, tiles * d.n0 / 1024));
  ich = (int)std::min<int64_t>(ich, ((int64_t)1 << 30) / pb - 8);
  const dim3 grid(kt, jt, div_up(d.n0, ich));
  hipLaunchKernelGGL((sweep2_kernel<T, TJ, TK>), grid, dim3(256), 0, s, (const T*)pin, (T*)pout,
                     (const T*)L.rhs, d, c0, c1, c2, ich);
  WF_LAUNCHED();
  return 0;
}

template <class T>
int sweep2_t(const wf_mg_level& L, const void* pin, void* pout, hipStream_t s) {
  const int tile = tune_get(sizeof(T) == 8 ? WF_TUNE_SWEEP2_TILE_F64 : WF_TUNE_SWEEP2_TILE_F32,
                            L.n0, L.n1, L.n2);
  if constexpr (sizeof(T) == 4) {  // two pairs per thread: FP32 only (LDS, registers)
    if (tile == 3) return launch_sweep2<T, 16, 64>(L, pin, pout, s);
    if (tile == 4) return launch_sweep2<T, 32, 32>(L, pin, pout, s);
  }
  return tile == 1 ? launch_sweep2<T, 8, 64>(L, pin, pout, s)
                   : launch_sweep2<T, 16, 32>(L, pin, pout, s);
}

// nsweeps RBGS sweeps of L.  Pairs of sweeps run as two-sweep launches
// (phi -> tmp, then tmp -> phi, ...); with allow_swap an odd number of them
// may leave the result in tmp, which *swapped reports (the caller then
// treats tmp as the level's phi); without it the pairs go in even numbers
// and the rest as one-sweep launches, so the result is in phi.
// Which smoother relax_t runs for nsweeps sweeps of level L.
enum class RelaxPath { LDS2, LDS_RHS, LDS, SWEEPS };
template <class T>
RelaxPath relax_path(const wf_mg_level& L, int nsweeps) {
  const size_t cells = (size_t)L.n0 * L.n1 * L.n2;
  const size_t lds = (size_t)max_lds();
  // levels whose zero-ringed copy fits LDS: all sweeps in one workgroup
  if (relax_lds2_cpt(L) > 0 && relax_lds2_bytes(L, sizeof(T)) <= lds) return RelaxPath::LDS2;
  // many sweeps in one workgroup pay off only with rhs in LDS too: with rhs
  // read from global memory every colour pass waits on its latency
  if ((cells <= LDS_RELAX_MAX_CELLS || nsweeps >= LDS_RELAX_MIN_SWEEPS) &&
      2 * cells * sizeof(T) <= lds)
    return RelaxPath::LDS_RHS;
  if (cells <= LDS_RELAX_MAX_CELLS && cells * sizeof(T) <= lds) return RelaxPath::LDS;
  return RelaxPath::SWEEPS;
}
// True when relax_t's first launch for nsweeps sweeps of L is a full-sweep
// kernel, which can take a zero iterate without reading it (zero_start).
template <class T>
bool relax_zero_start_ok(const wf_mg_level& L, int nsweeps) {
  return nsweeps >= 2 && L.tmp && relax_path<T>(L, nsweeps) == RelaxPath::SWEEPS;
}

// zero_start: the iterate in L.phi is to be taken as zero (the caller has
// not zeroed it); only valid when relax_zero_start_ok.
template <class T>
int relax_t(const wf_mg_level& L, int nsweeps, hipStream_t s, bool allow_swap = false,
            bool* swapped = nullptr, bool zero_start = false) {
  if (swapped) *swapped = false;
  if (nsweeps <= 0) return 0;
  const size_t cells = (size_t)L.n0 * L.n1 * L.n2;
  switch (relax_path<T>(L, nsweeps)) {
    case RelaxPath::LDS2:
      return launch_relax_lds2<T>(L, nsweeps, s);
    case RelaxPath::LDS_RHS:
      return launch_relax_lds<T, true>(L, nsweeps, 2 * cells * sizeof(T), s);
    case RelaxPath::LDS:
      return launch_relax_lds<T, false>(L, nsweeps, cells * sizeof(T), s);
    case RelaxPath::SWEEPS:
      break;
  }
  if (zero_start && !relax_zero_start_ok<T>(L, nsweeps)) return WF_EINVAL;
  wf_mg_level C = L;  // C.phi holds the current iterate
  // the first launch reads the zero iterate through an empty descriptor
  auto src = [&](const void* p) -> const void* {
    const void* r = zero_start ? nullptr : p;
    zero_start = false;
    return r;
  };
  int q = 0;
  const int pairs = tune_get(sizeof(T) == 8 ? WF_TUNE_SWEEP_PAIRS_F64 : WF_TUNE_SWEEP_PAIRS_F32,
                             L.n0, L.n1, L.n2);
  // built-in: two-sweep launches where they measured faster (FP64, >= 2^23 cells)
  if (C.tmp && (pairs == 2 || (pairs == 0 && sizeof(T) == 8 && cells >= ((size_t)1 << 23)))) {
    int npair = nsweeps / 2;
    if (!allow_swap) npair &= ~1;
    for (int f = 0; f < npair; ++f) {
      WF_TRY((hipError_t)sweep2_t<T>(C, src(C.phi), C.tmp, s));
      std::swap(C.phi, C.tmp);
    }
    q = 2 * npair;
  }
  if (C.tmp) {
    for (; q + 1 < nsweeps; q += 2) {
      WF_TRY((hipError_t)sweep_t<T>(C, src(C.phi), C.tmp, s));
      WF_TRY((hipError_t)sweep_t<T>(C, C.tmp, C.phi, s));
    }
  }
  for (; q < nsweeps; ++q) {
    WF_TRY((hipError_t)launch_rbgs<T>(C, 0, 1, s));
    WF_TRY((hipError_t)launch_rbgs<T>(C, 1, 1, s));
  }
  if (swapped) *swapped = C.phi != L.phi;
  return 0;
}

template <class T>
int restrict_t(const wf_mg_level& F, const wf_mg_level& C, hipStream_t s) {
  const MgDev f = mg_dev(F), c = mg_dev(C);
  T c0, c1, c2;
  inv_h2(F, c0, c1, c2);
  const int ich = plane_chunk(c.n0, c.n1, c.n2);
  hipLaunchKernelGGL(restrict_kernel<T>, chunk_grid(c.n0, c.n1, c.n2, ich), dim3(64, 4), 0, s,
                     (const T*)F.phi, (const T*)F.rhs, f, c0, c1, c2, (T*)C.phi, (T*)C.rhs, c,
                     ich);
  WF_LAUNCHED();
  return 0;
}

template <class T>
int prolong_t(const wf_mg_level& C, const wf_mg_level& F, hipStream_t s) {
  const MgDev f = mg_dev(F), c = mg_dev(C);
  // chunks sized for ~2048 workgroups (1..8 planes): measured on the 256^3
  // hierarchy, 8 planes at the 256^3 fine level (as plane_chunk), 4 instead
  // of 1 at 128^3 (8.9 vs 13.1 us), 1 below; each cell's value does not
  // depend on the chunking
  if (f.n0 == 2 * c.n0 && f.n1 == 2 * c.n1 && f.n2 == 2 * c.n2) {  // 2 x 2 x 2 blocks
    const int64_t tiles = (int64_t)div_up(c.n2, BX) * div_up(c.n1, BY);
    const int ich = (int)std::min<int64_t>(8, std::max<int64_t>(1, tiles * c.n0 / 2048));
    hipLaunchKernelGGL(prolong8_kernel<T>, chunk_grid(c.n0, c.n1, c.n2, ich), dim3(64, 4), 0,
                       s, (const T*)C.phi, c, (T*)F.phi, f, ich);
    WF_LAUNCHED();
    return 0;
  }
  const int64_t tiles = (int64_t)div_up(f.n2, BX) * div_up(f.n1, BY);
  const int ich = (int)std::min<int64_t>(8, std::max<int64_t>(1, tiles * f.n0 / 2048));
  hipLaunchKernelGGL(prolong_kernel<T>, chunk_grid(f.n0, f.n1, f.n2, ich), dim3(64, 4), 0, s,
                     (const T*)C.phi, c, (T*)F.phi, f, ich);
  WF_LAUNCHED();
  return 0;
}

template <class T>
int zero_t(void* p, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(zero_kernel<T>, dim3(div_up(n, 256)), dim3(256), 0, s, (T*)p, n);
  WF_LAUNCHED();
  return 0;
}

bool pair_ok(const wf_mg_level* F, const wf_mg_level* C) {
  return level_ok(F) && level_ok(C) && F->prec == C->prec && F->n0 == 2 * C->n0 &&
         F->n1 == 2 * C->n1 && F->n2 == 2 * C->n2;
}

int relax_any(const wf_mg_level& L, int n, hipStream_t s, bool allow_swap = false,
              bool* swapped = nullptr, bool zero_start = false) {
  return L.prec == 8 ? relax_t<double>(L, n, s, allow_swap, swapped, zero_start)
                     : relax_t<float>(L, n, s, allow_swap, swapped, zero_start);
}
bool relax_zero_start_any(const wf_mg_level& L, int n) {
  return L.prec == 8 ? relax_zero_start_ok<double>(L, n) : relax_zero_start_ok<float>(L, n);
}

// Coarse operators: up to COARSE_OP_SERIAL_MAX_CELLS cells each row is one
// thread's serial sum (the form the batched solve shares bit for bit); above
// it, up to COARSE_OP_MAX_CELLS, the split form (coarse_op_split_kernel),
// whose chunk sums sit in a tail of the operator buffer after the ncell^2
// matrix entries.
constexpr int COARSE_OP_SERIAL_MAX_CELLS = 4096;
constexpr int COARSE_OP_MAX_CELLS = 16384;

// column chunks of the split form: about 2048 workgroups, at most 64 chunks,
// at most COP_SPLIT_MAX_CLEN columns each
inline int cop_split_chunks(int ncell) {
  const int rb = div_up(ncell, COP_SPLIT_ROWS);
  int nch = std::max(1, std::min(64, div_up(2048, rb)));
  while (div_up(ncell, nch) > COP_SPLIT_MAX_CLEN) ++nch;
  return nch;
}
inline size_t cop_elems(size_t ncell) {
  size_t e = ncell * ncell;
  if (ncell > (size_t)COARSE_OP_SERIAL_MAX_CELLS) e += (size_t)cop_split_chunks((int)ncell) * ncell;
  return e;
}

template <class T>
int coarse_op_apply_t(const wf_mg_level& L, hipStream_t s) {
  const MgDev d = mg_dev(L);
  const int ncell = d.n0 * d.n1 * d.n2;
  const size_t bytes = (size_t)ncell * sizeof(T);
  if (ncell > COARSE_OP_MAX_CELLS) return WF_EINVAL;
  if (ncell > COARSE_OP_SERIAL_MAX_CELLS) {
    const int nch = cop_split_chunks(ncell), clen = div_up(ncell, nch);
    T* part = (T*)L.cop + (size_t)ncell * ncell;
    hipLaunchKernelGGL(coarse_op_split_kernel<T>,
                       dim3((unsigned)div_up(ncell, COP_SPLIT_ROWS), (unsigned)nch),
                       dim3(COP_SPLIT_ROWS), 0, s, (const T*)L.cop, (const T*)L.rhs, d, clen,
                       part);
    WF_LAUNCHED();
    hipLaunchKernelGGL(coarse_op_fold_kernel<T>, dim3((unsigned)div_up(ncell, 256)), dim3(256),
                       0, s, (const T*)part, nch, (T*)L.phi, d);
    WF_LAUNCHED();
    return 0;
  }
  if (bytes > (size_t)max_lds()) return WF_EINVAL;
  static bool attr_set = false;
  if (!attr_set) {
    WF_TRY(hipFuncSetAttribute((const void*)coarse_op_apply_rows_kernel<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, max_lds()));
    attr_set = true;
  }
  hipLaunchKernelGGL(coarse_op_apply_rows_kernel<T>, dim3((unsigned)div_up(ncell, 64)),
                     dim3(64), bytes, s, (const T*)L.cop, (T*)L.phi, (const T*)L.rhs, d);
  WF_LAUNCHED();
  return 0;
}

int coarse_op_apply_any(const wf_mg_level& L, hipStream_t s) {
  return L.prec == 8 ? coarse_op_apply_t<double>(L, s) : coarse_op_apply_t<float>(L, s);
}

template <class T>
int coarse_op_build_t(const wf_mg_level& L, int nsweeps, void* op, hipStream_t s) {
  const MgDev d = mg_dev(L);
  const int ncell = d.n0 * d.n1 * d.n2;
  const size_t bytes = (size_t)ncell * sizeof(T);
  if (ncell > COARSE_OP_MAX_CELLS || bytes > (size_t)max_lds()) return WF_EINVAL;
  static bool attr_set = false;
  if (!attr_set) {
    WF_TRY(hipFuncSetAttribute((const void*)coarse_op_build_kernel<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, max_lds()));
    attr_set = true;
  }
  T c0, c1, c2;
  inv_h2(L, c0, c1, c2);
  hipLaunchKernelGGL(coarse_op_build_kernel<T>, dim3(ncell), dim3(256), bytes, s, (T*)op, d, c0,
                     c1, c2, nsweeps);
  WF_LAUNCHED();
  return 0;
}

template <class Relax>
int vcycle_levels(const wf_mg_level* lv, int nlev, int npre, int npost, int ncoarse,
                  hipStream_t s, Relax&& relax_l, bool zero0 = false) {
  for (int l = 0; l < nlev; ++l) {
    if (l > 0) {
      const int r = lv[l].prec == 8 ? restrict_t<double>(lv[l - 1], lv[l], s)
                                    : restrict_t<float>(lv[l - 1], lv[l], s);
      if (r) return r;
    }
    if (l > 0 && l == nlev - 1 && lv[l].cop && lv[l].cop_sweeps == ncoarse && ncoarse > 0) {
      // coarsest level right after restriction zeroed phi: phi = M rhs
      const int r = coarse_op_apply_any(lv[l], s);
      if (r) return r;
      continue;
    }
    const int r = relax_l(l, l < nlev - 1 ? npre : ncoarse, l == 0 && zero0);
    if (r) return r;
  }
  for (int l = nlev - 2; l >= 0; --l) {
    const int r = lv[l].prec == 8 ? prolong_t<double>(lv[l + 1], lv[l], s)
                                  : prolong_t<float>(lv[l + 1], lv[l], s);
    if (r) return r;
    const int r2 = relax_l(l, npost, false);
    if (r2) return r2;
  }
  return 0;
}

// One V-cycle.  The levels are worked on through a local view whose phi and
// tmp swap when a relax leaves its result in tmp (an odd number of two-sweep
// launches); a level still swapped at the end is copied back, so the caller
// sees every level's iterate in its phi (with V(6,6): 3 + 3 launches, none).
// zero0: level 0's iterate is to be taken as zero without reading it (the
// caller checked relax_zero_start_any(lv0[0], npre) and skipped its zeroing).
int vcycle_any(const wf_mg_level* lv0, int nlev, int npre, int npost, int ncoarse,
               hipStream_t s, bool zero0 = false) {
  wf_mg_level lv[16];
  for (int l = 0; l < nlev; ++l) lv[l] = lv0[l];
  auto relax = [&](int l, int n, bool zs) -> int {
    bool sw = false;
    const int r = relax_any(lv[l], n, s, true, &sw, zs);
    if (sw) std::swap(lv[l].phi, lv[l].tmp);
    return r;
  };
  const int r = vcycle_levels(lv, nlev, npre, npost, ncoarse, s, relax, zero0);
  if (r) return r;
  for (int l = 0; l < nlev; ++l)
    if (lv[l].phi != lv0[l].phi)
      WF_TRY(hipMemcpyAsync(lv0[l].phi, lv[l].phi, (size_t)level_elems(lv0[l]) * lv0[l].prec,
                            hipMemcpyDeviceToDevice, s));
  return 0;
}

bool hier_ok(const wf_mg_level* lv, int nlev) {
  if (!lv || nlev < 1 || nlev > 16) return false;
  for (int l = 0; l < nlev; ++l) {
    if (!level_ok(&lv[l])) return false;
    if (l > 0 && !pair_ok(&lv[l - 1], &lv[l])) return false;
  }
  return true;
}

bool field_grid_ok(const wf_grid& g) {
  return g.ni > 0 && g.nj > 0 && g.nk > 0 && g.npl >= 1 &&
         g.sj >= (int64_t)(g.nk + 2 * g.npl) && g.si >= g.sj * (int64_t)(g.nj + 2 * g.npl);
}

}  // namespace

namespace wf {
bool mg_level_ok(const wf_mg_level* L) { return level_ok(L); }
int64_t mg_level_elems(const wf_mg_level& L) { return level_elems(L); }
dim3 mg_cell_grid(int n0, int n1, int n2) { return cell_grid(n0, n1, n2); }
int mg_vcycle_levels(const wf_mg_level* lv, int nlev, int npre, int npost, int ncoarse,
                     hipStream_t s, bool zero0) {
  return vcycle_any(lv, nlev, npre, npost, ncoarse, s, zero0);
}
bool mg_relax_zero_start_ok(const wf_mg_level& L, int nsweeps) {
  return relax_zero_start_any(L, nsweeps);
}
bool mg_hier_ok(const wf_mg_level* lv, int nlev) { return hier_ok(lv, nlev); }
bool mg_field_grid_ok(const wf_grid& g) { return field_grid_ok(g); }
template <class T>
int mg_zero(void* p, int64_t n, hipStream_t s) {
  return zero_t<T>(p, n, s);
}
template int mg_zero<double>(void*, int64_t, hipStream_t);
template int mg_zero<float>(void*, int64_t, hipStream_t);
}  // namespace wf

#define S_(s) ((hipStream_t)(s))

extern "C" {

int wf_mg_coeff(const wf_mg_level* L, void* aw, void* ae, void* as, void* an, void* ab,
                void* at, void* ap, void* stream) {
  if (!level_ok(L)) return WF_EINVAL;
  const MgDev d = mg_dev(*L);
  if (L->prec == 8) {
    double c0, c1, c2;
    inv_h2(*L, c0, c1, c2);
    hipLaunchKernelGGL(coeff_kernel<double>, cell_grid(d.n0, d.n1, d.n2), dim3(64, 4), 0,
                       S_(stream), d, c0, c1, c2, (double*)aw, (double*)ae, (double*)as,
                       (double*)an, (double*)ab, (double*)at, (double*)ap);
  } else {
    float c0, c1, c2;
    inv_h2(*L, c0, c1, c2);
    hipLaunchKernelGGL(coeff_kernel<float>, cell_grid(d.n0, d.n1, d.n2), dim3(64, 4), 0,
                       S_(stream), d, c0, c1, c2, (float*)aw, (float*)ae, (float*)as,
                       (float*)an, (float*)ab, (float*)at, (float*)ap);
  }
  WF_LAUNCHED();
  return 0;
}

int wf_mg_coarse_op_bytes(const wf_mg_level* L, size_t* bytes) {
  if (!level_ok(L) || !bytes) return WF_EINVAL;
  const size_t ncell = (size_t)L->n0 * L->n1 * L->n2;
  if (ncell > (size_t)COARSE_OP_MAX_CELLS || ncell * L->prec > (size_t)max_lds())
    return WF_EINVAL;
  *bytes = cop_elems(ncell) * (size_t)L->prec;
  return 0;
}

int wf_mg_coarse_op_build(const wf_mg_level* L, int nsweeps, void* op, void* stream) {
  if (!level_ok(L) || !op || nsweeps < 0) return WF_EINVAL;
  return L->prec == 8 ? coarse_op_build_t<double>(*L, nsweeps, op, S_(stream))
                      : coarse_op_build_t<float>(*L, nsweeps, op, S_(stream));
}

int wf_mg_coarse_op_apply(const wf_mg_level* L, void* stream) {
  if (!level_ok(L) || !L->cop) return WF_EINVAL;
  return coarse_op_apply_any(*L, S_(stream));
}

int wf_mg_rbgs(const wf_mg_level* L, int color, void* stream) {
  if (!level_ok(L) || (color != 0 && color != 1)) return WF_EINVAL;
  return L->prec == 8 ? launch_rbgs<double>(*L, color, 0, S_(stream))
                      : launch_rbgs<float>(*L, color, 0, S_(stream));
}

int wf_mg_refill(const wf_mg_level* L, void* stream) {
  if (!level_ok(L)) return WF_EINVAL;
  const MgDev d = mg_dev(*L);
  const int nthreads = 2 * (d.n1 * d.n2 + d.n0 * d.n2 + d.n0 * d.n1);
  if (L->prec == 8)
    hipLaunchKernelGGL(refill_kernel<double>, dim3(div_up(nthreads, 256)), dim3(256), 0,
                       S_(stream), (double*)L->phi, d);
  else
    hipLaunchKernelGGL(refill_kernel<float>, dim3(div_up(nthreads, 256)), dim3(256), 0,
                       S_(stream), (float*)L->phi, d);
  WF_LAUNCHED();
  return 0;
}

int wf_mg_relax(const wf_mg_level* L, int nsweeps, void* stream) {
  if (!level_ok(L) || nsweeps < 0) return WF_EINVAL;
  return relax_any(*L, nsweeps, S_(stream));
}

int wf_mg_restrict(const wf_mg_level* F, const wf_mg_level* C, void* stream) {
  if (!pair_ok(F, C)) return WF_EINVAL;
  return F->prec == 8 ? restrict_t<double>(*F, *C, S_(stream))
                      : restrict_t<float>(*F, *C, S_(stream));
}

int wf_mg_prolong(const wf_mg_level* C, const wf_mg_level* F, void* stream) {
  if (!pair_ok(F, C)) return WF_EINVAL;
  return F->prec == 8 ? prolong_t<double>(*C, *F, S_(stream))
                      : prolong_t<float>(*C, *F, S_(stream));
}

int wf_mg_copy_in(const double* src, wf_grid g, const wf_mg_level* L, void* stream) {
  if (!level_ok(L) || L->prec != 8 || !field_grid_ok(g) || g.ni != L->n0 || g.nj != L->n1 ||
      g.nk != L->n2)
    return WF_EINVAL;
  const MgDev d = mg_dev(*L);
  hipLaunchKernelGGL(copy_in_kernel, cell_grid(d.n0, d.n1, d.n2), dim3(64, 4), 0, S_(stream),
                     src, g, (double*)L->rhs, d);
  WF_LAUNCHED();
  return 0;
}

int wf_mg_copy_out_ring(const wf_mg_level* L, double* dst, wf_grid g, void* stream) {
  if (!level_ok(L) || L->prec != 8 || !field_grid_ok(g) || g.ni != L->n0 ||
      g.nj != L->n1 || g.nk != L->n2)
    return WF_EINVAL;
  const MgDev d = mg_dev(*L);
  hipLaunchKernelGGL(copy_out_kernel, cell_grid(d.n0 + 2, d.n1 + 2, d.n2 + 2), dim3(64, 4), 0,
                     S_(stream), (const double*)L->phi, d, dst, g);
  WF_LAUNCHED();
  return 0;
}

int wf_mg_vcycle(const wf_mg_level* levels, int nlev, int n_pre, int n_post, int n_coarse,
                 void* stream) {
  if (!hier_ok(levels, nlev) || n_pre < 0 || n_post < 0 || n_coarse < 0) return WF_EINVAL;
  return vcycle_any(levels, nlev, n_pre, n_post, n_coarse, S_(stream));
}

int wf_mg_solve(const wf_mg_level* levels, int nlev, const double* su, double* pp, wf_grid g,
                int n_pre, int n_post, int n_coarse, int n_vcycles, void* stream) {
  if (!hier_ok(levels, nlev) || levels[0].prec != 8 || n_vcycles < 0) return WF_EINVAL;
  const hipStream_t s = S_(stream);
  int r = wf_mg_copy_in(su, g, &levels[0], stream);
  if (r) return r;
  // pp: the copy-out below writes its real cells and its whole ring; its
  // deeper ghost layers are not written (they stay zero, as the reference's
  // zeroed pp leaves them).  phi = 0: the first sweep of the first V-cycle
  // takes the zero iterate without reading it where its smoother allows.
  const bool zs =
      n_vcycles > 0 && relax_zero_start_any(levels[0], nlev > 1 ? n_pre : n_coarse);
  if (!zs) {
    r = zero_t<double>(levels[0].phi, level_elems(levels[0]), s);
    if (r) return r;
  }
  for (int c = 0; c < n_vcycles; ++c) {
    r = vcycle_any(levels, nlev, n_pre, n_post, n_coarse, s, c == 0 && zs);
    if (r) return r;
  }
  return wf_mg_copy_out_ring(&levels[0], pp, g, stream);
}

}  // extern "C"
