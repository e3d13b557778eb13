// Mixed-precision pressure solve: FP32 RBGS V-cycles on the error equation
// inside FP64 defect correction.
//
//   x = 0 (FP64, the caller's pp field)
//   repeat:  r = b - A x            FP64 defect, matrix-free, on the field
//            [r -= mean(b)]          FP64 mean removal (optional)
//            e = V-cycle_32(r)       FP32, from e = 0, same V(pre, post) and
//                                    coarse sweeps as the plain form
//            x += e                  FP64 update
//   [x -= mean(x)]                   FP64 pressure-gauge removal (optional)
//   ring faces of x = Neumann mirror (the state the plain solve leaves)
//
// A V-cycle started from e = 0 on the defect is the plain iteration
// x <- x + M(b - A x), so with the same cycle count the result equals the
// FP64 V-cycles up to FP32 rounding of each correction, while every residual,
// update and norm stays FP64.  In tolerance mode the loop stops once
// max|b - A x| <= rtol * max|b| (the reference's residual norm,
// tests/test_mg.py:16-32).
//
// Sums are deterministic: fixed-shape block reductions into a partials
// buffer, then one workgroup folds the partials in a fixed order.
#include <algorithm>
#include <mutex>

#include "batch.h"
#include "mg_common.h"

namespace wf {

// MIX_MEAN_DEFERRED (with MIX_RESULT_MEAN): the mean of x is formed into
// work[2] but not subtracted; the caller applies it (wf_add_ring_shift).
// MIX_RHS_READY: levels[0].rhs already holds (float)b on the real cells.
enum : int { MIX_DEFECT_MEAN = 1, MIX_RESULT_MEAN = 2, MIX_MEAN_DEFERRED = 4,
             MIX_RHS_READY = 8 };

constexpr int WORK_HEAD = 4;  // work[0] mean, work[1] max|r|, work[2] sum, work[3] spare

inline dim3 field_grid(const wf_grid& g) {
  return dim3(div_up(g.nk, 64), div_up(g.nj, 4), g.ni);
}
inline int field_ich(const wf_grid& g) { return plane_chunk(g.ni, g.nj, g.nk); }
inline dim3 field_cgrid(const wf_grid& g) { return chunk_grid(g.ni, g.nj, g.nk, field_ich(g)); }
inline int64_t field_blocks(const wf_grid& g) {
  const dim3 b = field_grid(g);
  return (int64_t)b.x * b.y * b.z;
}

// Linear index of block B of a field_grid launch (x fastest).
__device__ inline int64_t block_linear(const wf_grid& g, Blk B) {
  const int gx = (g.nk + 63) / 64, gy = (g.nj + 3) / 4;
  return ((int64_t)B.z * gy + B.y) * gx + B.x;
}

// Block sum of 256 values in a fixed order; every thread must call it.
__device__ inline double block_sum_fixed(double v) {
  __shared__ double part[4];
  v = wave_sum(v);
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) part[t >> 6] = v;
  __syncthreads();
  return ((part[0] + part[1]) + part[2]) + part[3];
}

// Partial sums of f over the real cells of a field, one per block.
__device__ inline void field_partials_body(const double* __restrict__ f, const wf_grid& g,
                                           double* __restrict__ partials, Blk B) {
  const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
  double v = 0.0;
  if (k < g.nk && j < g.nj) v = f[at(g, i + g.npl, j + g.npl, k + g.npl)];
  const double s = block_sum_fixed(v);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[block_linear(g, B)] = s;
}
__global__ void __launch_bounds__(256)
field_partials_kernel(const double* __restrict__ f, wf_grid g, double* __restrict__ partials,
                      int ich) {
  for_planes<true>(g.ni, ich, [&](Blk B) { field_partials_body(f, g, partials, B); });
}

// out = scale * sum(partials[0..n)), folded by one 1024-thread workgroup in
// a fixed order.
__device__ inline void fold_body(const double* __restrict__ partials, int64_t n, double scale,
                                 double* __restrict__ out) {
  __shared__ double s[1024];
  const int t = threadIdx.x;
  double acc = 0.0;
  // the same additions in the same order as one load per step, with eight
  // loads issued ahead (the fold was bound by one load latency per partial)
  int64_t q = t;
  for (; q + 7 * 1024 < n; q += 8 * 1024) {
    double a[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = partials[q + e * 1024];
#pragma unroll
    for (int e = 0; e < 8; ++e) acc += a[e];
  }
  for (; q < n; q += 1024) acc += partials[q];
  s[t] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) s[t] = s[t] + s[t + w];
    __syncthreads();
  }
  if (t == 0) *out = s[0] * scale;
}
__global__ void __launch_bounds__(1024)
fold_kernel(const double* __restrict__ partials, int64_t n, double scale, double* __restrict__ out) {
  fold_body(partials, n, scale, out);
}

__global__ void set_scalar_kernel(double* p, double v) { *p = v; }

// b - A x at one cell from x's centre value and its six neighbours, with
// fixed roundings (no contraction), so that the plane-walking and the
// unrolled defect kernels, single-domain or batched, agree bit for
// bit.
__device__ inline double defect_val(double b, const Coef<double>& c, double xc, double xw,
                                    double xe, double xs, double xn, double xb, double xt) {
#pragma clang fp contract(off)
  double nb = c.aw * xw;
  nb = __builtin_fma(c.ae, xe, nb);
  nb = __builtin_fma(c.as, xs, nb);
  nb = __builtin_fma(c.an, xn, nb);
  nb = __builtin_fma(c.ab, xb, nb);
  nb = __builtin_fma(c.at, xt, nb);
  return b - __builtin_fma(c.ap, xc, nb);
}

// r = b - A x (FP64; x == nullptr means x = 0), minus *mean if given, cast to
// FP32 into the level-0 rhs; rmax (nullable): each workgroup's max|b - A x|
// into rmax[workgroup].
__device__ inline void defect_body(const double* __restrict__ b, const double* __restrict__ x,
                                   const wf_grid& g, float* __restrict__ r32, const MgDev& d,
                                   double c0, double c1, double c2,
                                   const double* __restrict__ mean, double* __restrict__ rmax,
                                   Blk B) {
  const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
  double av = 0.0;
  if (k < g.nk && j < g.nj) {
    const int64_t q = at(g, i + g.npl, j + g.npl, k + g.npl);
    double r = b[q];
    if (x) {
      const Coef<double> c = mg_coef(d, i, j, k, c0, c1, c2);
      r = defect_val(r, c, x[q], x[q - g.si], x[q + g.si], x[q - g.sj], x[q + g.sj], x[q - 1],
                     x[q + 1]);
    }
    av = fabs(r);
    if (mean) r = r - *mean;
    r32[mg_at(d, i, j, k)] = (float)r;
  }
  if (rmax) {  // per-workgroup max into the partials (folded by max_fold_kernel)
    __shared__ double part[4];
    double m = wave_max(av);
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) part[t >> 6] = m;
    __syncthreads();
    if (t == 0) rmax[block_linear(g, B)] = nan_max(nan_max(part[0], part[1]), nan_max(part[2], part[3]));
  }
}
__global__ void __launch_bounds__(256)
defect_kernel(const double* __restrict__ b, const double* __restrict__ x, wf_grid g,
              float* __restrict__ r32, MgDev d, double c0, double c1, double c2,
              const double* __restrict__ mean, double* __restrict__ rmax, int ich) {
  for_planes<true>(g.ni, ich,
                   [&](Blk B) { defect_body(b, x, g, r32, d, c0, c1, c2, mean, rmax, B); });
}

// The same defect (x given, no rmax), DEFECT_U planes at a time: a thread
// issues every load of its (j, k) column for U planes -- x at planes i0-1 ..
// i0+U, b and the four in-plane neighbours of x at each plane -- before the
// first use, so U planes of HBM reads are in flight per wait (the
// plane-at-a-time walk waits on each plane's loads in turn).  Plane indices
// past the chunk are clamped onto its last plane (x: onto plane z1, real or
// the first ghost plane), so every load stays inside the field; their values
// are not used.  Same addresses and the same bits as defect_kernel.
constexpr int DEFECT_U = 4;
__global__ void __launch_bounds__(256)
defect_stream_kernel(const double* __restrict__ b, const double* __restrict__ x, wf_grid g,
                     float* __restrict__ r32, MgDev d, double c0, double c1, double c2,
                     const double* __restrict__ mean, int ich) {
  const Blk C = this_blk();
  const int k = C.x * 64 + threadIdx.x, j = C.y * 4 + threadIdx.y;
  if (k >= g.nk || j >= g.nj) return;  // no barriers below
  const int z0 = C.z * ich, z1 = min(z0 + ich, g.ni);
  const int64_t q0 = at(g, g.npl, j + g.npl, k + g.npl);  // real plane 0 of the column
  for (int i0 = z0; i0 < z1; i0 += DEFECT_U) {
    double xc[DEFECT_U + 2], bv[DEFECT_U], xs[DEFECT_U], xn[DEFECT_U], xb[DEFECT_U],
        xt[DEFECT_U];
#pragma unroll
    for (int u = 0; u < DEFECT_U + 2; ++u)
      xc[u] = x[q0 + (int64_t)min(i0 - 1 + u, z1) * g.si];
#pragma unroll
    for (int u = 0; u < DEFECT_U; ++u) {
      const int64_t q = q0 + (int64_t)min(i0 + u, z1 - 1) * g.si;
      bv[u] = b[q];
      xs[u] = x[q - g.sj];
      xn[u] = x[q + g.sj];
      xb[u] = x[q - 1];
      xt[u] = x[q + 1];
    }
#pragma unroll
    for (int u = 0; u < DEFECT_U; ++u) {
      const int i = i0 + u;
      if (i < z1) {
        const Coef<double> c = mg_coef(d, i, j, k, c0, c1, c2);
        double r = defect_val(bv[u], c, xc[u + 1], xc[u], xc[u + 2], xs[u], xn[u], xb[u], xt[u]);
        if (mean) r = r - *mean;
        r32[mg_at(d, i, j, k)] = (float)r;
      }
    }
  }
}

// The first cycle's update fused with the second cycle's defect: x = 0.0 + e
// on the real cells (update_kernel's assign) and r = b - A x from those same
// values, formed from e directly, so x is written and never read.  Across
// the domain faces the stencil takes e's ring in place of x's ghosts; both
// enter with a zero coefficient and are finite, so r has the bits of the
// separate update and defect.  Loads are grouped DEFECT_U planes at a time as
// in defect_stream_kernel (plane indices past the chunk clamped onto plane
// z1, which the (n+2)^3 level array holds).
__global__ void __launch_bounds__(256)
assign_defect_kernel(const double* __restrict__ b, double* __restrict__ x, wf_grid g,
                     const float* __restrict__ e, float* __restrict__ r32, MgDev d, double c0,
                     double c1, double c2, const double* __restrict__ mean, int ich) {
  const Blk C = this_blk();
  const int k = C.x * 64 + threadIdx.x, j = C.y * 4 + threadIdx.y;
  if (k >= g.nk || j >= g.nj) return;  // no barriers below
  const int z0 = C.z * ich, z1 = min(z0 + ich, g.ni);
  const int64_t q0 = at(g, g.npl, j + g.npl, k + g.npl);  // real plane 0 of the column
  const int64_t m0 = mg_at(d, 0, j, k);
  for (int i0 = z0; i0 < z1; i0 += DEFECT_U) {
    double xc[DEFECT_U + 2], bv[DEFECT_U], xs[DEFECT_U], xn[DEFECT_U], xb[DEFECT_U],
        xt[DEFECT_U];
#pragma unroll
    for (int u = 0; u < DEFECT_U + 2; ++u)
      xc[u] = 0.0 + (double)e[m0 + (int64_t)min(i0 - 1 + u, z1) * d.s0];
#pragma unroll
    for (int u = 0; u < DEFECT_U; ++u) {
      const int p = min(i0 + u, z1 - 1);
      const int64_t mq = m0 + (int64_t)p * d.s0;
      bv[u] = b[q0 + (int64_t)p * g.si];
      xs[u] = 0.0 + (double)e[mq - d.s1];
      xn[u] = 0.0 + (double)e[mq + d.s1];
      xb[u] = 0.0 + (double)e[mq - 1];
      xt[u] = 0.0 + (double)e[mq + 1];
    }
#pragma unroll
    for (int u = 0; u < DEFECT_U; ++u) {
      const int i = i0 + u;
      if (i < z1) {
        x[q0 + (int64_t)i * g.si] = xc[u + 1];
        const Coef<double> c = mg_coef(d, i, j, k, c0, c1, c2);
        double r = defect_val(bv[u], c, xc[u + 1], xc[u], xc[u + 2], xs[u], xn[u], xb[u], xt[u]);
        if (mean) r = r - *mean;
        r32[m0 + (int64_t)i * d.s0] = (float)r;
      }
    }
  }
}

// out = max(partials[0..n)), one 1024-thread workgroup, fixed order.
__global__ void __launch_bounds__(1024)
max_fold_kernel(const double* __restrict__ partials, int64_t n, double* __restrict__ out) {
  __shared__ double s[16];
  const int t = threadIdx.x;
  double m = 0.0;
  for (int64_t q = t; q < n; q += 1024) m = nan_max(m, partials[q]);
  m = wave_max(m);
  if ((t & 63) == 0) s[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double r = s[0];
    for (int w = 1; w < 16; ++w) r = nan_max(r, s[w]);
    *out = r;
  }
}

// x += e (FP32 correction widened to FP64) on the real cells.
__device__ inline void update_body(double* __restrict__ x, const wf_grid& g,
                                   const float* __restrict__ e, const MgDev& d, Blk B) {
  const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
  if (k >= g.nk || j >= g.nj) return;
  const int64_t q = at(g, i + g.npl, j + g.npl, k + g.npl);
  x[q] = x[q] + (double)e[mg_at(d, i, j, k)];
}
// The same on the solve's own field: assign != 0 takes x = 0 (the first
// cycle; 0.0 + e keeps the bits of adding e to a zeroed x, signed zeros
// included) without reading x; partials != nullptr also stores the block
// sums of the new x, in the block shape and order of field_partials_body.
__global__ void __launch_bounds__(256)
update_kernel(double* __restrict__ x, wf_grid g, const float* __restrict__ e, MgDev d, int ich,
              int assign, double* __restrict__ partials) {
  if (!partials) {
    for_planes(g.ni, ich, [&](Blk B) {
      const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
      if (k >= g.nk || j >= g.nj) return;
      const int64_t q = at(g, i + g.npl, j + g.npl, k + g.npl);
      x[q] = (assign ? 0.0 : x[q]) + (double)e[mg_at(d, i, j, k)];
    });
    return;
  }
  for_planes<true>(g.ni, ich, [&](Blk B) {
    const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
    double v = 0.0;
    if (k < g.nk && j < g.nj) {
      const int64_t q = at(g, i + g.npl, j + g.npl, k + g.npl);
      v = (assign ? 0.0 : x[q]) + (double)e[mg_at(d, i, j, k)];
      x[q] = v;
    }
    const double s = block_sum_fixed(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) partials[block_linear(g, B)] = s;
  });
}

__device__ inline void shift_body(double* __restrict__ x, const wf_grid& g,
                                  const double* __restrict__ mean, Blk B) {
  const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
  if (k >= g.nk || j >= g.nj) return;
  const int64_t q = at(g, i + g.npl, j + g.npl, k + g.npl);
  x[q] = x[q] - *mean;
}
__global__ void __launch_bounds__(256)
shift_kernel(double* __restrict__ x, wf_grid g, const double* __restrict__ mean, int ich) {
  for_planes(g.ni, ich, [&](Blk B) { shift_body(x, g, mean, B); });
}

// First ghost layer of the field: faces = mirror of the adjacent real cell,
// edges and corners = 0 (the ring state of the plain solve's copy-out, whose
// level ring keeps zero edges and corners).  Thread t < field_ring_cells(g).
__host__ __device__ inline int field_ring_cells(const wf_grid& g) {
  return 2 * (g.nj * g.nk + g.ni * g.nk + g.ni * g.nj) + 4 * (g.ni + g.nj + g.nk) + 8;
}
__device__ inline void field_ring_body(double* __restrict__ x, const wf_grid& g, int t) {
  const int n0 = g.ni, n1 = g.nj, n2 = g.nk, P = g.npl;
  const int fx = n1 * n2, fy = n0 * n2, fz = n0 * n1;
  int q = t;
  if (q < 2 * fx) {
    const int hi = q >= fx;
    q -= hi * fx;
    const int j = q / n2, k = q % n2, i = hi ? n0 - 1 : 0;
    const int64_t p = at(g, i + P, j + P, k + P);
    x[p + (hi ? g.si : -g.si)] = x[p];
    return;
  }
  q -= 2 * fx;
  if (q < 2 * fy) {
    const int hi = q >= fy;
    q -= hi * fy;
    const int i = q / n2, k = q % n2, j = hi ? n1 - 1 : 0;
    const int64_t p = at(g, i + P, j + P, k + P);
    x[p + (hi ? g.sj : -g.sj)] = x[p];
    return;
  }
  q -= 2 * fy;
  if (q < 2 * fz) {
    const int hi = q >= fz;
    q -= hi * fz;
    const int i = q / n1, j = q % n1, k = hi ? n2 - 1 : 0;
    const int64_t p = at(g, i + P, j + P, k + P);
    x[p + (hi ? 1 : -1)] = x[p];
    return;
  }
  q -= 2 * fz;
  // edges along k, j, i (ring coordinate -1 or n on the other two axes), corners
  if (q < 4 * n2) {
    const int e = q / n2;
    x[at(g, (e & 1 ? n0 : -1) + P, (e & 2 ? n1 : -1) + P, q % n2 + P)] = 0.0;
    return;
  }
  q -= 4 * n2;
  if (q < 4 * n1) {
    const int e = q / n1;
    x[at(g, (e & 1 ? n0 : -1) + P, q % n1 + P, (e & 2 ? n2 : -1) + P)] = 0.0;
    return;
  }
  q -= 4 * n1;
  if (q < 4 * n0) {
    const int e = q / n0;
    x[at(g, q % n0 + P, (e & 1 ? n1 : -1) + P, (e & 2 ? n2 : -1) + P)] = 0.0;
    return;
  }
  q -= 4 * n0;
  if (q < 8) x[at(g, (q & 1 ? n0 : -1) + P, (q & 2 ? n1 : -1) + P, (q & 4 ? n2 : -1) + P)] = 0.0;
}
__global__ void __launch_bounds__(256) field_ring_kernel(double* __restrict__ x, wf_grid g) {
  field_ring_body(x, g, blockIdx.x * blockDim.x + threadIdx.x);
}

// max|rhs - A phi| over the real cells of one level (FP64 or FP32).
template <class T>
__global__ void __launch_bounds__(256)
level_resmax_kernel(const T* __restrict__ phi, const T* __restrict__ rhs, MgDev d, T c0, T c1,
                    T c2, double* __restrict__ out) {
  const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
  double av = 0.0;
  if (k < d.n2 && j < d.n1) {
    const int64_t x = mg_at(d, i, j, k);
    const Coef<T> c = mg_coef(d, i, j, k, c0, c1, c2);
    av = fabs((double)(rhs[x] - (nb_sum(phi, x, d, c) + c.ap * phi[x])));
  }
  __shared__ double part[4];
  double m = wave_max(av);
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  if ((t & 63) == 0) part[t >> 6] = m;
  __syncthreads();
  if (t == 0) atomic_max_nonneg(out, nan_max(nan_max(part[0], part[1]), nan_max(part[2], part[3])));
}

// ---------------------------------------------------------------------------
// Batched (ensemble) mixed solve: the single-domain kernels' bodies over
// every active member, level 0 of member m at level pool + loff[0].
// ---------------------------------------------------------------------------
__device__ inline MgDev level0_dev(const MemDev& M) {
  MgDev d;
  d.n0 = M.ln[0][0];
  d.n1 = M.ln[0][1];
  d.n2 = M.ln[0][2];
  d.s1 = d.n2 + 2;
  d.s0 = (int64_t)(d.n1 + 2) * d.s1;
  return d;
}

template <bool SYNC, class F>
__device__ inline void mx_items(const Plan& P, const BView& V, const int* active, F&& f) {
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    if (active && !active[m]) continue;
    f(m, it - P.start[m], V.mem[m]);
    if (SYNC) __syncthreads();
  }
}

__global__ void __launch_bounds__(256)
batch_defect_kernel(BView V, Plan P, const double* su, const double* pp, float* rhs0,
                    const int* active) {
  mx_items<false>(P, V, active, [&](int, int q, const MemDev& M) {
    const wf_grid& g = M.g;
    defect_body(su + M.off, pp ? pp + M.off : nullptr, g, rhs0 + M.loff[0], level0_dev(M),
                1.0 / (M.dx * M.dx), 1.0 / (M.dy * M.dy), 1.0 / (M.dz * M.dz), nullptr,
                nullptr, cells_blk(q, g.nj, g.nk));
  });
}

// One plane of assign_defect_kernel's pass (x = 0.0 + e on the real cells,
// r = b - A x from e), for the batched solve's one-plane work items.
__device__ inline void assign_defect_body(const double* __restrict__ b, double* __restrict__ x,
                                          const wf_grid& g, const float* __restrict__ e,
                                          float* __restrict__ r32, const MgDev& d, double c0,
                                          double c1, double c2, Blk B) {
  const int k = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y, i = B.z;
  if (k >= g.nk || j >= g.nj) return;
  const int64_t m = mg_at(d, i, j, k);
  const auto xe = [&](int64_t o) { return 0.0 + (double)e[m + o]; };
  const double xc = xe(0);
  x[at(g, i + g.npl, j + g.npl, k + g.npl)] = xc;
  const Coef<double> c = mg_coef(d, i, j, k, c0, c1, c2);
  r32[m] = (float)defect_val(b[at(g, i + g.npl, j + g.npl, k + g.npl)], c, xc, xe(-d.s0),
                             xe(d.s0), xe(-d.s1), xe(d.s1), xe(-1), xe(1));
}

__global__ void __launch_bounds__(256)
batch_assign_defect_kernel(BView V, Plan P, const double* su, double* pp, const float* phi0,
                           float* rhs0, const int* active) {
  mx_items<false>(P, V, active, [&](int, int q, const MemDev& M) {
    const wf_grid& g = M.g;
    assign_defect_body(su + M.off, pp + M.off, g, phi0 + M.loff[0], rhs0 + M.loff[0],
                       level0_dev(M), 1.0 / (M.dx * M.dx), 1.0 / (M.dy * M.dy),
                       1.0 / (M.dz * M.dz), cells_blk(q, g.nj, g.nk));
  });
}

__global__ void __launch_bounds__(256)
batch_update_kernel(BView V, Plan P, double* pp, const float* phi0, const int* active) {
  mx_items<false>(P, V, active, [&](int, int q, const MemDev& M) {
    update_body(pp + M.off, M.g, phi0 + M.loff[0], level0_dev(M), cells_blk(q, M.g.nj, M.g.nk));
  });
}

__global__ void __launch_bounds__(256)
batch_partials_kernel(BView V, Plan P, const double* f, double* part, const int* active) {
  mx_items<true>(P, V, active, [&](int, int q, const MemDev& M) {
    field_partials_body(f + M.off, M.g, part + M.sum0, cells_blk(q, M.g.nj, M.g.nk));
  });
}

__global__ void __launch_bounds__(1024)
batch_sum_fold_kernel(BView V, const double* part, double* out, const int* active) {
  const int m = blockIdx.x;
  if (active && !active[m]) return;
  const MemDev& M = V.mem[m];
  const wf_grid& g = M.g;
  fold_body(part + M.sum0, cells_blocks(g.ni, g.nj, g.nk), 1.0 / ((double)g.ni * g.nj * g.nk),
            out + m);
}

__global__ void __launch_bounds__(256)
batch_shift_kernel(BView V, Plan P, double* pp, const double* mean, const int* active) {
  mx_items<false>(P, V, active, [&](int m, int q, const MemDev& M) {
    shift_body(pp + M.off, M.g, mean + m, cells_blk(q, M.g.nj, M.g.nk));
  });
}

__global__ void __launch_bounds__(256)
batch_field_ring_kernel(BView V, double* pp, const int* active) {
  const int m = blockIdx.x;
  if (active && !active[m]) return;
  const MemDev& M = V.mem[m];
  const wf_grid& g = M.g;
  const int nring = field_ring_cells(g);
  for (int t = threadIdx.x; t < nring; t += blockDim.x) field_ring_body(pp + M.off, g, t);
}

__global__ void __launch_bounds__(256) batch_fill_kernel(double* p, int64_t n, double v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}
__global__ void __launch_bounds__(256) batch_fillf_kernel(float* p, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = 0.0f;
}

// wf_mg_solve_mixed over every active member with a fixed cycle count (the
// plain iteration x <- x + M(b - A x), no tolerance test: no host sync).
int batch_mg_mixed(const Batch& B, float* const* phi, float* const* rhs, float* const* tmp,
                   const float* cop, const double* su, double* pp, int ncycles, int npre,
                   int npost, int ncoarse, int flags, const int* active, hipStream_t s) {
  const BView V = B.view();
  const Plan& P = B.plan[PK_CELLS];
  const int fb = (int)std::min<int64_t>((B.field_elems + 255) / 256, 8192);
  hipLaunchKernelGGL(batch_fill_kernel, dim3(fb), dim3(256), 0, s, pp, B.field_elems, 0.0);
  WF_LAUNCHED();
  bool defect_formed = false;  // the fused first update formed this cycle's defect
  for (int c = 0; c < ncycles; ++c) {
    if (!defect_formed) {
      hipLaunchKernelGGL(batch_defect_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, su,
                         c == 0 ? (const double*)nullptr : (const double*)pp, rhs[0], active);
      WF_LAUNCHED();
    }
    defect_formed = false;
    const int lb = (int)std::min<int64_t>((B.level_elems[0] + 255) / 256, 8192);
    hipLaunchKernelGGL(batch_fillf_kernel, dim3(lb), dim3(256), 0, s, phi[0], B.level_elems[0]);
    WF_LAUNCHED();
    const int r = batch_vcycle(B, phi, rhs, tmp, cop, npre, npost, ncoarse, active, s);
    if (r) return r;
    if (c == 0 && ncycles > 1) {
      // pp = 0 + e and the second cycle's defect in one pass
      hipLaunchKernelGGL(batch_assign_defect_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, su,
                         pp, (const float*)phi[0], rhs[0], active);
      defect_formed = true;
    } else {
      hipLaunchKernelGGL(batch_update_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, pp,
                         (const float*)phi[0], active);
    }
    WF_LAUNCHED();
  }
  if (flags & MIX_RESULT_MEAN) {
    hipLaunchKernelGGL(batch_partials_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P,
                       (const double*)pp, B.sum_part, active);
    WF_LAUNCHED();
    hipLaunchKernelGGL(batch_sum_fold_kernel, dim3(B.nmem), dim3(1024), 0, s, V,
                       (const double*)B.sum_part, B.scal, active);
    WF_LAUNCHED();
    hipLaunchKernelGGL(batch_shift_kernel, dim3(P.grid), dim3(64, 4), 0, s, V, P, pp,
                       (const double*)B.scal, active);
    WF_LAUNCHED();
  }
  hipLaunchKernelGGL(batch_field_ring_kernel, dim3(B.nmem), dim3(256), 0, s, V, pp, active);
  WF_LAUNCHED();
  return 0;
}

}  // namespace wf

using namespace wf;

namespace {

MgDev dev_of(const wf_mg_level& L) {
  MgDev d;
  d.n0 = L.n0;
  d.n1 = L.n1;
  d.n2 = L.n2;
  d.s1 = L.n2 + 2;
  d.s0 = (int64_t)(L.n1 + 2) * d.s1;
  return d;
}

int field_sum(const double* f, const wf_grid& g, double scale, double* work, double* out,
              hipStream_t s) {
  double* partials = work + WORK_HEAD;
  hipLaunchKernelGGL(field_partials_kernel, field_cgrid(g), dim3(64, 4), 0, s, f, g, partials,
                     field_ich(g));
  WF_LAUNCHED();
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(1024), 0, s, partials, field_blocks(g), scale,
                     out);
  WF_LAUNCHED();
  return 0;
}

}  // namespace

extern "C" {

int wf_mg_mixed_workspace_bytes(wf_grid g, size_t* bytes) {
  if (!bytes || !mg_field_grid_ok(g)) return WF_EINVAL;
  *bytes = (size_t)(WORK_HEAD + field_blocks(g)) * sizeof(double);
  return 0;
}

int wf_mg_residual_max(const wf_mg_level* L, double* out_dev, void* stream) {
  if (!mg_level_ok(L) || !out_dev) return WF_EINVAL;
  const hipStream_t s = (hipStream_t)stream;
  const MgDev d = dev_of(*L);
  hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, s, out_dev, 0.0);
  WF_LAUNCHED();
  if (L->prec == 8)
    hipLaunchKernelGGL(level_resmax_kernel<double>, mg_cell_grid(d.n0, d.n1, d.n2), dim3(64, 4),
                       0, s, (const double*)L->phi, (const double*)L->rhs, d,
                       1.0 / (L->h0 * L->h0), 1.0 / (L->h1 * L->h1), 1.0 / (L->h2 * L->h2),
                       out_dev);
  else
    hipLaunchKernelGGL(level_resmax_kernel<float>, mg_cell_grid(d.n0, d.n1, d.n2), dim3(64, 4),
                       0, s, (const float*)L->phi, (const float*)L->rhs, d,
                       (float)(1.0 / (L->h0 * L->h0)), (float)(1.0 / (L->h1 * L->h1)),
                       (float)(1.0 / (L->h2 * L->h2)), out_dev);
  WF_LAUNCHED();
  return 0;
}

int wf_mg_solve_mixed(const wf_mg_level* levels, int nlev, const double* su, double* pp,
                      wf_grid g, int n_pre, int n_post, int n_coarse, int max_cycles,
                      double rtol, int flags, double* work, size_t work_bytes, int* cycles_out,
                      double* res_out, void* stream) {
  if (!mg_hier_ok(levels, nlev) || !mg_field_grid_ok(g) || levels[0].prec != 4) return WF_EINVAL;
  if (levels[0].n0 != g.ni || levels[0].n1 != g.nj || levels[0].n2 != g.nk) return WF_EINVAL;
  if (max_cycles < 0 || n_pre < 0 || n_post < 0 || n_coarse < 0 || !work) return WF_EINVAL;
  size_t need = 0;
  wf_mg_mixed_workspace_bytes(g, &need);
  if (work_bytes < need) return WF_EINVAL;
  const hipStream_t s = (hipStream_t)stream;
  const wf_mg_level& L0 = levels[0];
  const MgDev d0 = dev_of(L0);
  const double c0 = 1.0 / (L0.h0 * L0.h0), c1 = 1.0 / (L0.h1 * L0.h1),
               c2 = 1.0 / (L0.h2 * L0.h2);
  const double ncell = (double)g.ni * g.nj * g.nk;
  const bool tol_mode = rtol > 0.0;
  double* mean = work + 0;
  double* rmax = work + 1;
  double* xmean = work + 2;

  // x = 0 is implicit: the first defect reads no x and the first update
  // assigns x = e, so the real cells are never read before being written;
  // the ring is rewritten at the end.  Only a solve that runs no cycle
  // zeroes x explicitly.
  int r = 0;
  if (flags & MIX_DEFECT_MEAN) {
    r = field_sum(su, g, 1.0 / ncell, work, mean, s);
    if (r) return r;
  }
  double bmax = 0.0, res = 0.0;
  int cycles = 0;
  // with a fixed cycle count the last update also forms the block sums of
  // x that the result mean needs (the same sums field_partials would form)
  bool summed = false;
  const double* dmean = (flags & MIX_DEFECT_MEAN) ? mean : nullptr;
  bool defect_formed = false;  // the fused first update formed this cycle's defect
  for (int c = 0; c <= max_cycles; ++c) {
    const bool last_check = c == max_cycles;
    if (last_check && !tol_mode) break;
    const bool rhs_ready =
        defect_formed ||
        (c == 0 && (flags & MIX_RHS_READY) && !tol_mode && !(flags & MIX_DEFECT_MEAN));
    defect_formed = false;
    if (!rhs_ready) {
      if (c > 0 && !tol_mode)
        hipLaunchKernelGGL(defect_stream_kernel, field_cgrid(g), dim3(64, 4), 0, s, su, pp, g,
                           (float*)L0.rhs, d0, c0, c1, c2, dmean, field_ich(g));
      else
        hipLaunchKernelGGL(defect_kernel, field_cgrid(g), dim3(64, 4), 0, s, su,
                           c == 0 ? nullptr : pp, g, (float*)L0.rhs, d0, c0, c1, c2, dmean,
                           tol_mode ? work + WORK_HEAD : nullptr, field_ich(g));
      WF_LAUNCHED();
    }
    if (tol_mode) {
      hipLaunchKernelGGL(max_fold_kernel, dim3(1), dim3(1024), 0, s, work + WORK_HEAD,
                         field_blocks(g), rmax);
      WF_LAUNCHED();
      double h = 0.0;
      WF_TRY(hipMemcpyAsync(&h, rmax, sizeof(double), hipMemcpyDeviceToHost, s));
      WF_TRY(hipStreamSynchronize(s));
      if (c == 0) bmax = h;
      res = bmax > 0.0 ? h / bmax : 0.0;
      if (res <= rtol || last_check) break;
    }
    // the V-cycle starts from e = 0: the first sweep of level 0 takes it
    // without reading phi where its smoother allows, else phi is zeroed
    const bool zs = mg_relax_zero_start_ok(L0, nlev > 1 ? n_pre : n_coarse);
    if (!zs) {
      r = mg_zero<float>(L0.phi, mg_level_elems(L0), s);
      if (r) return r;
    }
    r = mg_vcycle_levels(levels, nlev, n_pre, n_post, n_coarse, s, zs);
    if (r) return r;
    summed = (flags & MIX_RESULT_MEAN) && !tol_mode && c == max_cycles - 1;
    if (c == 0 && !tol_mode && c + 1 < max_cycles) {
      // x = e and the next cycle's defect in one pass
      hipLaunchKernelGGL(assign_defect_kernel, field_cgrid(g), dim3(64, 4), 0, s, su, pp, g,
                         (const float*)L0.phi, (float*)L0.rhs, d0, c0, c1, c2, dmean,
                         field_ich(g));
      defect_formed = true;
    } else {
      hipLaunchKernelGGL(update_kernel, field_cgrid(g), dim3(64, 4), 0, s, pp, g,
                         (const float*)L0.phi, d0, field_ich(g), c == 0 ? 1 : 0,
                         summed ? work + WORK_HEAD : nullptr);
    }
    WF_LAUNCHED();
    ++cycles;
  }
  if (cycles == 0) {
    r = mg_zero<double>(pp, (int64_t)(g.ni + 2 * g.npl) * g.si, s);
    if (r) return r;
  }
  if (flags & MIX_RESULT_MEAN) {
    if (summed) {
      hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(1024), 0, s, work + WORK_HEAD,
                         field_blocks(g), 1.0 / ncell, xmean);
      WF_LAUNCHED();
    } else {
      r = field_sum(pp, g, 1.0 / ncell, work, xmean, s);
      if (r) return r;
    }
    if (!(flags & MIX_MEAN_DEFERRED)) {
      hipLaunchKernelGGL(shift_kernel, field_cgrid(g), dim3(64, 4), 0, s, pp, g, xmean,
                         field_ich(g));
      WF_LAUNCHED();
    }
  }
  hipLaunchKernelGGL(field_ring_kernel, dim3(div_up(field_ring_cells(g), 256)), dim3(256), 0, s,
                     pp, g);
  WF_LAUNCHED();
  if (cycles_out) *cycles_out = cycles;
  if (res_out) *res_out = res;
  return 0;
}

int wf_field_mean(const double* f, wf_grid g, double* work, size_t work_bytes, double* out_dev,
                  void* stream) {
  size_t need = 0;
  if (!mg_field_grid_ok(g) || wf_mg_mixed_workspace_bytes(g, &need) || work_bytes < need)
    return WF_EINVAL;
  return field_sum(f, g, 1.0 / ((double)g.ni * g.nj * g.nk), work, out_dev, (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation form of wf_mg_solve_mixed: its workspace (four scalars and
// one partial sum per workgroup of a field pass) lives in `scratch`, a static
// device array of WF_STATIC_SCRATCH_BYTES that every GPU the library runs on
// holds until the process ends.  Host code never passes the array's name to
// a kernel: the name is a host-side stand-in, so the device address of each
// GPU's copy is looked up once per device with hipGetSymbolAddress and kept.
// Every kernel writes its partials and scalars before any kernel reads them,
// so nothing in the scratch needs clearing between calls.  Calls that overlap
// on one GPU (two streams, two threads) share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[WF_STATIC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[SCRATCH_MAX_DEVICES];  // device address of scratch, per device

int scratch_of_current_device(double** out) {
  int dev = 0;
  WF_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return WF_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    WF_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return WF_EINVAL;
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace

extern "C" int wf_mg_solve_mixed_static(const wf_mg_level* levels, int nlev, const double* su,
                                        double* pp, wf_grid g, int n_pre, int n_post,
                                        int n_coarse, int max_cycles, double rtol, int flags,
                                        int* cycles_out, double* res_out, void* stream) {
  // the deferred gauge mean would be left inside the library's scratch, where
  // the caller cannot reach it and the next call overwrites it
  if (flags & MIX_MEAN_DEFERRED) return WF_EINVAL;
  size_t need = 0;
  if (wf_mg_mixed_workspace_bytes(g, &need)) return WF_EINVAL;
  if (need > WF_STATIC_SCRATCH_BYTES) return WF_ESCRATCH;
  double* work = nullptr;
  const int r = scratch_of_current_device(&work);
  if (r) return r;
  return wf_mg_solve_mixed(levels, nlev, su, pp, g, n_pre, n_post, n_coarse, max_cycles, rtol,
                           flags, work, need, cycles_out, res_out, stream);
}
