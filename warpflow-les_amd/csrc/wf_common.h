// Shared definitions for the warpflow gfx950 kernels.
//
// Layout: every field is a C-ordered [i][j][k] array with k unit-stride;
// `wf_grid` carries the real-cell counts, the ghost depth and the element
// strides of j and i, so padded pitches and batched members use the same
// kernels.  Full-array index (i, j, k) -> i*si + j*sj + k.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/warpflow_hip.h"

#define WF_TRY(expr)                                   \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return (int)e_;              \
  } while (0)

#define WF_LAUNCHED() WF_TRY(hipPeekAtLastError())

namespace wf {

__host__ __device__ inline int64_t at(const wf_grid& g, int i, int j, int k) {
  return (int64_t)i * g.si + (int64_t)j * g.sj + (int64_t)k;
}

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Launch geometry for a 3-D index box of extent (d0, d1, d2): 64 lanes along
// k (one wavefront per k-row), 4 j-rows per workgroup, one i-plane per
// grid.z.  Every kernel bounds-checks against the box.
struct Box3 {
  int lo0, lo1, lo2;
  int d0, d1, d2;
};

constexpr int BX = 64, BY = 4;

inline dim3 box_grid(const Box3& b) {
  return dim3(div_up(b.d2, BX), div_up(b.d1, BY), b.d0);
}
inline dim3 box_block() { return dim3(BX, BY, 1); }

// Block coordinates of a workgroup: blockIdx for a single-domain launch, or
// the decoded work item of a batched (ensemble) launch, so that both run the
// same kernel body.
struct Blk {
  int x, y, z;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b, linear
// in x-fastest order, runs on XCD b % 8), each XCD with its own L2.  This
// bijection hands XCD x the contiguous run of logical blocks
// [x*q + min(x, r), ...), so the blocks that share halo rows and planes
// share an L2.
__host__ __device__ inline int xcd_linear(int b, int nb) {
  const int q = nb / 8, r = nb % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}
__device__ inline Blk this_blk() {
  const int gx = gridDim.x, gy = gridDim.y;
  const int nb = gx * gy * (int)gridDim.z;
  const int L = xcd_linear(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
  return Blk{L % gx, (L / gx) % gy, L / (gx * gy)};
}

// Elementwise kernels give each workgroup a chunk of `ich` i-planes of its
// (k, j) tile: the block decode above is paid once per chunk, and the
// i-neighbours a plane reads were just read by the same CU.  The chunk is
// sized to keep >= 8192 workgroups in flight.
inline int plane_chunk(int d0, int d1, int d2) {
  const int64_t tiles = (int64_t)div_up(d2, BX) * div_up(d1, BY);
  const int64_t c = tiles * d0 / 8192;
  return (int)(c < 1 ? 1 : (c > 8 ? 8 : c));
}
inline dim3 chunk_grid(int d0, int d1, int d2, int ich) {
  return dim3(div_up(d2, BX), div_up(d1, BY), div_up(d0, ich));
}
// f(B) for every plane B.z of this workgroup's chunk; SYNC: a barrier after
// each plane, for bodies that reduce through LDS.
template <bool SYNC = false, class F>
__device__ inline void for_planes(int d0, int ich, F&& f) {
  const Blk C = this_blk();
  const int z0 = C.z * ich, z1 = min(z0 + ich, d0);
  for (int z = z0; z < z1; ++z) {
    f(Blk{C.x, C.y, z});
    if (SYNC) __syncthreads();
  }
}

__device__ inline bool box_thread(const Box3& b, Blk B, int& i, int& j, int& k) {
  k = B.x * BX + threadIdx.x;
  j = B.y * BY + threadIdx.y;
  i = B.z;
  if (k >= b.d2 || j >= b.d1 || i >= b.d0) return false;
  i += b.lo0;
  j += b.lo1;
  k += b.lo2;
  return true;
}
__device__ inline bool box_thread(const Box3& b, int& i, int& j, int& k) {
  return box_thread(b, this_blk(), i, j, k);
}

// Non-negative doubles order like their bit patterns, so max over them is an
// integer max; order-independent and therefore deterministic.
__device__ inline void atomic_max_nonneg(double* p, double v) {
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}
__device__ inline void atomic_min_nonneg(double* p, double v) {
  atomicMin((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

// Wavefront (64-lane) reductions via cross-lane shuffles.  max/min propagate
// NaN (a blown-up field must not vanish from rmax or the dt limits).
__device__ inline double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ inline double nan_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = nan_min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Tuning table (api.hip): launch-shape choices per (knob, shape) set by the
// autotuner; 0 = the built-in choice.  They change launch shapes only, never
// results.
int tune_get(int knob, int n0, int n1, int n2);

// Fused conv4th + diff4th of component comp over box b (convdiff.hip); ich =
// i-planes per workgroup, 0 picks it from the box size.  coo_save != nullptr:
// the base value is the component itself (coo unused), written to coo_save
// on the box.
int convdiff_launch(int comp, const double* u, const double* v, const double* w,
                    const double* coo, const double* vis, double* cstar, const wf_grid& g,
                    const Box3& b, double dx, double dy, double dz, double dta, int ich,
                    hipStream_t s, double* coo_save = nullptr, int nch = 0, int cstep = 1);
// The first RK stage's conv+diff of u, v and w in one pass over the union of
// their operator boxes (convdiff.hip), saving u, v, w into uoo, voo, woo.
// The second RK stage's conv+diff of v and w in one pass; v's z-term reads
// w_pre where w's fill writes (convdiff.hip).
int convdiff_vw_launch(const double* u, const double* v, const double* w, const double* voo,
                       const double* woo, const double* vis, const double* wpre, double* vs,
                       double* ws, const wf_grid& g, double dx, double dy, double dz,
                       double dta, int ich, hipStream_t s);
int save_cross_ghosts_launch(const double* w, double* wpre, const wf_grid& g, hipStream_t s);
int convdiff3_save_launch(const double* u, const double* v, const double* w,
                          const double* vis, double* us, double* vs, double* ws, double* uoo,
                          double* voo, double* woo, const wf_grid& g, double dx, double dy,
                          double dz, double dta, int ich, hipStream_t s);
// The second RK stage's conv+diff of u, v and w in one pass after the three
// fills: v's z-term reads w_pre as convdiff_vw_launch's does; u's planes 0 and
// ni-2 are not written (convdiff_u_edges_launch writes them before the fills
// of v and w).
int convdiff3_launch(const double* u, const double* v, const double* w, const double* uoo,
                     const double* voo, const double* woo, const double* vis,
                     const double* wpre, double* us, double* vs, double* ws, const wf_grid& g,
                     double dx, double dy, double dz, double dta, int ich, hipStream_t s);
int convdiff_u_edges_launch(const double* u, const double* v, const double* w,
                            const double* uoo, const double* vis, double* us, const wf_grid& g,
                            double dx, double dy, double dz, double dta, hipStream_t s);

// Raw buffer access: a scalar descriptor holds the base (and a byte range),
// each lane supplies a 32-bit byte offset, so a load or store needs no 64-bit
// vector address.  A lane offset at or past the range (BUF_DEAD, with ranges
// kept below 2^30) reads zero and drops its store.  Descriptor word 3 =
// 0x00020000: untyped 32-bit data format, the value the gfx9 family (gfx950
// included) expects for raw buffer instructions.
constexpr uint32_t BUF_DEAD = 0xc0000000u;
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
template <class T>
__device__ inline T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t boff, int soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)boff, soff, 0));
}
// AUX: the cache-policy operand; BUF_NT (the gfx940-family "nt" bit) marks a
// store non-temporal, so streamed outputs do not displace lines in the L2 that
// other workgroups are about to read.
constexpr int BUF_NT = 2;
template <class T, int AUX = 0>
__device__ inline void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t boff, int soff, T v) {
  if constexpr (sizeof(T) == 8) {
    using U2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, (int)boff, soff, AUX);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)boff, soff,
                                          AUX);
  }
}

}  // namespace wf
