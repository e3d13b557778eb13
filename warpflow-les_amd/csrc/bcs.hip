// Ghost-layer fills of the cavity boundary set (the reference's
// warpflow/bcs.py): W/E/B no-slip walls, S/N slip planes, T moving lid.
//
// The reference issues one 2-D launch per side in the order W, E, S, N, B,
// T.  A side operation reads and writes only the line of cells along its
// axis through one perpendicular position (a, b), so the two sides of one
// axis can share a launch: a thread runs the low-side then the high-side
// operations of its (a, b) line, which is exactly the reference's per-cell
// order.  The x fill has a launch of its own, because the y and z fills
// read x ghosts it writes in other planes; the y and z fills of one i-plane
// touch only that plane, so one workgroup per plane runs its y lines, a
// barrier, then its z lines (bc_yz_kernel): two launches per fill.
#include <algorithm>

#include "batch.h"
#include "wf_common.h"

namespace wf {

struct BcOp {
  int32_t kind;  // 0: ghost = val; 1: ghost = val * source
  int32_t high;
  int32_t g0, s0;
  int32_t nlay;
  double val;
};

struct BcOps {
  BcOp op[4];
  int32_t n;
};

// Whether every source cell of the op list lies outside the ghost cells
// written by the same or an earlier op (and every op has at most 3 layers):
// then all sources can be loaded before the first store and the stores keep
// their order, which leaves every cell with the value of the op-by-op walk.
// Uniform over a launch (it depends on the ops only).
__host__ __device__ inline bool bc_ops_hoistable(const BcOps& ops) {
  if (ops.n > 4) return false;
  for (int q = 0; q < ops.n; ++q) {
    const BcOp& o = ops.op[q];
    if (o.nlay > 3) return false;
    if (o.kind == 0 || o.nlay <= 0) continue;
    const int slo = o.high ? o.s0 - o.nlay + 1 : o.s0, shi = o.high ? o.s0 : o.s0 + o.nlay - 1;
    for (int p = 0; p <= q; ++p) {
      const BcOp& w = ops.op[p];
      if (w.nlay <= 0) continue;
      const int glo = w.high ? w.g0 : w.g0 - w.nlay + 1, ghi = w.high ? w.g0 + w.nlay - 1 : w.g0;
      if (slo <= ghi && glo <= shi) return false;
    }
  }
  return true;
}

// The ops of one (a, b) line along `axis`, low side then high side.  When
// the ops allow it (bc_ops_hoistable) a thread issues all its loads before
// its first store, so a line costs one memory round trip instead of one per
// layer and side (a load after a store to the same array cannot be moved
// ahead of it by the compiler).
__device__ inline void bc_line(double* __restrict__ f, const wf_grid& g, int axis, int a, int b,
                               const BcOps& ops) {
  int64_t base, s;
  if (axis == 0) {
    base = at(g, 0, a, b);
    s = g.si;
  } else if (axis == 1) {
    base = at(g, a, 0, b);
    s = g.sj;
  } else {
    base = at(g, a, b, 0);
    s = 1;
  }
  if (bc_ops_hoistable(ops)) {
    // every op field, destination and source value is in registers before
    // the first store (a kernel-argument read placed after a store would
    // wait for that store to complete)
    double v[4][3];
    int64_t dst[4][3];
    bool on[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const BcOp o = ops.op[q];
      const bool used = q < ops.n;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        on[q][m] = used && m < o.nlay;
        dst[q][m] = base + (int64_t)(o.high ? o.g0 + m : o.g0 - m) * s;
        // unconditional load (cell 0 of the line when no source is read), so
        // no branch separates the loads
        const bool src = on[q][m] && o.kind != 0;
        const double x = f[base + (int64_t)(src ? (o.high ? o.s0 - m : o.s0 + m) : 0) * s];
        v[q][m] = src ? o.val * x : o.val;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int m = 0; m < 3; ++m)
        if (on[q][m]) f[dst[q][m]] = v[q][m];
    }
    return;
  }
  for (int q = 0; q < ops.n; ++q) {
    const BcOp o = ops.op[q];
    for (int m = 0; m < o.nlay; ++m) {
      const int gi = o.high ? o.g0 + m : o.g0 - m;
      if (o.kind == 0) {
        f[base + gi * s] = o.val;
      } else {
        const int si = o.high ? o.s0 - m : o.s0 + m;
        f[base + gi * s] = o.val * f[base + si * s];
      }
    }
  }
}

__global__ void __launch_bounds__(256)
bc_axis_kernel(double* __restrict__ f, wf_grid g, int axis, int a0, int na, int b0, int nb,
               BcOps ops) {
  const int bb = blockIdx.x * blockDim.x + threadIdx.x;
  const int aa = blockIdx.y * blockDim.y + threadIdx.y;
  if (aa >= na || bb >= nb) return;
  bc_line(f, g, axis, a0 + aa, b0 + bb, ops);
}

// The same axis launch for the u, v and w fills at once (blockIdx.z picks
// the field): the three fields are independent, so each keeps its own
// per-cell order of operations.
struct Bc3 {
  double* f[3];
  int32_t a0[3], na[3], b0[3], nb[3];
  BcOps ops[3];
  int32_t edges;  // refill: only the lines in another axis's written band
  int32_t yl[3], yh[3], zl[3], zh[3];  // field z's y and z bands (see in_band)
};

// Refill of fields filled since their last change: an op rewrites the values
// it wrote before unless one of its sources is a cell another axis's ops
// wrote, and a line along one axis has such a source only if it lies in
// another axis's written band: array index <= lo (the highest index the low
// side's ops write) or >= hi (the lowest the high side's write), ghosts and
// wall faces.  The other lines are skipped.
__device__ inline bool in_band(int x, int lo, int hi) { return x <= lo || x >= hi; }

__global__ void __launch_bounds__(256) bc_axis3_kernel(Bc3 A, wf_grid g, int axis) {
  const int z = blockIdx.z;
  const int bb = blockIdx.x * blockDim.x + threadIdx.x;
  const int aa = blockIdx.y * blockDim.y + threadIdx.y;
  if (aa >= A.na[z] || bb >= A.nb[z]) return;
  const int a = A.a0[z] + aa, b = A.b0[z] + bb;  // (j, k) lines of the x ops
  if (A.edges && !in_band(a, A.yl[z], A.yh[z]) && !in_band(b, A.zl[z], A.zh[z])) return;
  bc_line(A.f[z], g, axis, a, b, A.ops[z]);
}

// The y and z ops of a fill, one workgroup per array plane i (blockIdx.x =
// i - i_base; blockIdx.y picks the field of a multi-field fill): the y ops of
// every (i, k) line of the plane, a barrier, then the z ops of every (i, j)
// line.  The workgroup has a thread per line up to 1024 lines, so each phase
// is one round trip to memory.  A y or z op reads and writes only cells of
// its own plane, and the z ops read the y ghosts the y ops have just
// written, as in the reference's side order (bcs.py:201-311); the x ops,
// which read other planes, run in the launch before.  Each cell sees the
// op-by-op order of the three axis launches, so the bits are theirs.
// The ghost cells (k indices) a field's z ops write, when there are at most
// ZLANES of them and no op reads a cell the list writes (n = 0 otherwise),
// each with the op that writes it last: ghost g[c] = val[c] * (cell src[c])
// or, for src[c] < 0, val[c].
constexpr int ZLANES = 8;
struct BcCells {
  int32_t g[ZLANES], src[ZLANES];
  double val[ZLANES];
  int32_t n;
};
struct BcYZ {
  double* f[3];
  int32_t ay0[3], nay[3], by0[3], nby[3];  // y ops: planes [ay0, ay0+nay) x k lines
  int32_t az0[3], naz[3], bz0[3], nbz[3];  // z ops: planes [az0, az0+naz) x j lines
  BcOps oy[3], oz[3];
  BcCells zc[3];
  int32_t ibase;
  int32_t edges;  // refill (see in_band)
  int32_t xl[3], xh[3], yl[3], yh[3], zl[3], zh[3];  // field z's written bands
};

// The z ops of ZLANES-lane groups, a lane per ghost cell of an (i, j) line:
// the lane stores the value of the last op that writes its cell, which is
// what the cell holds after the op-by-op walk (no op reads a cell the list
// writes, so the earlier writes are never read).  A wave then touches the
// cache lines of 64 / ZLANES lines per access, where a thread per line
// touches 64 (the ghosts and sources of a line share its first and last
// lines).  A thread issues the loads of ZU groups before their stores.
__device__ inline void bc_zcells(double* __restrict__ f, const wf_grid& g, int i, int b0, int nb,
                                 const BcCells& zc) {
  constexpr int ZU = 4;
  const int total = nb * ZLANES, nt = blockDim.x;
  for (int t0 = threadIdx.x; t0 < total; t0 += ZU * nt) {
    double v[ZU];
    int64_t dst[ZU];
    bool on[ZU];
#pragma unroll
    for (int u = 0; u < ZU; ++u) {
      const int t = t0 + u * nt, c = t % ZLANES;
      on[u] = t < total && c < zc.n;
      const int cc = on[u] ? c : 0;
      const int64_t base = at(g, i, b0 + (on[u] ? t / ZLANES : 0), 0);
      dst[u] = base + zc.g[cc];
      const int sc = zc.src[cc];
      // unconditional load (cell 0 of the line for a constant ghost)
      const double x = f[base + (sc >= 0 ? sc : 0)];
      v[u] = sc >= 0 ? zc.val[cc] * x : zc.val[cc];
    }
#pragma unroll
    for (int u = 0; u < ZU; ++u)
      if (on[u]) f[dst[u]] = v[u];
  }
}

__global__ void __launch_bounds__(1024) bc_yz_kernel(BcYZ A, wf_grid g) {
  const int z = blockIdx.y;
  const int i = A.ibase + (int)blockIdx.x;
  // refill: a plane outside the x ghost band keeps only its band lines
  const bool part = A.edges && !in_band(i, A.xl[z], A.xh[z]);
  if (i >= A.ay0[z] && i < A.ay0[z] + A.nay[z])
    for (int b = threadIdx.x; b < A.nby[z]; b += blockDim.x)
      if (!part || in_band(A.by0[z] + b, A.zl[z], A.zh[z]))
        bc_line(A.f[z], g, 1, i, A.by0[z] + b, A.oy[z]);
  __syncthreads();
  if (i >= A.az0[z] && i < A.az0[z] + A.naz[z]) {
    if (part) {
      for (int b = threadIdx.x; b < A.nbz[z]; b += blockDim.x)
        if (in_band(A.bz0[z] + b, A.yl[z], A.yh[z])) bc_line(A.f[z], g, 2, i, A.bz0[z] + b, A.oz[z]);
    } else if (A.zc[z].n > 0)
      bc_zcells(A.f[z], g, i, A.bz0[z], A.nbz[z], A.zc[z]);
    else
      for (int b = threadIdx.x; b < A.nbz[z]; b += blockDim.x)
        bc_line(A.f[z], g, 2, i, A.bz0[z] + b, A.oz[z]);
  }
}

}  // namespace wf

using namespace wf;

static int axis_extent(const wf_grid& g, int axis) {
  return (axis == 0 ? g.ni : (axis == 1 ? g.nj : g.nk)) + 2 * g.npl;
}

// Check every index an op list touches against the array before launching.
static bool ops_ok(const wf_grid& g, int axis, int a0, int na, int b0, int nb,
                   const BcOps& ops) {
  if (g.ni <= 0 || g.nj <= 0 || g.nk <= 0 || axis < 0 || axis > 2) return false;
  if (g.sj < (int64_t)(g.nk + 2 * g.npl) || g.si < g.sj * (int64_t)(g.nj + 2 * g.npl))
    return false;
  const int pa = axis == 0 ? 1 : 0, pb = axis == 2 ? 1 : 2;
  if (na < 0 || nb < 0 || a0 < 0 || b0 < 0) return false;
  if (na > 0 && a0 + na > axis_extent(g, pa)) return false;
  if (nb > 0 && b0 + nb > axis_extent(g, pb)) return false;
  const int n = axis_extent(g, axis);
  for (int q = 0; q < ops.n; ++q) {
    const BcOp& o = ops.op[q];
    if (o.nlay < 0) return false;
    for (int m = 0; m < o.nlay; ++m) {
      const int gi = o.high ? o.g0 + m : o.g0 - m;
      if (gi < 0 || gi >= n) return false;
      if (o.kind == 1) {
        const int si = o.high ? o.s0 - m : o.s0 + m;
        if (si < 0 || si >= n) return false;
      }
    }
  }
  return true;
}

static int launch_axis(double* f, const wf_grid& g, int axis, int a0, int na, int b0, int nb,
                       const BcOps& ops, hipStream_t s) {
  if (!ops_ok(g, axis, a0, na, b0, nb, ops)) return WF_EINVAL;
  if (na == 0 || nb == 0 || ops.n == 0) return 0;
  dim3 block(64, 4), grid(div_up(nb, 64), div_up(na, 4));
  hipLaunchKernelGGL(bc_axis_kernel, grid, block, 0, s, f, g, axis, a0, na, b0, nb, ops);
  WF_LAUNCHED();
  return 0;
}

__host__ __device__ inline BcOp C_(int high, int g0, double val, int nlay) {
  return BcOp{0, high, g0, 0, nlay, val};
}
__host__ __device__ inline BcOp L_(int high, int g0, int s0, double sign, int nlay) {
  return BcOp{1, high, g0, s0, nlay, sign};
}
__host__ __device__ inline BcOps ops2(BcOp a, BcOp b) {
  BcOps o{};
  o.op[0] = a;
  o.op[1] = b;
  o.n = 2;
  return o;
}

// Launch range (a0, na) x (b0, nb) and op list of the `axis` launch of
// BoundaryKernels.apply_u / _v / _w / _vis / _pbound / _p_mirror
// (which = 0..5, bcs.py:201-311) on an N-ghost field of ni x nj x nk real
// cells.  The reference uses one n for every axis; here each range uses its
// own axis' count, identical for a cubic grid.  False for an unknown `which`.
__host__ __device__ inline bool bc_axis_ops(int which, int axis, int N, int ni, int nj, int nk,
                                            double param, int& a0, int& na, int& b0, int& nb,
                                            BcOps& o) {
  a0 = b0 = N - 1;
  const int n3[3] = {ni, nj, nk};
  const int pa = axis == 0 ? nj : ni, pb = axis == 2 ? nj : nk;  // perpendicular counts
  switch (which) {
    case 0:  // apply_u, bcs.py:201-223; param = lid velocity
      if (axis == 0) {
        na = nj + 2, nb = nk + 2;
        o = ops2(C_(0, N - 1, 0.0, 3), C_(1, N + ni - 1, 0.0, 3));
      } else if (axis == 1) {
        na = ni + 1, nb = nk + 2;
        o = ops2(L_(0, N - 1, N, 1.0, 3), L_(1, N + nj, N + nj - 1, 1.0, 3));
      } else {
        na = ni + 1, nb = nj + 2;
        o = ops2(L_(0, N - 1, N, -1.0, 3), C_(1, N + nk, param, 3));
      }
      return true;
    case 1:  // apply_v, bcs.py:225-244
      if (axis == 0) {
        na = nj + 1, nb = nk + 2;
        o = ops2(L_(0, N - 1, N, -1.0, 3), L_(1, N + ni, N + ni - 1, -1.0, 3));
      } else if (axis == 1) {
        na = ni + 2, nb = nk + 2;
        o = ops2(C_(0, N - 1, 0.0, 3), C_(1, N + nj - 1, 0.0, 3));
      } else {
        na = ni + 2, nb = nj + 1;
        o = ops2(L_(0, N - 1, N, -1.0, 3), L_(1, N + nk, N + nk - 1, 1.0, 3));
      }
      return true;
    case 2:  // apply_w, bcs.py:246-265
      if (axis == 0) {
        na = nj + 2, nb = nk + 1;
        o = ops2(L_(0, N - 1, N, -1.0, 3), L_(1, N + ni, N + ni - 1, -1.0, 3));
      } else if (axis == 1) {
        na = ni + 2, nb = nk + 1;
        o = ops2(L_(0, N - 1, N, 1.0, 3), L_(1, N + nj, N + nj - 1, 1.0, 3));
      } else {
        na = ni + 2, nb = nj + 2;
        o = ops2(C_(0, N - 1, 0.0, 3), C_(1, N + nk - 1, 0.0, 3));
      }
      return true;
    case 3:  // apply_vis, bcs.py:267-293; param = 1/Re
      na = pa + 2, nb = pb + 2;
      if (axis == 0) {
        o = BcOps{};
        o.op[0] = C_(0, N, param, 1);
        o.op[1] = L_(0, N - 1, N, 1.0, 1);
        o.op[2] = C_(1, N + ni - 1, param, 1);
        o.op[3] = L_(1, N + ni, N + ni - 1, 1.0, 1);
        o.n = 4;
      } else if (axis == 1) {
        o = ops2(L_(0, N - 1, N, 1.0, 1), L_(1, N + nj, N + nj - 1, 1.0, 1));
      } else {
        o = BcOps{};
        o.op[0] = C_(0, N, param, 1);
        o.op[1] = L_(0, N - 1, N, 1.0, 1);
        o.op[2] = L_(1, N + nk, N + nk - 1, 1.0, 1);
        o.n = 3;
      }
      return true;
    case 4:  // apply_pbound, bcs.py:295-302
      na = pa + 3, nb = pb + 3;
      o = ops2(L_(0, N - 1, N, 1.0, 1), L_(1, N + n3[axis], N + n3[axis] - 1, 1.0, 1));
      return true;
    case 5: {  // apply_p_mirror, bcs.py:304-311; param = layers
      const int nl = (int)param;
      a0 = b0 = 0;
      na = pa + 2 * N, nb = pb + 2 * N;
      o = ops2(L_(0, N - 1, N, 1.0, nl), L_(1, N + n3[axis], N + n3[axis] - 1, 1.0, nl));
      return true;
    }
    default:
      return false;
  }
}

namespace wf {

// Line blocks of one axis launch per member: 64 x 4 (b, a) blocks over the
// widest range any fill uses ((perpendicular count + 6) per side).
__host__ __device__ inline void bc_extent(int axis, int ni, int nj, int nk, int& namax,
                                          int& nbmax) {
  namax = (axis == 0 ? nj : ni) + 6;
  nbmax = (axis == 2 ? nj : nk) + 6;
}
int bc_blocks(int axis, int ni, int nj, int nk) {
  int na, nb;
  bc_extent(axis, ni, nj, nk, na, nb);
  return div_up(nb, 64) * div_up(na, 4);
}

// Fields of a multi-field batched fill: blockIdx.y picks field z.
struct BcFields {
  double* f[3];
  int which[3];
};

__global__ void __launch_bounds__(256)
batch_bc_kernel(BView V, Plan P, BcFields F, int axis, const int* active) {
  double* __restrict__ f = F.f[blockIdx.y];
  const int which = F.which[blockIdx.y];
  for (int it = blockIdx.x; it < P.total; it += gridDim.x) {
    const int m = member_of(P.start, V.nmem, it);
    if (active && !active[m]) continue;
    const int q = it - P.start[m];
    const MemDev& M = V.mem[m];
    const wf_grid g = M.g;
    const double param = which == 0 ? M.ulid : (which == 3 ? M.rey : (double)g.npl);
    int a0, na, b0, nb;
    BcOps o{};
    bc_axis_ops(which, axis, g.npl, g.ni, g.nj, g.nk, param, a0, na, b0, nb, o);
    int namax, nbmax;
    bc_extent(axis, g.ni, g.nj, g.nk, namax, nbmax);
    const int nx = (nbmax + 63) / 64;
    const int bb = (q % nx) * 64 + threadIdx.x, aa = (q / nx) * 4 + threadIdx.y;
    if (aa < na && bb < nb) bc_line(f + M.off, g, axis, a0 + aa, b0 + bb, o);
  }
}

// The three axis launches of the fills which[z] of fields f[z] (z < nf,
// independent fields) over every (active) member: each launch covers the nf
// fields, one per grid row, so the axes keep their order within each field.
int batch_bc_fill_n(const Batch& B, double* const* f, const int* which, int nf,
                    const int* active, hipStream_t s) {
  if (nf < 1 || nf > 3) return WF_EINVAL;
  BcFields F{};
  for (int z = 0; z < nf; ++z) {
    if (!f[z] || which[z] < 0 || which[z] > 5) return WF_EINVAL;
    F.f[z] = f[z];
    F.which[z] = which[z];
  }
  for (int axis = 0; axis < 3; ++axis) {
    const Plan& P = B.plan[PK_BC0 + axis];
    hipLaunchKernelGGL(batch_bc_kernel, dim3(P.grid, nf), dim3(64, 4), 0, s, B.view(), P, F,
                       axis, active);
    WF_LAUNCHED();
  }
  return 0;
}

int batch_bc_fill(const Batch& B, double* f, int which, const int* active, hipStream_t s) {
  return batch_bc_fill_n(B, &f, &which, 1, active, s);
}

}  // namespace wf

extern "C" {

int wf_bc_lin(double* fld, wf_grid g, int axis, int high, int g0, int s0, double sign,
              int a0, int na, int b0, int nb, int nlay, void* stream) {
  BcOps o{};
  o.op[0] = L_(high != 0, g0, s0, sign, nlay);
  o.n = 1;
  return launch_axis(fld, g, axis, a0, na, b0, nb, o, (hipStream_t)stream);
}

int wf_bc_const(double* fld, wf_grid g, int axis, int high, int g0, double val, int a0,
                int na, int b0, int nb, int nlay, void* stream) {
  BcOps o{};
  o.op[0] = C_(high != 0, g0, val, nlay);
  o.n = 1;
  return launch_axis(fld, g, axis, a0, na, b0, nb, o, (hipStream_t)stream);
}

// The y-then-z launch of fills `which[z]` of fields f[z] (z < nf): ranges and
// ops of the y and z axis launches, checked against the array, one
// workgroup per plane any of them touches.
// The distinct ghost cells of a z op list, for bc_zcells (n = 0: the list
// is not hoistable, some op reads a cell some op writes -- the lanes of a
// line load and store concurrently -- or it writes more than ZLANES cells).
static BcCells z_cells(const BcOps& ops) {
  BcCells c{};
  if (!bc_ops_hoistable(ops)) return c;
  for (int q = 0; q < ops.n; ++q) {
    const BcOp& o = ops.op[q];
    if (o.kind == 0) continue;
    for (int m = 0; m < o.nlay; ++m) {
      const int si = o.high ? o.s0 - m : o.s0 + m;
      for (int p = 0; p < ops.n; ++p)
        for (int u = 0; u < ops.op[p].nlay; ++u)
          if ((ops.op[p].high ? ops.op[p].g0 + u : ops.op[p].g0 - u) == si) return c;
    }
  }
  int n = 0;
  for (int q = 0; q < ops.n; ++q)
    for (int m = 0; m < ops.op[q].nlay; ++m) {
      const BcOp& o = ops.op[q];
      const int gi = o.high ? o.g0 + m : o.g0 - m;
      int u = 0;
      while (u < n && c.g[u] != gi) ++u;
      if (u == n) {
        if (n == ZLANES) return BcCells{};
        c.g[n++] = gi;
      }
      // a later op overwrites an earlier one's entry: the last writer stays
      c.src[u] = o.kind != 0 ? (o.high ? o.s0 - m : o.s0 + m) : -1;
      c.val[u] = o.val;
    }
  c.n = n;
  return c;
}

// The band an op list writes along its axis: lo = the highest index a low-side
// op writes, hi = the lowest a high-side op writes.
static void op_band(const BcOps& o, int& lo, int& hi) {
  lo = -1;
  hi = 1 << 30;
  for (int q = 0; q < o.n; ++q) {
    if (o.op[q].nlay < 1) continue;
    if (o.op[q].high)
      hi = std::min(hi, (int)o.op[q].g0);
    else
      lo = std::max(lo, (int)o.op[q].g0);
  }
}

// The written bands of fill `which` (param) along each axis.  *self_dep is
// set when an op reads a cell of its own axis's band (tiny grids, where one
// side's sources reach the other side's ghosts): a line can then change on a
// refill without lying in another axis's band, and the refill runs in full.
static bool fill_bands(int which, double param, const wf_grid& g, int lo[3], int hi[3],
                       bool* self_dep = nullptr) {
  for (int axis = 0; axis < 3; ++axis) {
    int a0, na, b0, nb;
    BcOps o{};
    if (!bc_axis_ops(which, axis, g.npl, g.ni, g.nj, g.nk, param, a0, na, b0, nb, o))
      return false;
    op_band(o, lo[axis], hi[axis]);
    if (self_dep)
      for (int q = 0; q < o.n; ++q)
        if (o.op[q].kind == 1)
          for (int m = 0; m < o.op[q].nlay; ++m) {
            const int si = o.op[q].high ? o.op[q].s0 - m : o.op[q].s0 + m;
            if (si <= lo[axis] || si >= hi[axis]) *self_dep = true;
          }
  }
  return true;
}

static int launch_yz(double* const* f, const int* which, const double* param, int nf,
                     const wf_grid& g, hipStream_t s, bool edges = false) {
  BcYZ A{};
  A.edges = edges;
  for (int z = 0; z < nf; ++z) {
    int lo[3], hi[3];
    if (!fill_bands(which[z], param[z], g, lo, hi)) return WF_EINVAL;
    A.xl[z] = lo[0], A.xh[z] = hi[0], A.yl[z] = lo[1], A.yh[z] = hi[1];
    A.zl[z] = lo[2], A.zh[z] = hi[2];
  }
  int lo = 1 << 30, hi = -1, lines = 64;
  for (int z = 0; z < nf; ++z) {
    A.f[z] = f[z];
    int a0, na, b0, nb;
    if (!bc_axis_ops(which[z], 1, g.npl, g.ni, g.nj, g.nk, param[z], a0, na, b0, nb, A.oy[z]) ||
        !ops_ok(g, 1, a0, na, b0, nb, A.oy[z]))
      return WF_EINVAL;
    A.ay0[z] = a0, A.nay[z] = na, A.by0[z] = b0, A.nby[z] = nb;
    if (na > 0 && nb > 0 && A.oy[z].n > 0) {
      lo = std::min(lo, a0), hi = std::max(hi, a0 + na);
      lines = std::max(lines, nb);
    } else {
      A.nay[z] = 0;
    }
    if (!bc_axis_ops(which[z], 2, g.npl, g.ni, g.nj, g.nk, param[z], a0, na, b0, nb, A.oz[z]) ||
        !ops_ok(g, 2, a0, na, b0, nb, A.oz[z]))
      return WF_EINVAL;
    A.az0[z] = a0, A.naz[z] = na, A.bz0[z] = b0, A.nbz[z] = nb;
    A.zc[z] = z_cells(A.oz[z]);
    if (na > 0 && nb > 0 && A.oz[z].n > 0) {
      lo = std::min(lo, a0), hi = std::max(hi, a0 + na);
      lines = std::max(lines, nb);
    } else {
      A.naz[z] = 0;
    }
  }
  if (hi <= lo) return 0;
  A.ibase = lo;
  const int threads = std::min(1024, (lines + 63) / 64 * 64);
  hipLaunchKernelGGL(bc_yz_kernel, dim3(hi - lo, nf), dim3(threads), 0, s, A, g);
  WF_LAUNCHED();
  return 0;
}

// Per-axis restatement of BoundaryKernels.apply_u/v/w/vis/pbound/p_mirror
// (bcs.py:201-311): the x launch of bc_axis_ops, then its y and z ops in one
// launch of a workgroup per plane (bc_yz_kernel).
int wf_bc_fill(double* f, wf_grid g, int which, double param, void* stream) {
  const hipStream_t s = (hipStream_t)stream;
  if (g.npl < 1 || which < 0 || which > 5) return WF_EINVAL;
  int a0, na, b0, nb;
  BcOps o{};
  if (!bc_axis_ops(which, 0, g.npl, g.ni, g.nj, g.nk, param, a0, na, b0, nb, o))
    return WF_EINVAL;
  WF_TRY((hipError_t)launch_axis(f, g, 0, a0, na, b0, nb, o, s));
  return launch_yz(&f, &which, &param, 1, g, s);
}

// apply_u(u), apply_v(v), apply_w(w) (bcs.py:201-265) on three fields of
// one geometry: one x launch covering the three fields, then one y-then-z
// launch covering them.
static int fill_uvw(double* u, double* v, double* w, const wf_grid& g, double ulid, hipStream_t s,
                    bool edges) {
  if (edges) {  // the band rule needs sources outside every band
    bool dep = false;
    int lo[3], hi[3];
    for (int which = 0; which < 3; ++which)
      if (!fill_bands(which, which == 0 ? ulid : 0.0, g, lo, hi, &dep)) return WF_EINVAL;
    if (dep) edges = false;
  }
  if (g.npl < 1 || !v || !w) return WF_EINVAL;
  // the fields filled, in the reference's order (u may be absent)
  double* fs[3];
  int which[3], nf = 0;
  if (u) {
    fs[nf] = u;
    which[nf++] = 0;
  }
  fs[nf] = v;
  which[nf++] = 1;
  fs[nf] = w;
  which[nf++] = 2;
  {  // the x launch
    const int axis = 0;
    Bc3 A{};
    A.edges = edges;
    for (int z = 0; z < nf; ++z) {
      int lo[3], hi[3];
      if (!fill_bands(which[z], which[z] == 0 ? ulid : 0.0, g, lo, hi)) return WF_EINVAL;
      A.yl[z] = lo[1], A.yh[z] = hi[1], A.zl[z] = lo[2], A.zh[z] = hi[2];
    }
    int namax = 0, nbmax = 0;
    for (int z = 0; z < nf; ++z) {
      int a0, na, b0, nb;
      A.f[z] = fs[z];
      if (!bc_axis_ops(which[z], axis, g.npl, g.ni, g.nj, g.nk, which[z] == 0 ? ulid : 0.0, a0,
                       na, b0, nb, A.ops[z]) ||
          !ops_ok(g, axis, a0, na, b0, nb, A.ops[z]))
        return WF_EINVAL;
      A.a0[z] = a0;
      A.na[z] = na;
      A.b0[z] = b0;
      A.nb[z] = nb;
      namax = std::max(namax, na);
      nbmax = std::max(nbmax, nb);
    }
    if (namax > 0 && nbmax > 0) {
      hipLaunchKernelGGL(bc_axis3_kernel, dim3(div_up(nbmax, 64), div_up(namax, 4), nf),
                         dim3(64, 4), 0, s, A, g, axis);
      WF_LAUNCHED();
    }
  }
  double param[3];
  for (int z = 0; z < nf; ++z) param[z] = which[z] == 0 ? ulid : 0.0;
  return launch_yz(fs, which, param, nf, g, s, edges);
}

int wf_bc_fill_uvw(double* u, double* v, double* w, wf_grid g, double ulid, void* stream) {
  return fill_uvw(u, v, w, g, ulid, (hipStream_t)stream, false);
}

int wf_bc_refill_uvw(double* u, double* v, double* w, wf_grid g, double ulid, void* stream) {
  return fill_uvw(u, v, w, g, ulid, (hipStream_t)stream, true);
}

}  // extern "C"
