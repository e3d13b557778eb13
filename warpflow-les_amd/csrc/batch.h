// Ensemble (batched) execution: many independent cavities of different
// sizes advanced by one launch per operator.
//
// Every field of every member lives in one pool per field name (member m's
// (n0+6)(n1+6)(n2+6) array at element offset off); every multigrid level l
// in one pool per level array (member m's (n_l+2)^3 array at loff[l]).  A
// batched launch walks a list of work items -- (member, local block) pairs
// numbered member after member -- with a grid-stride loop: the grid is at
// most the number of workgroups the device holds at once (a persistent
// launch), and each item is decoded into the block coordinates the
// single-domain kernel body expects, so both paths run the same arithmetic.
// A per-member `active` mask (nullable) lets converged members skip the
// projection iterations of the others inside one captured graph.
#pragma once
#include "wf_common.h"

namespace wf {

constexpr int BATCH_MAX_LEVELS = 8;

// Device view of one member, derived from wf_member by wf_batch_create.
struct MemDev {
  wf_grid g;  // field geometry, 3 ghost layers
  int64_t off;
  double dx, dy, dz, rey, ulid, noise, dt;
  double ls;    // WALE length scale cw * delta
  double rsum;  // 1/dx^2 + 1/dy^2 + 1/dz^2 (the checkdt viscous limit)
  uint32_t seed;
  int nlev;
  int ln[BATCH_MAX_LEVELS][3];
  int64_t loff[BATCH_MAX_LEVELS];
  int64_t coff;  // coarse-operator offset (elements) in the operator pool
  int red0;      // first reduction-partial slot of this member
  int sum0;      // first sum-partial slot of this member
};

// A work-item list: start[m] .. start[m+1]-1 are member m's items.
struct Plan {
  const int* start;
  int total;
  int grid;  // workgroups launched (<= total)
};

struct BView {
  const MemDev* mem;
  int nmem;
};

// Largest m with start[m] <= w (start[0] = 0, start[nmem] = total > w).
__device__ inline int member_of(const int* __restrict__ start, int nmem, int w) {
  int lo = 0, hi = nmem;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Block coordinates of local block q of a 64 x 4 (k, j) x plane launch over
// a box of extent (d0, d1, d2): the box_grid order (x fastest).
__host__ __device__ inline int cells_blocks(int d0, int d1, int d2) {
  return ((d2 + 63) / 64) * ((d1 + 3) / 4) * d0;
}
__device__ inline Blk cells_blk(int q, int d1, int d2) {
  const int nx = (d2 + 63) / 64, ny = (d1 + 3) / 4;
  return Blk{q % nx, (q / nx) % ny, q / (nx * ny)};
}

// Reduction items: 64 x 4 (k, j) tiles x i-chunks of RED_ICH planes.
constexpr int BATCH_RED_ICH = 8;
__host__ __device__ inline int red_blocks(int d0, int d1, int d2) {
  return ((d2 + 63) / 64) * ((d1 + 3) / 4) * ((d0 + BATCH_RED_ICH - 1) / BATCH_RED_ICH);
}

// Work-item counts per member for each kind of batched launch.
enum PlanKind : int {
  PK_CELLS = 0,     // 64x4 blocks over the real cells
  PK_RING,          // 64x4 blocks over real cells + 1 ring layer
  PK_RED,           // reduction tiles over the real cells
  PK_CD0,           // convdiff tiles, components u, v, w
  PK_CD1,
  PK_CD2,
  PK_BC0,           // ghost-fill lines of the x, y, z axis launches
  PK_BC1,
  PK_BC2,
  PK_WALE,          // WALE tiles (8 x 32 (j, k) x BATCH_CD_ICH planes) over the real cells;
                    // also the three-component conv+diff passes
  PK_CDE,           // u's edge planes 0 and ni-2: 8 x 32 (j, k) tiles x 2
  PK_XG,            // w's cross-ghost cells of each real plane, 256 per item
  PK_MEMBER,        // one item per member
  PK_LCELLS0,       // 64x4 blocks over level-l real cells, l = 0 .. 7
  PK_SWEEP0 = PK_LCELLS0 + BATCH_MAX_LEVELS,  // fused-sweep tiles of level l
  PK_COUNT = PK_SWEEP0 + BATCH_MAX_LEVELS
};

// Host side of a batch (defined in batch.hip).
struct Batch {
  int nmem = 0, nlev = 0;
  MemDev* mem_dev = nullptr;        // nmem descriptors (device)
  MemDev* mem_host = nullptr;       // host copy
  Plan plan[PK_COUNT];
  int* plan_mem[PK_COUNT] = {};     // device start arrays
  double* red_part = nullptr;       // 4 * nred reduction partials (device)
  int nred = 0;
  double* sum_part = nullptr;       // nsum sum partials (device)
  int nsum = 0;
  double* scal = nullptr;           // 4 * nmem scalars (device): means, maxima
  int64_t field_elems = 0;
  int64_t level_elems[BATCH_MAX_LEVELS] = {};
  int64_t cop_elems = 0;
  int max_coarse_cells = 0, max_lds_cells[BATCH_MAX_LEVELS] = {};
  int lds = 0;  // dynamic LDS bytes available to one workgroup
  BView view() const { return BView{mem_dev, nmem}; }
};

// Convdiff tiles (8 x 32 (j, k) tiles x i-chunks): shared with convdiff.hip.
constexpr int BATCH_CD_ICH = 16;
__host__ __device__ inline int cd_tiles(int d0, int d1, int d2) {
  return ((d2 + 31) / 32) * ((d1 + 7) / 8) * ((d0 + BATCH_CD_ICH - 1) / BATCH_CD_ICH);
}
// Operator box of component c over n0 x n1 x n2 cells (solver.py:99-110).
__host__ __device__ inline void op_dims(int c, int n0, int n1, int n2, int& d0, int& d1,
                                        int& d2) {
  d0 = n0 - (c == 0);
  d1 = n1 - (c == 1);
  d2 = n2 - (c == 2);
}
// Fused-sweep tiles of a level (the single-domain tile choice).
__host__ __device__ inline int sweep_tj(int n2) { return n2 > 32 ? 8 : 16; }
__host__ __device__ inline int sweep_tk(int n2) { return n2 > 32 ? 64 : 32; }
constexpr int BATCH_SWEEP_ICH = 8;
__host__ __device__ inline int sweep_tiles(int n0, int n1, int n2) {
  return ((n2 + sweep_tk(n2) - 1) / sweep_tk(n2)) * ((n1 + sweep_tj(n2) - 1) / sweep_tj(n2)) *
         ((n0 + BATCH_SWEEP_ICH - 1) / BATCH_SWEEP_ICH);
}

// Batched launchers, one per source file (argument checks done by the C ABI).
int batch_noise(const Batch& B, double* u, double* v, double* w, hipStream_t s);
int batch_fill_vis(const Batch& B, double* vis, hipStream_t s);
int batch_wale(const Batch& B, const double* u, const double* v, const double* w, double* vis,
               hipStream_t s);
int batch_checkdt(const Batch& B, const double* u, const double* v, const double* w,
                  const double* vis, double* out4, hipStream_t s);
int batch_calvel3(const Batch& B, const double* us, const double* vs, const double* ws,
                  const double* p, double* u, double* v, double* w, double alpha,
                  const int* active, hipStream_t s);
int batch_calmas(const Batch& B, const double* u, const double* v, const double* w, double* su,
                 double* rmax, const int* active, hipStream_t s);
int batch_calvel3_calmas(const Batch& B, const double* us, const double* vs, const double* ws,
                         const double* p, double* u, double* v, double* w, double* su,
                         double* rmax, double alpha, const int* active, hipStream_t s);
int batch_add_ring(const Batch& B, double* p, const double* pp, const int* active,
                   hipStream_t s);
int batch_convdiff(const Batch& B, int comp, const double* u, const double* v,
                   const double* w, const double* coo, const double* vis, double* cstar,
                   double alpha, hipStream_t s, double* coo_save = nullptr);
// The three-component conv+diff passes of the two RK stages (convdiff3_save_
// kernel, convdiff3_kernel), u's edge-plane pass and w's cross-ghost copy, for
// every member (convdiff.hip).
int batch_convdiff3(const Batch& B, const double* u, const double* v, const double* w,
                    const double* vis, const double* wpre, double* us, double* vs, double* ws,
                    double* uoo, double* voo, double* woo, double alpha, hipStream_t s);
int batch_convdiff_u_edges(const Batch& B, const double* u, const double* v, const double* w,
                           const double* uoo, const double* vis, double* us, double alpha,
                           hipStream_t s);
int batch_save_cross_ghosts(const Batch& B, const double* w, double* wpre, hipStream_t s);
int batch_bc_fill(const Batch& B, double* f, int which, const int* active, hipStream_t s);
int batch_bc_fill_n(const Batch& B, double* const* f, const int* which, int nf,
                    const int* active, hipStream_t s);
int batch_vcycle(const Batch& B, float* const* phi, float* const* rhs, float* const* tmp,
                 const float* cop, int npre, int npost, int ncoarse, const int* active,
                 hipStream_t s);
int bc_blocks(int axis, int ni, int nj, int nk);
int batch_mg_init(int* lds_out);
int batch_mg_mixed(const Batch& B, float* const* phi, float* const* rhs, float* const* tmp,
                   const float* cop, const double* su, double* pp, int ncycles, int npre,
                   int npost, int ncoarse, int flags, const int* active, hipStream_t s);

}  // namespace wf
