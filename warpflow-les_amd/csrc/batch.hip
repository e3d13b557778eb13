// Ensemble batches: member descriptors, pool layout, work-item plans and the
// C ABI of the batched operators (the launchers live beside the kernels they
// share their bodies with: ops.hip, convdiff.hip, bcs.hip, mg.hip, mixed.hip).
//
// Everything a batched step touches is allocated here, at creation, so that a
// step can be captured into a HIP graph and replayed without allocation.
#include <algorithm>
#include <vector>

#include "batch.h"

namespace wf {

// Projection-loop bookkeeping per member (solver.py:178-212): a member still
// active at outer iteration `it` records it, then leaves the loop when
// rmax < eps and it > 1; rmax > 10 marks it diverged (status 1) and stops it.
// A NaN rmax counts as diverged too: the reference's `rmax > 10.0` test lets a
// NaN pass, but in a batch such a member would hold every other member in
// host-synchronised outer iterations up to the limit.
__global__ void batch_mask_kernel(int nmem, const double* __restrict__ rmax, double eps, int it,
                                  int* __restrict__ active, int* __restrict__ iters,
                                  int* __restrict__ status) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmem || !active[m]) return;
  iters[m] = it;
  const double r = rmax[m];
  if (r < eps && it > 1) {
    active[m] = 0;
  } else if (!(r <= 10.0)) {
    active[m] = 0;
    status[m] = 1;
  }
}

__global__ void batch_set_int_kernel(int* __restrict__ p, int n, int v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

}  // namespace wf

using namespace wf;

namespace {

constexpr int64_t ALIGN = 32;  // elements: member arrays start on 256-byte (FP64) boundaries

int64_t align_up(int64_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

int member_count(int kind, const MemDev& M) {
  const int n0 = M.g.ni, n1 = M.g.nj, n2 = M.g.nk;
  switch (kind) {
    case PK_CELLS:
      return cells_blocks(n0, n1, n2);
    case PK_RING:
      return cells_blocks(n0 + 2, n1 + 2, n2 + 2);
    case PK_RED:
      return red_blocks(n0, n1, n2);
    case PK_CD0:
    case PK_CD1:
    case PK_CD2: {
      int d0, d1, d2;
      op_dims(kind - PK_CD0, n0, n1, n2, d0, d1, d2);
      return cd_tiles(d0, d1, d2);
    }
    case PK_BC0:
    case PK_BC1:
    case PK_BC2:
      return bc_blocks(kind - PK_BC0, n0, n1, n2);
    case PK_WALE:
      return cd_tiles(n0, n1, n2);
    case PK_CDE:
      return ((n2 + 31) / 32) * ((n1 + 7) / 8) * (n0 - 1 > 1 ? 2 : (n0 - 1 > 0 ? 1 : 0));
    case PK_XG:
      return ((2 * (n2 + 3) + 4 * n1 + 255) / 256) * n0;
    case PK_MEMBER:
      return 1;
    default:
      break;
  }
  if (kind >= PK_LCELLS0 && kind < PK_LCELLS0 + BATCH_MAX_LEVELS) {
    const int l = kind - PK_LCELLS0;
    if (l >= M.nlev) return 0;
    return cells_blocks(M.ln[l][0], M.ln[l][1], M.ln[l][2]);
  }
  if (kind >= PK_SWEEP0 && kind < PK_SWEEP0 + BATCH_MAX_LEVELS) {
    const int l = kind - PK_SWEEP0;
    if (l >= M.nlev) return 0;
    return sweep_tiles(M.ln[l][0], M.ln[l][1], M.ln[l][2]);
  }
  return 0;
}

void batch_free(Batch* B) {
  if (!B) return;
  for (int k = 0; k < PK_COUNT; ++k)
    if (B->plan_mem[k]) (void)hipFree(B->plan_mem[k]);
  if (B->mem_dev) (void)hipFree(B->mem_dev);
  if (B->red_part) (void)hipFree(B->red_part);
  if (B->sum_part) (void)hipFree(B->sum_part);
  if (B->scal) (void)hipFree(B->scal);
  delete[] B->mem_host;
  delete B;
}

int create(const wf_member* mem, int nmem, int nlev, Batch** out) {
  if (!mem || !out || nmem < 1 || nlev < 1 || nlev > BATCH_MAX_LEVELS) return WF_EINVAL;
  const int f = 1 << (nlev - 1);
  for (int m = 0; m < nmem; ++m) {
    const wf_member& a = mem[m];
    const int n[3] = {a.n0, a.n1, a.n2};
    for (int d = 0; d < 3; ++d)
      if (n[d] < 2 * f || n[d] % f) return WF_EINVAL;
    if (!(a.dx > 0 && a.dy > 0 && a.dz > 0 && a.dt > 0)) return WF_EINVAL;
  }
  int rc = 0;
  Batch* B = new Batch();
  B->nmem = nmem;
  B->nlev = nlev;
  B->mem_host = new MemDev[nmem];
  int64_t foff = 0, loff[BATCH_MAX_LEVELS] = {}, coff = 0;
  int red = 0, sum = 0, maxc = 0;
  for (int m = 0; m < nmem; ++m) {
    const wf_member& a = mem[m];
    MemDev& M = B->mem_host[m];
    M = MemDev{};
    M.g.ni = a.n0;
    M.g.nj = a.n1;
    M.g.nk = a.n2;
    M.g.npl = 3;
    M.g.sj = a.n2 + 6;
    M.g.si = (int64_t)(a.n1 + 6) * (a.n2 + 6);
    M.off = foff;
    foff = align_up(foff + (int64_t)(a.n0 + 6) * M.g.si);
    M.dx = a.dx;
    M.dy = a.dy;
    M.dz = a.dz;
    M.rey = a.rey;
    M.ulid = a.ulid;
    M.noise = a.noise;
    M.dt = a.dt;
    M.ls = a.ls;
    M.rsum = 1.0 / (a.dx * a.dx) + 1.0 / (a.dy * a.dy) + 1.0 / (a.dz * a.dz);
    M.seed = a.seed;
    M.nlev = nlev;
    for (int l = 0; l < nlev; ++l) {
      M.ln[l][0] = a.n0 >> l;
      M.ln[l][1] = a.n1 >> l;
      M.ln[l][2] = a.n2 >> l;
      M.loff[l] = loff[l];
      loff[l] = align_up(loff[l] + (int64_t)(M.ln[l][0] + 2) * (M.ln[l][1] + 2) * (M.ln[l][2] + 2));
      const int cells = M.ln[l][0] * M.ln[l][1] * M.ln[l][2];
      B->max_lds_cells[l] = std::max(B->max_lds_cells[l], cells);
    }
    const int cc = M.ln[nlev - 1][0] * M.ln[nlev - 1][1] * M.ln[nlev - 1][2];
    if (nlev > 1 && cc <= 4096) {
      M.coff = coff;
      coff = align_up(coff + (int64_t)cc * cc);
    } else {
      M.coff = -1;
    }
    maxc = std::max(maxc, cc);
    M.red0 = red;
    red += red_blocks(a.n0, a.n1, a.n2);
    M.sum0 = sum;
    sum += cells_blocks(a.n0, a.n1, a.n2);
  }
  B->field_elems = foff;
  for (int l = 0; l < nlev; ++l) B->level_elems[l] = loff[l];
  B->cop_elems = coff;
  B->max_coarse_cells = maxc;
  B->nred = red;
  B->nsum = sum;

  rc = batch_mg_init(&B->lds);
  if (rc) {
    batch_free(B);
    return rc;
  }
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  auto fail = [&](hipError_t e) {
    rc = (int)e;
    return e != hipSuccess;
  };
  if (fail(hipMalloc((void**)&B->mem_dev, sizeof(MemDev) * nmem)) ||
      fail(hipMemcpy(B->mem_dev, B->mem_host, sizeof(MemDev) * nmem, hipMemcpyHostToDevice)) ||
      fail(hipMalloc((void**)&B->red_part, sizeof(double) * 4 * std::max(red, 1))) ||
      fail(hipMalloc((void**)&B->sum_part, sizeof(double) * std::max(sum, 1))) ||
      fail(hipMalloc((void**)&B->scal, sizeof(double) * 4 * nmem))) {
    batch_free(B);
    return rc;
  }
  std::vector<int> start(nmem + 1);
  for (int k = 0; k < PK_COUNT; ++k) {
    start[0] = 0;
    for (int m = 0; m < nmem; ++m) start[m + 1] = start[m] + member_count(k, B->mem_host[m]);
    Plan& P = B->plan[k];
    P.total = start[nmem];
    // a persistent grid: at most 8 workgroups per CU, the most a CU holds
    P.grid = std::max(1, std::min(P.total, 8 * ncu));
    if (fail(hipMalloc((void**)&B->plan_mem[k], sizeof(int) * (nmem + 1))) ||
        fail(hipMemcpy(B->plan_mem[k], start.data(), sizeof(int) * (nmem + 1),
                       hipMemcpyHostToDevice))) {
      batch_free(B);
      return rc;
    }
    P.start = B->plan_mem[k];
  }
  *out = B;
  return 0;
}

bool ok(wf_batch b) { return b != nullptr; }
Batch& B_(wf_batch b) { return *reinterpret_cast<Batch*>(b); }

}  // namespace

extern "C" {

int wf_batch_create(const wf_member* members, int nmem, int nlev, wf_batch* out) {
  Batch* B = nullptr;
  const int r = create(members, nmem, nlev, &B);
  if (r) return r;
  *out = reinterpret_cast<wf_batch>(B);
  return 0;
}

int wf_batch_destroy(wf_batch b) {
  batch_free(reinterpret_cast<Batch*>(b));
  return 0;
}

int wf_batch_layout(wf_batch b, int64_t* field_off, int64_t* field_elems, int64_t* level_off,
                    int64_t* level_elems, int64_t* cop_off, int64_t* cop_elems) {
  if (!ok(b)) return WF_EINVAL;
  const Batch& B = B_(b);
  for (int m = 0; m < B.nmem; ++m) {
    const MemDev& M = B.mem_host[m];
    if (field_off) field_off[m] = M.off;
    if (cop_off) cop_off[m] = M.coff;
    if (level_off)
      for (int l = 0; l < B.nlev; ++l) level_off[l * B.nmem + m] = M.loff[l];
  }
  if (field_elems) *field_elems = B.field_elems;
  if (level_elems)
    for (int l = 0; l < B.nlev; ++l) level_elems[l] = B.level_elems[l];
  if (cop_elems) *cop_elems = B.cop_elems;
  return 0;
}

int wf_batch_init_noise(wf_batch b, double* u, double* v, double* w, void* stream) {
  if (!ok(b) || !u || !v || !w) return WF_EINVAL;
  return batch_noise(B_(b), u, v, w, (hipStream_t)stream);
}

int wf_batch_fill_vis(wf_batch b, double* vis, void* stream) {
  if (!ok(b) || !vis) return WF_EINVAL;
  return batch_fill_vis(B_(b), vis, (hipStream_t)stream);
}

int wf_batch_wale(wf_batch b, const double* u, const double* v, const double* w, double* vis,
                  void* stream) {
  if (!ok(b) || !u || !v || !w || !vis) return WF_EINVAL;
  return batch_wale(B_(b), u, v, w, vis, (hipStream_t)stream);
}

int wf_batch_checkdt(wf_batch b, const double* u, const double* v, const double* w,
                     const double* vis, double* out4, void* stream) {
  if (!ok(b) || !u || !v || !w || !vis || !out4) return WF_EINVAL;
  return batch_checkdt(B_(b), u, v, w, vis, out4, (hipStream_t)stream);
}

int wf_batch_bc_fill(wf_batch b, double* f, int which, const int* active, void* stream) {
  if (!ok(b) || !f) return WF_EINVAL;
  return batch_bc_fill(B_(b), f, which, active, (hipStream_t)stream);
}

int wf_batch_bc_fill_uvw(wf_batch b, double* u, double* v, double* w, const int* active,
                         void* stream) {
  if (!ok(b) || !v || !w) return WF_EINVAL;
  double* f[3];
  int which[3], nf = 0;
  if (u) {
    f[nf] = u;
    which[nf++] = 0;
  }
  f[nf] = v;
  which[nf++] = 1;
  f[nf] = w;
  which[nf++] = 2;
  return batch_bc_fill_n(B_(b), f, which, nf, active, (hipStream_t)stream);
}

int wf_batch_convdiff(wf_batch b, int comp, const double* u, const double* v, const double* w,
                      const double* coo, const double* vis, double* cstar, double alpha,
                      void* stream) {
  if (!ok(b) || comp < 0 || comp > 2 || !u || !v || !w || !coo || !vis || !cstar)
    return WF_EINVAL;
  return batch_convdiff(B_(b), comp, u, v, w, coo, vis, cstar, alpha, (hipStream_t)stream);
}

int wf_batch_convdiff_save(wf_batch b, int comp, const double* u, const double* v,
                           const double* w, double* coo_save, const double* vis, double* cstar,
                           double alpha, void* stream) {
  if (!ok(b) || comp < 0 || comp > 2 || !u || !v || !w || !coo_save || !vis || !cstar)
    return WF_EINVAL;
  return batch_convdiff(B_(b), comp, u, v, w, nullptr, vis, cstar, alpha, (hipStream_t)stream,
                        coo_save);
}

int wf_batch_convdiff3_save(wf_batch b, const double* u, const double* v, const double* w,
                            const double* vis, double* ustar, double* vstar, double* wstar,
                            double* uoo, double* voo, double* woo, double alpha, void* stream) {
  if (!ok(b) || !u || !v || !w || !vis || !ustar || !vstar || !wstar || !uoo || !voo || !woo)
    return WF_EINVAL;
  return batch_convdiff3(B_(b), u, v, w, vis, nullptr, ustar, vstar, wstar, uoo, voo, woo,
                         alpha, (hipStream_t)stream);
}

int wf_batch_convdiff3(wf_batch b, const double* u, const double* v, const double* w,
                       const double* uoo, const double* voo, const double* woo,
                       const double* vis, const double* w_pre, double* ustar, double* vstar,
                       double* wstar, double alpha, void* stream) {
  if (!ok(b) || !u || !v || !w || !uoo || !voo || !woo || !vis || !w_pre || !ustar || !vstar ||
      !wstar)
    return WF_EINVAL;
  return batch_convdiff3(B_(b), u, v, w, vis, w_pre, ustar, vstar, wstar,
                         const_cast<double*>(uoo), const_cast<double*>(voo),
                         const_cast<double*>(woo), alpha, (hipStream_t)stream);
}

int wf_batch_convdiff4th_u_edges(wf_batch b, const double* u, const double* v, const double* w,
                                 const double* uoo, const double* vis, double* ustar,
                                 double alpha, void* stream) {
  if (!ok(b) || !u || !v || !w || !uoo || !vis || !ustar) return WF_EINVAL;
  return batch_convdiff_u_edges(B_(b), u, v, w, uoo, vis, ustar, alpha, (hipStream_t)stream);
}

int wf_batch_save_cross_ghosts(wf_batch b, const double* w, double* w_pre, void* stream) {
  if (!ok(b) || !w || !w_pre) return WF_EINVAL;
  return batch_save_cross_ghosts(B_(b), w, w_pre, (hipStream_t)stream);
}

int wf_batch_calvel3(wf_batch b, const double* us, const double* vs, const double* ws,
                     const double* p, double* u, double* v, double* w, double alpha,
                     const int* active, void* stream) {
  if (!ok(b) || !us || !vs || !ws || !p || !u || !v || !w) return WF_EINVAL;
  return batch_calvel3(B_(b), us, vs, ws, p, u, v, w, alpha, active, (hipStream_t)stream);
}

int wf_batch_calmas(wf_batch b, const double* u, const double* v, const double* w, double* su,
                    double* rmax, const int* active, void* stream) {
  if (!ok(b) || !u || !v || !w || !su || !rmax) return WF_EINVAL;
  return batch_calmas(B_(b), u, v, w, su, rmax, active, (hipStream_t)stream);
}

int wf_batch_calvel3_calmas(wf_batch b, const double* us, const double* vs, const double* ws,
                            const double* p, double* u, double* v, double* w, double* su,
                            double* rmax, double alpha, const int* active, void* stream) {
  if (!ok(b) || !us || !vs || !ws || !p || !u || !v || !w || !su || !rmax) return WF_EINVAL;
  return batch_calvel3_calmas(B_(b), us, vs, ws, p, u, v, w, su, rmax, alpha, active,
                              (hipStream_t)stream);
}

int wf_batch_add_ring(wf_batch b, double* p, const double* pp, const int* active,
                      void* stream) {
  if (!ok(b) || !p || !pp) return WF_EINVAL;
  return batch_add_ring(B_(b), p, pp, active, (hipStream_t)stream);
}

int wf_batch_mg_solve_mixed(wf_batch b, float* const* phi, float* const* rhs, float* const* tmp,
                            const float* cop, const double* su, double* pp, int ncycles,
                            int n_pre, int n_post, int n_coarse, int flags, const int* active,
                            void* stream) {
  if (!ok(b) || !phi || !rhs || !tmp || !su || !pp || ncycles < 0 || n_pre < 0 ||
      n_post < 0 || n_coarse < 0 || (flags & ~2))
    return WF_EINVAL;
  const Batch& B = B_(b);
  for (int l = 0; l < B.nlev; ++l)
    if (!phi[l] || !rhs[l] || !tmp[l]) return WF_EINVAL;
  if (cop)
    for (int m = 0; m < B.nmem; ++m)
      if (B.mem_host[m].coff < 0) return WF_EINVAL;
  return batch_mg_mixed(B, phi, rhs, tmp, cop, su, pp, ncycles, n_pre, n_post, n_coarse, flags,
                        active, (hipStream_t)stream);
}

int wf_batch_project_mask(wf_batch b, const double* rmax, double eps, int it, int* active,
                          int* iters, int* status, void* stream) {
  if (!ok(b) || !rmax || !active || !iters || !status) return WF_EINVAL;
  const int n = B_(b).nmem;
  hipLaunchKernelGGL(batch_mask_kernel, dim3(div_up(n, 64)), dim3(64), 0, (hipStream_t)stream,
                     n, rmax, eps, it, active, iters, status);
  WF_LAUNCHED();
  return 0;
}

int wf_batch_set_int(int* p, int n, int value, void* stream) {
  if (n <= 0) return 0;
  if (!p) return WF_EINVAL;
  hipLaunchKernelGGL(batch_set_int_kernel, dim3(div_up(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, p, n, value);
  WF_LAUNCHED();
  return 0;
}

}  // extern "C"
