// Device helpers shared by the multigrid kernels (plain FP64 V-cycle and
// the mixed-precision defect-correction driver).
#pragma once
#include "wf_common.h"

namespace wf {

struct MgDev {
  int n0, n1, n2;
  int64_t s0, s1;  // strides of i and j in the (n+2)^3 array
};

inline MgDev mg_dev(const wf_mg_level& L) {
  MgDev d;
  d.n0 = L.n0;
  d.n1 = L.n1;
  d.n2 = L.n2;
  d.s1 = L.n2 + 2;
  d.s0 = (int64_t)(L.n1 + 2) * d.s1;
  return d;
}

// Array index of real cell (i, j, k), 0-based.
__host__ __device__ inline int64_t mg_at(const MgDev& d, int i, int j, int k) {
  return (int64_t)(i + 1) * d.s0 + (int64_t)(j + 1) * d.s1 + (k + 1);
}

template <class T>
struct Coef {
  T aw, ae, as, an, ab, at, ap;
};

// _mgcoeff at real cell (i, j, k); c0..c2 = 1/h^2 in precision T.
template <class T>
__device__ inline Coef<T> mg_coef(const MgDev& d, int i, int j, int k, T c0, T c1, T c2) {
  Coef<T> c;
  c.aw = i == 0 ? T(0) : c0;
  c.ae = i == d.n0 - 1 ? T(0) : c0;
  c.as = j == 0 ? T(0) : c1;
  c.an = j == d.n1 - 1 ? T(0) : c1;
  c.ab = k == 0 ? T(0) : c2;
  c.at = k == d.n2 - 1 ? T(0) : c2;
  c.ap = -(c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

// aw*W + ae*E + as*S + an*N + ab*B + at*T (multigrid.py:85-90)
template <class T>
__device__ inline T nb_sum(const T* __restrict__ phi, int64_t x, const MgDev& d,
                           const Coef<T>& c) {
  return c.aw * phi[x - d.s0] + c.ae * phi[x + d.s0] + c.as * phi[x - d.s1] +
         c.an * phi[x + d.s1] + c.ab * phi[x - 1] + c.at * phi[x + 1];
}

// Write the face-ring images of real cell (i, j, k) holding value v.
template <class T>
__device__ inline void ring_images(T* __restrict__ phi, const MgDev& d, int i, int j, int k,
                                   int64_t x, T v) {
  if (i == 0) phi[x - d.s0] = v;
  if (i == d.n0 - 1) phi[x + d.s0] = v;
  if (j == 0) phi[x - d.s1] = v;
  if (j == d.n1 - 1) phi[x + d.s1] = v;
  if (k == 0) phi[x - 1] = v;
  if (k == d.n2 - 1) phi[x + 1] = v;
}

}  // namespace wf

namespace wf {
// Host-side helpers defined in mg.hip.
bool mg_level_ok(const wf_mg_level* L);
int64_t mg_level_elems(const wf_mg_level& L);
dim3 mg_cell_grid(int n0, int n1, int n2);
// zero0: level 0's iterate is taken as zero without being read or zeroed;
// only when mg_relax_zero_start_ok(lv[0], first sweep count of level 0).
int mg_vcycle_levels(const wf_mg_level* lv, int nlev, int npre, int npost, int ncoarse,
                     hipStream_t s, bool zero0 = false);
bool mg_relax_zero_start_ok(const wf_mg_level& L, int nsweeps);
bool mg_hier_ok(const wf_mg_level* lv, int nlev);
bool mg_field_grid_ok(const wf_grid& g);
template <class T>
int mg_zero(void* p, int64_t n, hipStream_t s);
}  // namespace wf
