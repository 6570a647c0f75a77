// Device types and helpers shared by the fused 2D step (mpm_tile.hip) and its
// adjoint (mpm_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_common.h"
#include "mpm_plan.h"

namespace dmpm {

constexpr int kThreads = 256;  // workgroup size of the particle-parallel kernels

struct Mat2 {
  int kind;
  int per_E;
  double de[3];
};

template <typename T> struct Grid2 {
  int nx, ny;
  T hx, hy, ihx, ihy;
  int ncells, nnodes;
  __device__ __forceinline__ T X(int i) const { return T(i) * hx; }
  __device__ __forceinline__ T Y(int j) const { return T(j) * hy; }
  __device__ __forceinline__ Geo<T> geo() const {
    Geo<T> g;
    g.nx = nx;
    g.ny = ny;
    g.hx = hx;
    g.hy = hy;
    g.ihx = ihx;
    g.ihy = ihy;
    g.ncells = ncells;
    g.nnodes = nnodes;
    return g;
  }
};

template <typename T>
__device__ __forceinline__ void stress_inc(const Mat2& m, T E_p, const T d[6], T ds[4]) {
  if (m.kind == DMPM_SIMPLE_MATERIAL) {
    T E = T(m.de[0]);
    for (int k = 0; k < 4; ++k) ds[k] = d[k] * E;
    return;
  }
  T a1 = T(m.de[0]), a2 = T(m.de[1]), G = T(m.de[2]);
  if (m.per_E) {
    a1 = a1 * E_p;
    a2 = a2 * E_p;
    G = G * E_p;
  }
  ds[0] = (a1 * d[0] + a2 * d[1]) + a2 * d[2];
  ds[1] = (a2 * d[0] + a1 * d[1]) + a2 * d[2];
  ds[2] = (a2 * d[0] + a2 * d[1]) + a1 * d[2];
  ds[3] = G * d[3];
}

// Snapped strain rate [r0, r1, 0, r3] at (xi, eta) of cell (I, J) from the
// four nodal velocities (particle.py:281-313).
template <typename T>
__device__ __forceinline__ void rate2(const Grid2<T>& g, int I, int J, T xi, T eta, const T vx[4],
                                      const T vy[4], T r[6]) {
  T ix, iy;
  q4_invJ(xi, eta, g.X(I), g.X(I + 1), g.Y(J), g.Y(J + 1), ix, iy);
  T L00 = T(0), L01 = T(0), L10 = T(0), L11 = T(0);
  for (int k = 0; k < 4; ++k) {
    T dx, dy;
    q4_dNdx(k, xi, eta, ix, iy, dx, dy);
    L00 = L00 + dx * vx[k];
    L01 = L01 + dx * vy[k];
    L10 = L10 + dy * vx[k];
    L11 = L11 + dy * vy[k];
  }
  r[0] = snap(L00);
  r[1] = snap(L11);
  r[2] = T(0);
  r[3] = snap(L01 + L10);
  r[4] = T(0);
  r[5] = T(0);
}

// Same as rate2 with the inverse-Jacobian scales given.
template <typename T>
__device__ __forceinline__ void rate2s(T ix, T iy, T xi, T eta, const T vx[4], const T vy[4],
                                       T r[6]) {
  T L00 = T(0), L01 = T(0), L10 = T(0), L11 = T(0);
  for (int k = 0; k < 4; ++k) {
    T dx, dy;
    q4_dNdx(k, xi, eta, ix, iy, dx, dy);
    L00 = L00 + dx * vx[k];
    L01 = L01 + dx * vy[k];
    L10 = L10 + dy * vx[k];
    L11 = L11 + dy * vy[k];
  }
  r[0] = snap(L00);
  r[1] = snap(L11);
  r[2] = T(0);
  r[3] = snap(L01 + L10);
  r[4] = T(0);
  r[5] = T(0);
}

// The simulations of a plan as the kernels see them.  A plan holds one
// simulation or, for a batch plan, several with their own meshes, materials,
// particles and constraints, laid end to end in global index spaces: the
// particles of simulation s are [part_off, part_off + n), its cells
// [cell_off, cell_off + nx*ny) and its nodes [node_off, node_off +
// (nx+1)(ny+1)).  Particles outside their own mesh all share the key
// ncells (the total).  One launch of each kernel serves every simulation.
template <typename T> struct Mesh2 {
  const SimRec* rec;  // device table, nsims entries
  int nsims;
  int ncells, nnodes, n;  // totals
  Grid2<T> g0;            // geometry of simulation 0 (the only one when nsims == 1)

  // index of the simulation owning global index v in the space with offsets
  // rec[s].*off (nondecreasing)
  template <typename F> __device__ __forceinline__ int find(int v, F off) const {
    if (nsims == 1) return 0;
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off(rec[mid]) <= v) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
  __device__ __forceinline__ int sim_p(int p) const {
    return find(p, [](const SimRec& r) { return r.part_off; });
  }
  __device__ __forceinline__ int sim_c(int c) const {
    return find(c, [](const SimRec& r) { return r.cell_off; });
  }
  __device__ __forceinline__ int sim_n(int nd) const {
    return find(nd, [](const SimRec& r) { return r.node_off; });
  }
  __device__ __forceinline__ int coff(int s) const { return nsims == 1 ? 0 : rec[s].cell_off; }
  __device__ __forceinline__ int noff(int s) const { return nsims == 1 ? 0 : rec[s].node_off; }
  __device__ __forceinline__ Grid2<T> grid(int s) const {
    if (nsims == 1) return g0;
    const SimRec& r = rec[s];
    Grid2<T> g;
    g.nx = r.nx;
    g.ny = r.ny;
    g.hx = T(r.hx);
    g.hy = T(r.hy);
    g.ihx = T(1) / g.hx;
    g.ihy = T(1) / g.hy;
    g.ncells = r.nx * r.ny;
    g.nnodes = (r.nx + 1) * (r.ny + 1);
    return g;
  }
  // sort key of local element id e of simulation s
  __device__ __forceinline__ int key(int s, const Grid2<T>& g, int e) const {
    return (e >= 0 && e < g.ncells) ? coff(s) + e : ncells;
  }
  // material coefficients of simulation s
  __device__ __forceinline__ Mat2 mat(const Mat2& m, int s) const {
    if (nsims == 1) return m;
    Mat2 r = m;
    for (int k = 0; k < 3; ++k) r.de[k] = rec[s].de[k];
    return r;
  }
};

template <typename T> Mesh2<T> mesh_of(const dmpm_plan* P) {
  Mesh2<T> M;
  M.rec = P->sims;
  M.nsims = P->nsims;
  M.ncells = P->ncells;
  M.nnodes = P->nnodes;
  M.n = (int)P->n;
  const SimRec& r = P->sims_h[0];
  Grid2<T>& g = M.g0;
  g.nx = r.nx;
  g.ny = r.ny;
  g.hx = T(r.hx);
  g.hy = T(r.hy);
  g.ihx = T(1) / g.hx;
  g.ihy = T(1) / g.hy;
  g.ncells = r.nx * r.ny;
  g.nnodes = (r.nx + 1) * (r.ny + 1);
  return M;
}

// Nodes of cell (I, J) in corner order ll, lr, ur, ul (element.py:692-723),
// offset by the simulation's first node.
__device__ __forceinline__ void cell_nodes(int nx, int I, int J, int ids[4], int noff = 0) {
  const int ll = noff + J * (nx + 1) + I;
  ids[0] = ll;
  ids[1] = ll + 1;
  ids[2] = ll + nx + 2;
  ids[3] = ll + nx + 1;
}

// Sum over the (up to four) cells of node (i, j) of a mesh g whose cells
// start at global index coff, canonical order, of the partials of quantity q
// (rows of stride `stride`): corner 2 of cell (i-1,j-1), 3 of (i,j-1), 1 of
// (i-1,j), 0 of (i,j).
template <typename T, int NQ>
__device__ __forceinline__ void node_gather(const Grid2<T>& g, int coff, int stride,
                                            const T* __restrict__ part, int i, int j, T acc[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = T(0);
  const int cells[4] = {(j - 1) * g.nx + i - 1, (j - 1) * g.nx + i, j * g.nx + i - 1,
                        j * g.nx + i};
  const bool ok[4] = {j > 0 && i > 0, j > 0 && i < g.nx, j < g.ny && i > 0, j < g.ny && i < g.nx};
  const int corner[4] = {2, 3, 1, 0};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (!ok[c]) continue;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      acc[q] = acc[q] + part[(int64_t)(NQ * corner[c] + q) * stride + coff + cells[c]];
  }
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace dmpm
