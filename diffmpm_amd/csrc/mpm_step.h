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

template <typename T> Grid2<T> grid_of(const dmpm_plan* P) {
  Grid2<T> g;
  g.nx = P->d.nx;
  g.ny = P->d.ny;
  g.hx = T(P->d.hx);
  g.hy = T(P->d.hy);
  g.ihx = T(1) / g.hx;
  g.ihy = T(1) / g.hy;
  g.ncells = P->ncells;
  g.nnodes = P->nnodes;
  return g;
}

// Nodes of cell (I, J) in corner order ll, lr, ur, ul (element.py:692-723).
__device__ __forceinline__ void cell_nodes(int nx, int I, int J, int ids[4]) {
  const int ll = J * (nx + 1) + I;
  ids[0] = ll;
  ids[1] = ll + 1;
  ids[2] = ll + nx + 2;
  ids[3] = ll + nx + 1;
}

// Sum over the (up to four) cells of node (i, j), canonical order, of the
// partials of quantity q: corner 2 of cell (i-1,j-1), 3 of (i,j-1), 1 of
// (i-1,j), 0 of (i,j).
template <typename T, int NQ>
__device__ __forceinline__ void node_gather(const Grid2<T>& g, const T* __restrict__ part, int i,
                                            int j, T acc[NQ]) {
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
      acc[q] = acc[q] + part[(int64_t)(NQ * corner[c] + q) * g.ncells + cells[c]];
  }
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace dmpm
