// Tiles of cells for the fused 2D kernels (gfx950): region arithmetic, the
// tile numbering (XCD ranges, strips), per-tile slot tables and node sums.
// Shared by the forward step (mpm_tile.hip) and its adjoint
// (mpm_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_plan.h"
#include "mpm_step.h"

namespace dmpm {

// n / d for 0 <= n < 1024 and 1 <= d <= 64 without a division:
// floor(n * ceil(2^16 / d) / 2^16) (exact there: the rounding error of the
// multiplier, n * (ceil(2^16/d) d - 2^16) / (d 2^16), stays below 1/d).
struct SmallDiv {
  int d;
  uint32_t m;
};
__device__ __forceinline__ SmallDiv small_div(int d) {
  return SmallDiv{d, (65536u + (uint32_t)d - 1u) / (uint32_t)d};
}
__device__ __forceinline__ int sdiv(const SmallDiv& s, int n) {
  return (int)(__umul24((uint32_t)n, s.m) >> 16);
}

struct Rgn {
  int i0, i1, j0, j1;  // cells [i0, i1) x [j0, j1)
};
__device__ __forceinline__ Rgn grow(const Rgn& r, int k, int nx, int ny) {
  return Rgn{max(r.i0 - k, 0), min(r.i1 + k, nx), max(r.j0 - k, 0), min(r.j1 + k, ny)};
}
__device__ __forceinline__ bool in_rgn(const Rgn& r, int I, int J) {
  return I >= r.i0 && I < r.i1 && J >= r.j0 && J < r.j1;
}

// Position of workgroup b in the tile numbering: the tiles are split into 8
// contiguous ranges, one per XCD (workgroups are dispatched to the XCDs
// round-robin), so that an XCD steps neighbouring tiles.
__device__ __forceinline__ int xcd_pos(int b, int ntiles) {
  const int q = ntiles >> 3, r = ntiles & 7;
  const int x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}

// Row-major tile of position g of a mesh's strip order: strips of kStrip tile
// columns, row-major within a strip, so that the ~64 tiles an XCD runs at the
// same time form a block of kStrip columns whose halos overlap and share its
// L2 (16 columns measured fastest of 4, 8, 16 and 32).
constexpr int kStrip = 16;
__device__ __forceinline__ int strip_tile(int g, int ntiles, int tiles_x) {
  const int tiles_y = ntiles / tiles_x;
  const int full = tiles_x / kStrip, per = kStrip * tiles_y;
  int row, col;
  if (g < full * per) {
    const int st = g / per, rem = g - st * per;
    row = rem / kStrip;
    col = st * kStrip + (rem - row * kStrip);
  } else {
    const int w = tiles_x - full * kStrip, rem = g - full * per;
    row = rem / w;
    col = full * kStrip + (rem - row * w);
  }
  return row * tiles_x + col;
}

// The tile a workgroup steps and its simulation: mesh, material, first global
// cell and node.  tile is the global tile id (the simulation's first tile plus
// the row-major tile of its mesh); (ti, tj) its column and row in the mesh.
template <typename T> struct TileCtx {
  Grid2<T> g;
  Mat2 mat;
  int coff, noff, tile, ti, tj;
};
// The tile tables of a plan: the simulations, the simulation of each tile
// (batch plans), the tile columns and tiles of simulation 0, its geometry and
// material.
template <typename T> struct TileTab {
  const SimRec* sims;
  const int* tile_sim;
  int tiles_x, ntiles;
  Grid2<T> g;
  Mat2 mat;
};
// Context of global tile position gid (B: batch plan), read as a position of
// each mesh's strip order (strip) or as a global tile id.
template <typename T, bool B>
__device__ __forceinline__ TileCtx<T> tile_ctx_of(const TileTab<T>& a, int gid, bool strip) {
  TileCtx<T> c;
  int base = 0, tx = a.tiles_x, nt = a.ntiles;
  if constexpr (B) {
    const SimRec& r = a.sims[a.tile_sim[gid]];
    c.g.nx = r.nx;
    c.g.ny = r.ny;
    c.g.hx = T(r.hx);
    c.g.hy = T(r.hy);
    c.g.ihx = T(1) / c.g.hx;
    c.g.ihy = T(1) / c.g.hy;
    c.g.ncells = r.nx * r.ny;
    c.g.nnodes = (r.nx + 1) * (r.ny + 1);
    c.mat = a.mat;
    for (int k = 0; k < 3; ++k) c.mat.de[k] = r.de[k];
    c.coff = r.cell_off;
    c.noff = r.node_off;
    base = r.tile_off;
    tx = r.tiles_x;
    nt = tx * ((r.ny + kStepTile - 1) / kStepTile);
  } else {
    c.g = a.g;
    c.mat = a.mat;
    c.coff = 0;
    c.noff = 0;
  }
  const int lt = strip ? strip_tile(gid - base, nt, tx) : gid - base;
  c.ti = lt % tx;
  c.tj = lt / tx;
  c.tile = base + lt;
  return c;
}

// Half-open node region of tile R0: the nodes the tile owns.
__device__ __forceinline__ Rgn owned_nodes(const Rgn& R0, int nx, int ny) {
  return Rgn{R0.i0, R0.i1 + (R0.i1 == nx ? 1 : 0), R0.j0, R0.j1 + (R0.j1 == ny ? 1 : 0)};
}

// Slot segments of the rows of region R, computed by the calling wave (lane r:
// row r): A[r] = first slot of row r's cells, P[r] = slots in the rows before
// r (P[rows] = all).  With rowOf (nrow32 entries): rowOf[c] = the row holding
// virtual index 32 c, for every 32 c below min(P[rows], 32 nrow32).
__device__ __forceinline__ void row_segments(const Rgn& R, int nx, int coff, const int* cs,
                                             int* A, int* P, int* rowOf = nullptr,
                                             int nrow32 = 0) {
  const int r = threadIdx.x & 63;
  int len = 0;
  if (r < R.j1 - R.j0) {
    const int k = coff + (R.j0 + r) * nx + R.i0;
    const int s0 = cs[k];
    A[r] = s0;
    len = cs[k + (R.i1 - R.i0)] - s0;
  }
  const int own_len = len;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(len, o, 64);
    if (r >= o) len += y;
  }
  if (r < R.j1 - R.j0) P[r + 1] = len;
  if (r == 0) P[0] = 0;
  // row r holds the virtual indices [len - own_len, len)
  if (rowOf && r < R.j1 - R.j0)
    for (int c = (len - own_len + 31) >> 5; c < min((len + 31) >> 5, nrow32); ++c) rowOf[c] = r;
}

// Slot of virtual index v (below P[rows]) from the row table (nrow32 entries)
// and the slot segments of row_segments.
__device__ __forceinline__ int slot_of_virtual(int v, const int* rowOf, int nrow32, const int* A,
                                               const int* P) {
  int r = rowOf[min(v >> 5, nrow32 - 1)];
  while (P[r + 1] <= v) ++r;
  return A[r] + (v - P[r]);
}

// Cell centres and scales 2/(X1 - X0) of the columns and rows of region R
// (natural coordinates and inverse Jacobians, particle.py:180-206), by the
// calling wave.
template <typename T>
__device__ __forceinline__ void cell_tables(const Grid2<T>& g, const Rgn& R, T* colC, T* colS,
                                            T* rowC, T* rowS) {
  const int r = threadIdx.x & 63;
  if (r < R.i1 - R.i0) {
    const T X0 = g.X(R.i0 + r), X1 = g.X(R.i0 + r + 1);
    colC[r] = (X0 + X1) / T(2);
    colS[r] = T(2) / (X1 - X0);
  }
  if (r < R.j1 - R.j0) {
    const T Y0 = g.Y(R.j0 + r), Y1 = g.Y(R.j0 + r + 1);
    rowC[r] = (Y0 + Y1) / T(2);
    rowS[r] = T(2) / (Y1 - Y0);
  }
}

// Row (in R; -1: none), column and slot range [s0, s1) of each of this
// thread's cells of region R (cell lc = t + u * NT, row-major), loaded ahead
// of the row segments.
template <int C, int NT>
__device__ __forceinline__ void cell_slots_rc(const Rgn& R, int nx, int coff, const int* cs,
                                              int (&r)[C], int (&col)[C], int (&s0)[C],
                                              int (&s1)[C]) {
  const int w = R.i1 - R.i0, nc = w * (R.j1 - R.j0);
  const SmallDiv dw = small_div(w);
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int lc = threadIdx.x + u * NT;
    r[u] = -1;
    col[u] = 0;
    s0[u] = s1[u] = 0;
    if (lc < nc) {
      r[u] = sdiv(dw, lc);
      col[u] = lc - r[u] * w;
      const int k = coff + (R.j0 + r[u]) * nx + R.i0 + col[u];
      s0[u] = cs[k];
      s1[u] = cs[k + 1];
    }
  }
}

// Virtual slot range [v0, v1) of those cells from the row segments A, P.
template <int C>
__device__ __forceinline__ void cell_ranges_of(const int (&r)[C], const int (&s0)[C],
                                               const int (&s1)[C], const int* A, const int* P,
                                               int (&v0)[C], int (&v1)[C]) {
#pragma unroll
  for (int u = 0; u < C; ++u) {
    v0[u] = v1[u] = 0;
    if (r[u] >= 0) {
      v0[u] = P[r[u]] + (s0[u] - A[r[u]]);
      v1[u] = v0[u] + (s1[u] - s0[u]);
    }
  }
}

// Node sums (node_gather's order) of the nodes of the half-open node region
// NR from the cell partials of region RC (which holds every mesh cell of those
// nodes), through LDS buffer sl (4 * NQ * cells of RC).  Called by every
// thread; ends with a barrier.
template <typename T, int NQ, int C, int NPT, int NT>
__device__ __forceinline__ void node_sums_nt(T* sl, const T (&acc)[C][4 * NQ], const Rgn& RC,
                                             const Rgn& NR, int nx, int ny,
                                             T (&nacc)[NPT][NQ]) {
  const int t = threadIdx.x;
  const int wc = RC.i1 - RC.i0, nc = wc * (RC.j1 - RC.j0);
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const int dI[4] = {-1, 0, -1, 0}, dJ[4] = {-1, -1, 0, 0}, corner[4] = {2, 3, 1, 0};
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int lc = t + u * NT;
    if (lc < nc)
#pragma unroll
      for (int v = 0; v < 4 * NQ; ++v) sl[v * nc + lc] = acc[u][v];
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) nacc[w][q] = T(0);
    const int ln = t + w * NT;
    if (ln >= nn) continue;
    const int jr = sdiv(dwn, ln);
    const int i = NR.i0 + (ln - jr * wn), j = NR.j0 + jr;
    T pv[4][NQ];
    bool in[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int ci = i + dI[c], cj = j + dJ[c];
      in[c] = ci >= 0 && ci < nx && cj >= 0 && cj < ny;
      const int lc = in[c] ? (ci - RC.i0) + (cj - RC.j0) * wc : 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) pv[c][q] = sl[(NQ * corner[c] + q) * nc + lc];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c])
#pragma unroll
        for (int q = 0; q < NQ; ++q) nacc[w][q] = nacc[w][q] + pv[c][q];
  }
  __syncthreads();
}

// Slot table of the round [v0, vend) of a region's slot list: entry q (slot
// v0 + q, at tab[q * stride]) holds (row << 8) | column of its cell, written
// by the threads that own the cells (virtual ranges [cv0, cv1)).
template <int C>
__device__ __forceinline__ void fill_slot_cells(uint16_t* tab, const int (&r)[C],
                                                const int (&col)[C], const int (&cv0)[C],
                                                const int (&cv1)[C], int v0, int vend,
                                                int stride = 1) {
#pragma unroll
  for (int u = 0; u < C; ++u) {
    if (r[u] < 0) continue;
    const uint16_t code = (uint16_t)((r[u] << 8) | col[u]);
    const int q1 = min(cv1[u], vend) - v0;
    for (int q = max(cv0[u], v0) - v0; q < q1; ++q) tab[q * stride] = code;
  }
}

}  // namespace dmpm
