// Fused 2D explicit step on node tiles (gfx950).
//
// One step is four launches:
//   k_sort_coop     cooperative sort by cell: nothing when no particle changed
//                   cell, an O(N) merge of the recorded changes into the stable
//                   order when a few did, a full counting sort otherwise
//   k_tile<KIN>     P2G of mass and momentum, nodal velocity, constraints
//   k_tile<FORCE_*> (USF: strain/volume/stress update of the tile's particles)
//                   P2G of body and internal forces, nodal forces, nodal
//                   acceleration/velocity update, constraints
//   k_g2p_*         G2P velocity/position (USL: + strain/volume/stress update),
//                   flags cell changes for the next sort
//
// A workgroup owns a tile of kTX x ty nodes and streams the particles of the
// cells around it (cell rows j0-1 .. j0+ty-1, columns i0-1 .. i0+kTX-1) row by
// row through LDS.  Node (i, j) sums its contributions in the same order as
// the per-phase kernels — cells (i-1,j-1), (i,j-1), (i-1,j), (i,j), ascending
// particle index within a cell — with the same per-term arithmetic, so both
// paths produce bit-identical nodal sums, without atomics.
//
// Particles in the one-cell halo are read by two tiles.  In USF the force
// kernel updates the stress of the particles it owns; so that a neighbour
// still reads the step-start values of shared halo particles, the ten updated
// fields are double-buffered between the caller's block and a plan workspace,
// and the kinematic nodal velocity lives in its own workspace.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>

#include "mpm_common.h"
#include "mpm_plan.h"

namespace cg = cooperative_groups;
using namespace dmpm;

namespace {

constexpr int kThreads = 256;  // sort / G2P workgroups
constexpr int kDB = 10;       // double-buffered particle fields (USF)
constexpr int kSmallCell = 64;
constexpr int kBigCellMax = 8192;

// LDS particle capacity per chunk: ~4 rows of particles per node column.
template <typename T, int TXN> struct Chunk {
  static constexpr int value = (sizeof(T) == 4 ? 8 : 4) * TXN;
};

// eps_xx, eps_yy, gamma_xy, sig_xx, sig_yy, sig_zz, sig_xy, V, rho, eps_v
__host__ __device__ constexpr int db_row(int k) {
  return k < 2 ? DMPM_P_STRAIN + k
         : k == 2 ? DMPM_P_STRAIN + 3
         : k < 7 ? DMPM_P_STRESS + (k - 3)
         : k == 7 ? DMPM_P_VOLUME
         : k == 8 ? DMPM_P_DENSITY
                  : DMPM_P_VOLSTRAIN_C;
}

template <typename T> struct DB {
  T* cur[kDB];
  T* nxt[kDB];
};

struct Mat2 {
  int kind;
  int per_E;
  double de[3];
};

struct TileShape {
  int ntx, nty, ty;
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

// Workgroup -> tile such that consecutive tiles (which share halo cell rows)
// land on the same XCD (workgroups are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ int tile_of_block(int b, int nb) {
  const int x = b % 8, k = b / 8, q = nb / 8, rem = nb % 8;
  const int base = x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q;
  return base + k;
}

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

#define PR(r) (pb + (int64_t)(r) * ld)
#define NR(r) (nb + (int64_t)(r) * nld)

enum { MODE_KIN = 0, MODE_FORCE_USF = 1, MODE_FORCE_USL = 2 };

template <typename T, int MODE, int kTX>
__global__ __launch_bounds__(2 * kTX) void k_tile(
    Grid2<T> g, TileShape tl, T* __restrict__ pb, int64_t ld, SortRef sr,
    T* __restrict__ nb, int64_t nld,
    T* __restrict__ vkin, const int* __restrict__ cidx, const int* __restrict__ cmask,
    const double* __restrict__ cval, DB<T> db, Mat2 mat, T gx, T gy, T dt, T t,
    const int* __restrict__ nfd, const int* __restrict__ nf_off, const int* __restrict__ nf_fn,
    const double* __restrict__ nf_val, FnTable fns, int temps) {
  constexpr int CH = Chunk<T, kTX>::value;
  constexpr int NT = 2 * kTX;
  constexpr bool KIN = MODE == MODE_KIN;
  constexpr bool USF = MODE == MODE_FORCE_USF;
  constexpr int NQ = KIN ? 3 : 5;
  constexpr int CHF = KIN ? 1 : CH;  // gradient scales are staged for force modes only
  __shared__ T s_xi[CH], s_eta[CH], s_ix[CHF], s_q[NQ][CH];
  __shared__ int s_cs[2][kTX + 2];
  __shared__ T s_mx[kTX + 2], s_sx[kTX + 2];  // per cell column: (X0+X1)/2, 2/(X1-X0)
  __shared__ T s_v[USF ? 2 : 1][2][2][kTX + 2];  // [buf][node row r|r+1][x|y][col]

  const int ident = sr.flags[1];
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  const int tile = tile_of_block(blockIdx.x, gridDim.x);
  const int tyi = tile % tl.nty, txi = tile / tl.nty;
  const int i0 = txi * kTX, j0 = tyi * tl.ty;
  const int tid = threadIdx.x, grp = tid / kTX, col = tid % kTX;  // blockDim.x == 2*kTX
  const int i = i0 + col;
  const bool icol = i <= g.nx;
  const int c0 = max(i0 - 1, 0), c1 = min(i0 + kTX, g.nx);
  const int ncell = c1 - c0;

  if (tid < ncell) {
    const T X0 = g.X(c0 + tid), X1 = g.X(c0 + tid + 1);
    s_mx[tid] = (X0 + X1) / T(2);
    s_sx[tid] = T(2) / (X1 - X0);
  }

  T acc[4] = {T(0), T(0), T(0), T(0)};

  // cell offsets (and, for USF, nodal velocities) of the row segment,
  // prefetched one cell row ahead
  const int rfirst = max(j0 - 1, 0), rlast = min(j0 + tl.ty, g.ny);  // [rfirst, rlast)
  const bool loader = tid <= ncell && rfirst < rlast;
  int cs_next = loader ? cs[rfirst * g.nx + c0 + tid] : 0;
  T v_next[4] = {T(0), T(0), T(0), T(0)};
  if (USF && loader) {
    const int nd = rfirst * (g.nx + 1) + c0 + tid;
    v_next[0] = vkin[nd];
    v_next[1] = vkin[g.nnodes + nd];
    v_next[2] = vkin[nd + g.nx + 1];
    v_next[3] = vkin[g.nnodes + nd + g.nx + 1];
  }

  for (int r = j0 - 1; r < j0 + tl.ty; ++r) {
    const bool top = grp == ((r + 1) & 1);
    const int my_row = top ? r + 1 : r;
    const bool node_ok = icol && my_row >= j0 && my_row < j0 + tl.ty && my_row <= g.ny;
    if (r >= 0 && r < g.ny) {
      const int cb_ = (r - rfirst) & 1;
      __syncthreads();
      if (tid <= ncell) {
        s_cs[cb_][tid] = cs_next;
        if (USF) {
          s_v[cb_ & (USF ? 1 : 0)][0][0][tid] = v_next[0];
          s_v[cb_ & (USF ? 1 : 0)][0][1][tid] = v_next[1];
          s_v[cb_ & (USF ? 1 : 0)][1][0][tid] = v_next[2];
          s_v[cb_ & (USF ? 1 : 0)][1][1][tid] = v_next[3];
        }
        if (r + 1 < rlast) {
          cs_next = cs[(r + 1) * g.nx + c0 + tid];
          if (USF) {
            const int nd = (r + 1) * (g.nx + 1) + c0 + tid;
            v_next[0] = vkin[nd];
            v_next[1] = vkin[g.nnodes + nd];
            v_next[2] = vkin[nd + g.nx + 1];
            v_next[3] = vkin[g.nnodes + nd + g.nx + 1];
          }
        }
      }
      __syncthreads();
      const int* __restrict__ rcs = s_cs[cb_];
      const int sb = rcs[0], se = rcs[ncell];
      const bool row_owned = r >= j0 && r < j0 + tl.ty;
      const T Y0 = g.Y(r), Y1 = g.Y(r + 1);
      const T my = (Y0 + Y1) / T(2), sy = T(2) / (Y1 - Y0);
      for (int cb = sb; cb < se; cb += CH) {
        const int ce = min(cb + CH, se);
        if (cb != sb) __syncthreads();
        // ---- stage the chunk's particles -------------------------------------
        for (int s = cb + tid; s < ce; s += NT) {
          const int p = ident ? s : perm[s];
          const int q = s - cb;
          const T x = PR(DMPM_P_X)[p], y = PR(DMPM_P_X + 1)[p];
          // the sort put p in cell row r, columns [c0, c1); its column is that
          // of its element id (A1: x >= X_nx wraps to column 0 of the next
          // row, points below or left of the mesh belong to element 0)
          const int e = element_id<T, 2>(g.geo(), x, y);
          const int I = e % g.nx;
          const int kk = min(max(I - c0, 0), ncell - 1);
          const T xi = (x - s_mx[kk]) * s_sx[kk];  // particle.py:203-205
          const T eta = (y - my) * sy;
          const bool owned = row_owned && I >= i0 && I < i0 + kTX;
          s_xi[q] = xi;
          s_eta[q] = eta;
          const T m = PR(DMPM_P_MASS)[p];
          s_q[0][q] = m;
          if (KIN) {
            s_q[1][q] = m * PR(DMPM_P_V)[p];
            s_q[2][q] = m * PR(DMPM_P_V + 1)[p];
            if (temps && owned) {
              PR(DMPM_P_XI)[p] = xi;
              PR(DMPM_P_XI + 1)[p] = eta;
            }
          } else {
            constexpr int q3 = KIN ? 0 : 3, q4 = KIN ? 0 : 4;
            s_ix[q % CHF] = s_sx[kk];
            if (MODE == MODE_FORCE_USL) {
              s_q[1][q] = -PR(DMPM_P_VOLUME)[p];
              s_q[2][q] = PR(DMPM_P_STRESS + 0)[p];
              s_q[q3][q] = PR(DMPM_P_STRESS + 1)[p];
              s_q[q4][q] = PR(DMPM_P_STRESS + 3)[p];
            } else {
              // USF strain / volume / stress update (particle.py:249-327)
              const int vb = cb_ & (USF ? 1 : 0);
              const T vx[4] = {s_v[vb][0][0][kk], s_v[vb][0][0][kk + 1], s_v[vb][1][0][kk + 1],
                               s_v[vb][1][0][kk]};
              const T vy[4] = {s_v[vb][0][1][kk], s_v[vb][0][1][kk + 1], s_v[vb][1][1][kk + 1],
                               s_v[vb][1][1][kk]};
              const T ix = s_sx[kk], iy = sy;
              T rr[6], rc[6], d[6];
              rate2s(ix, iy, xi, eta, vx, vy, rr);
              rate2s(ix, iy, T(0), T(0), vx, vy, rc);
              for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
              const T dvol = dt * (rc[0] + rc[1]);
              T ds[4];
              stress_inc(mat, mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
              const T f = T(1) + dvol;
              T nv[kDB];
              nv[0] = db.cur[0][p] + d[0];
              nv[1] = db.cur[1][p] + d[1];
              nv[2] = db.cur[2][p] + d[3];
              for (int k = 0; k < 4; ++k) nv[3 + k] = db.cur[3 + k][p] + ds[k];
              nv[7] = db.cur[7][p] * f;
              nv[8] = db.cur[8][p] / f;
              nv[9] = db.cur[9][p] + dvol;
              if (owned) {
                for (int k = 0; k < kDB; ++k) db.nxt[k][p] = nv[k];
                if (temps) {
                  for (int k = 0; k < 6; ++k) {
                    PR(DMPM_P_STRAIN_RATE + k)[p] = rr[k];
                    PR(DMPM_P_DSTRAIN + k)[p] = d[k];
                  }
                  PR(DMPM_P_DVOLSTRAIN)[p] = dvol;
                }
              }
              s_q[1][q] = -nv[7];
              s_q[2][q] = nv[3];
              s_q[q3][q] = nv[4];
              s_q[q4][q] = nv[6];
            }
          }
        }
        __syncthreads();
        // ---- node threads accumulate cells i-1 then i of this cell row ------------
        if (node_ok) {
          for (int dc = -1; dc <= 0; ++dc) {
            const int c = i + dc;
            if (c < c0 || c >= c1) continue;
            const int lo = max(rcs[c - c0], cb) - cb, hi = min(rcs[c - c0 + 1], ce) - cb;
            const int corner = top ? (dc < 0 ? 2 : 3) : (dc < 0 ? 1 : 0);
            for (int q = lo; q < hi; ++q) {
              const T xi = s_xi[q], eta = s_eta[q];
              const T N = q4_N<T>(corner, xi, eta);
              const T m = s_q[0][q];
              if (KIN) {
                acc[0] = acc[0] + m * N;                 // element.py:140
                acc[1] = acc[1] + N * s_q[1][q];         // element.py:172
                acc[2] = acc[2] + N * s_q[2][q];
              } else {
                constexpr int q3 = KIN ? 0 : 3, q4 = KIN ? 0 : 4;
                acc[0] = acc[0] + N * (m * gx);          // element.py:254-256
                acc[1] = acc[1] + N * (m * gy);
                T dx, dy;
                q4_dNdx(corner, xi, eta, s_ix[q % CHF], sy, dx, dy);
                const T nV = s_q[1][q];
                acc[2] = acc[2] + nV * (dx * s_q[2][q] + dy * s_q[q4][q]);  // element.py:868-878
                acc[3] = acc[3] + nV * (dy * s_q[q3][q] + dx * s_q[q4][q]);
              }
            }
          }
        }
      }
    }
    // ---- node row r is complete: epilogue ---------------------------------------
    if (!top) {
      if (node_ok) {
        const int nd = my_row * (g.nx + 1) + i;
        const int cr = cidx ? cidx[nd] : -1;
        const int mask = cr >= 0 ? cmask[cr] : 0;
        if (MODE == MODE_KIN) {  // element.py:122-202, constraint.py:29-44
          const T m = acc[0];
          T p[2] = {snap(acc[1]), snap(acc[2])};
          NR(DMPM_N_MASS)[nd] = m;
          for (int d = 0; d < 2; ++d) {
            T v = (m == T(0)) ? NR(DMPM_N_V + d)[nd] : p[d] / m;
            v = snap(v);
            if (mask & (1 << d)) {
              const T c = T(cval[2 * cr + d]);
              v = c;
              p[d] = m * c;
            }
            NR(DMPM_N_P + d)[nd] = p[d];
            vkin[(int64_t)d * g.nnodes + nd] = v;
          }
        } else {  // element.py:204-329
          const T m = NR(DMPM_N_MASS)[nd];
          for (int d = 0; d < 2; ++d) {
            T fe = acc[d];
            const int qg = nfd ? nfd[(int64_t)d * g.nnodes + nd] : -1;
            if (qg >= 0)
              for (int k = nf_off[qg]; k < nf_off[qg + 1]; ++k)
                fe = fe + fn_value<T>(fns, nf_fn[k], t) * T(nf_val[k]);
            const T fi = acc[2 + d];
            NR(DMPM_N_FEXT + d)[nd] = fe;
            NR(DMPM_N_FINT + d)[nd] = fi;
            const T F = (fi + fe) + NR(DMPM_N_FDAMP + d)[nd];
            T a = nan_to_num(F / m);
            T v = vkin[(int64_t)d * g.nnodes + nd] + a * dt;
            T pm = m * v;
            v = snap(v);
            a = snap(a);
            if (mask & (1 << d)) {
              const T c = T(cval[2 * cr + d]);
              v = snap(c);
              pm = m * c;
              a = T(0);
            }
            NR(DMPM_N_V + d)[nd] = v;
            NR(DMPM_N_A + d)[nd] = a;
            NR(DMPM_N_P + d)[nd] = pm;
          }
        }
      }
      acc[0] = acc[1] = acc[2] = acc[3] = T(0);
    }
  }
}

// G2P (particle.py:208-247); USL also updates strain/volume/stress with the
// step-start natural coordinates and the updated nodal velocities
// (scheme.py:127-133).  Sets flags[0] when a particle changes cell.
// Records a particle whose cell (sort key) changed: (particle, old key, new
// element id) is appended for the incremental re-sort; past movcap records a
// full sort is requested.
__device__ __forceinline__ void record_move(int* __restrict__ flags, int* __restrict__ mov,
                                            int movcap, int p, int kold, int enew) {
  const int k = atomicAdd(&flags[5], 1);
  if (k < movcap) {
    mov[kMovRec * k] = p;
    mov[kMovRec * k + 1] = kold;
    mov[kMovRec * k + 2] = enew;
  } else {
    flags[6] = 1;
  }
}

template <typename T, bool USL>
__global__ __launch_bounds__(kThreads) void k_g2p_tile(Grid2<T> g, T* __restrict__ pb, int64_t ld,
                                                       int n, const T* __restrict__ nb,
                                                       int64_t nld, int* __restrict__ flags,
                                                       int* __restrict__ mov, int movcap,
                                                       DB<T> db, Mat2 mat, T dt, int vu,
                                                       int temps, int usf_db) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const T x = PR(DMPM_P_X)[p], y = PR(DMPM_P_X + 1)[p];
  const int e = element_id<T, 2>(g.geo(), x, y);
  const bool inside = e >= 0 && e < g.ncells;
  if (!inside) {
    // Outside the mesh the reference's gathers degenerate to NaN; the state
    // of such a particle becomes NaN (it is in no cell, so no tile updated it).
    const T nan = qnan<T>();
    PR(DMPM_P_X)[p] = nan;
    PR(DMPM_P_X + 1)[p] = nan;
    PR(DMPM_P_V)[p] = nan;
    PR(DMPM_P_V + 1)[p] = nan;
    if (usf_db || USL)
      for (int k = 0; k < kDB; ++k) (USL ? PR(db_row(k)) : db.nxt[k])[p] = nan;
    if (usf_db || USL) {
      PR(DMPM_P_STRAIN + 2)[p] = nan;
      PR(DMPM_P_STRAIN + 4)[p] = nan;
      PR(DMPM_P_STRAIN + 5)[p] = nan;
      PR(DMPM_P_STRESS + 4)[p] = nan;
      PR(DMPM_P_STRESS + 5)[p] = nan;
    }
    if (temps) {
      PR(DMPM_P_XI)[p] = nan;
      PR(DMPM_P_XI + 1)[p] = nan;
      PR(DMPM_P_MOMENTUM)[p] = nan;
      PR(DMPM_P_MOMENTUM + 1)[p] = nan;
      for (int k = 0; k < 6; ++k) {
        PR(DMPM_P_STRAIN_RATE + k)[p] = nan;
        PR(DMPM_P_DSTRAIN + k)[p] = nan;
      }
      PR(DMPM_P_DVOLSTRAIN)[p] = nan;
    }
    // NaN positions search to element 0 (element.py:827-837)
    record_move(flags, mov, movcap, p, g.ncells, 0);
    return;
  }
  const int I = e % g.nx, J = e / g.nx;
  const T xi = natural(x, g.X(I), g.X(I + 1));
  const T eta = natural(y, g.Y(J), g.Y(J + 1));
  const int ll = J * (g.nx + 1) + I;
  const int ids[4] = {ll, ll + 1, ll + g.nx + 2, ll + g.nx + 1};
  T N[4];
  q4_N4(xi, eta, N);
  T vx[4], vy[4];
  T nvx = T(0), nvy = T(0), nax = T(0), nay = T(0);
  for (int k = 0; k < 4; ++k) {
    vx[k] = NR(DMPM_N_V)[ids[k]];
    vy[k] = NR(DMPM_N_V + 1)[ids[k]];
    nvx = nvx + N[k] * vx[k];
    nvy = nvy + N[k] * vy[k];
    nax = nax + N[k] * NR(DMPM_N_A)[ids[k]];
    nay = nay + N[k] * NR(DMPM_N_A + 1)[ids[k]];
  }
  const T pvx = vu ? nvx : PR(DMPM_P_V)[p] + nax * dt;
  const T pvy = vu ? nvy : PR(DMPM_P_V + 1)[p] + nay * dt;
  const T x1 = x + nvx * dt, y1 = y + nvy * dt;
  PR(DMPM_P_V)[p] = pvx;
  PR(DMPM_P_V + 1)[p] = pvy;
  PR(DMPM_P_X)[p] = x1;
  PR(DMPM_P_X + 1)[p] = y1;
  if (temps) {
    const T m = PR(DMPM_P_MASS)[p];
    PR(DMPM_P_MOMENTUM)[p] = m * pvx;
    PR(DMPM_P_MOMENTUM + 1)[p] = m * pvy;
  }
  if (USL) {
    T rr[6], rc[6], d[6];
    rate2(g, I, J, xi, eta, vx, vy, rr);
    rate2(g, I, J, T(0), T(0), vx, vy, rc);
    for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
    const T dvol = dt * (rc[0] + rc[1]);
    T ds[4];
    stress_inc(mat, mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
    const T f = T(1) + dvol;
    PR(DMPM_P_STRAIN + 0)[p] = PR(DMPM_P_STRAIN + 0)[p] + d[0];
    PR(DMPM_P_STRAIN + 1)[p] = PR(DMPM_P_STRAIN + 1)[p] + d[1];
    PR(DMPM_P_STRAIN + 3)[p] = PR(DMPM_P_STRAIN + 3)[p] + d[3];
    for (int k = 0; k < 4; ++k) PR(DMPM_P_STRESS + k)[p] = PR(DMPM_P_STRESS + k)[p] + ds[k];
    PR(DMPM_P_VOLUME)[p] = PR(DMPM_P_VOLUME)[p] * f;
    PR(DMPM_P_DENSITY)[p] = PR(DMPM_P_DENSITY)[p] / f;
    PR(DMPM_P_VOLSTRAIN_C)[p] = PR(DMPM_P_VOLSTRAIN_C)[p] + dvol;
    if (temps) {
      for (int k = 0; k < 6; ++k) {
        PR(DMPM_P_STRAIN_RATE + k)[p] = rr[k];
        PR(DMPM_P_DSTRAIN + k)[p] = d[k];
      }
      PR(DMPM_P_DVOLSTRAIN)[p] = dvol;
    }
  }
  const int e1 = element_id<T, 2>(g.geo(), x1, y1);
  const int k1 = (e1 >= 0 && e1 < g.ncells) ? e1 : g.ncells;
  if (k1 != e) record_move(flags, mov, movcap, p, e, e1);
}

// USF strain / volume / stress update (particle.py:249-327) streamed over the
// particles in storage order, in place, from the kinematic nodal velocities;
// used before k_tile<FORCE_USL> when USF runs split (DMPM_OPT_USF_SPLIT).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_stress_tile(Grid2<T> g, T* __restrict__ pb,
                                                          int64_t ld, int n,
                                                          const T* __restrict__ vkin, Mat2 mat,
                                                          T dt, int temps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const T x = PR(DMPM_P_X)[p], y = PR(DMPM_P_X + 1)[p];
  const int e = element_id<T, 2>(g.geo(), x, y);
  if (!(e >= 0 && e < g.ncells)) {  // outside the mesh: the state becomes NaN
    const T nan = qnan<T>();
    for (int k = 0; k < 6; ++k) {
      PR(DMPM_P_STRAIN + k)[p] = nan;
      PR(DMPM_P_STRESS + k)[p] = nan;
    }
    PR(DMPM_P_VOLUME)[p] = nan;
    PR(DMPM_P_DENSITY)[p] = nan;
    PR(DMPM_P_VOLSTRAIN_C)[p] = nan;
    if (temps) {
      for (int k = 0; k < 6; ++k) {
        PR(DMPM_P_STRAIN_RATE + k)[p] = nan;
        PR(DMPM_P_DSTRAIN + k)[p] = nan;
      }
      PR(DMPM_P_DVOLSTRAIN)[p] = nan;
    }
    return;
  }
  const int I = e % g.nx, J = e / g.nx;
  const T xi = natural(x, g.X(I), g.X(I + 1));
  const T eta = natural(y, g.Y(J), g.Y(J + 1));
  const int ll = J * (g.nx + 1) + I;
  const int ids[4] = {ll, ll + 1, ll + g.nx + 2, ll + g.nx + 1};
  T vx[4], vy[4];
  for (int k = 0; k < 4; ++k) {
    vx[k] = vkin[ids[k]];
    vy[k] = vkin[g.nnodes + ids[k]];
  }
  T rr[6], rc[6], d[6];
  rate2(g, I, J, xi, eta, vx, vy, rr);
  rate2(g, I, J, T(0), T(0), vx, vy, rc);
  for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
  const T dvol = dt * (rc[0] + rc[1]);
  T ds[4];
  stress_inc(mat, mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
  const T f = T(1) + dvol;
  PR(DMPM_P_STRAIN + 0)[p] = PR(DMPM_P_STRAIN + 0)[p] + d[0];
  PR(DMPM_P_STRAIN + 1)[p] = PR(DMPM_P_STRAIN + 1)[p] + d[1];
  PR(DMPM_P_STRAIN + 3)[p] = PR(DMPM_P_STRAIN + 3)[p] + d[3];
  for (int k = 0; k < 4; ++k) PR(DMPM_P_STRESS + k)[p] = PR(DMPM_P_STRESS + k)[p] + ds[k];
  PR(DMPM_P_VOLUME)[p] = PR(DMPM_P_VOLUME)[p] * f;
  PR(DMPM_P_DENSITY)[p] = PR(DMPM_P_DENSITY)[p] / f;
  PR(DMPM_P_VOLSTRAIN_C)[p] = PR(DMPM_P_VOLSTRAIN_C)[p] + dvol;
  if (temps) {
    for (int k = 0; k < 6; ++k) {
      PR(DMPM_P_STRAIN_RATE + k)[p] = rr[k];
      PR(DMPM_P_DSTRAIN + k)[p] = d[k];
    }
    PR(DMPM_P_DVOLSTRAIN)[p] = dvol;
  }
}

// Element ids and sort keys from the positions at the start of a run; marks
// the sort dirty when a key changed.
template <typename T>
__global__ void k_tile_check(Grid2<T> g, const T* __restrict__ pb, int64_t ld, int n,
                             int* __restrict__ eid, int* __restrict__ key, int* __restrict__ flags) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
  eid[p] = e;
  const int k = (e >= 0 && e < g.ncells) ? e : g.ncells;
  if (key[p] != k) {
    key[p] = k;
    flags[0] = 1;
  }
}

// ---- cooperative counting sort ------------------------------------------------
__device__ __forceinline__ int blk_excl_scan(int v, int* sh, int& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const int a = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const int incl = sh[t];
  total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

// Number of entries of the ascending array a[0..n) that are < v.
template <typename K>
__device__ __forceinline__ int lower_count(const K* __restrict__ a, int n, K v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

enum { SORT_NONE = 0, SORT_INCREMENTAL = 1, SORT_FULL = 2 };

// Block-level exclusive scan of f(c) over [lo, hi) written as out[c] = base[c]
// + carry + prefix, for the chunked grid scans of k_sort_coop.
template <typename F>
__device__ __forceinline__ void chunk_scan(int lo, int hi, int carry, int* sh, F f,
                                           const int* __restrict__ base, int* __restrict__ out) {
  for (int b0 = lo; b0 < hi; b0 += kThreads) {
    const int c = b0 + threadIdx.x;
    const int v = c < hi ? f(c) : 0;
    int tot;
    const int ex = blk_excl_scan(v, sh, tot);
    if (c < hi) out[c] = (base ? base[c] : 0) + carry + ex;
    carry += tot;
  }
}

// Stable sort of the particles by cell in one cooperative launch (phases
// separated by grid barriers).  The mode is read from the flags, which no
// block writes before the first barrier, so all blocks agree:
//  * nothing to do: every block returns;
//  * SORT_INCREMENTAL (1..movcap recorded cell changes): cell counts change
//    by the recorded leavers/arrivals; one scan gives the new cell offsets;
//    every particle of an untouched cell moves by its cell's offset change,
//    every particle of a touched cell also by the leavers before it and the
//    arrivals with a smaller index (short per-cell lists), every mover goes
//    before the first particle of its new cell with a larger index.  One
//    streaming pass; the result is the stable order, in the other buffer;
//  * SORT_FULL: counting sort (histogram, scan, scatter, per-cell fix-up to
//    ascending index) into buffer 0.
// Also maintains element ids, keys, iperm and the buffer flag.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sort_coop(
    Grid2<T> g, const T* __restrict__ pb, int64_t ld, int n, int* __restrict__ eid,
    int* __restrict__ key, int* __restrict__ counts, int* __restrict__ fill,
    int* __restrict__ perm0, int* __restrict__ perm1, int* __restrict__ cs0,
    int* __restrict__ cs1, int* __restrict__ iperm, int* __restrict__ bs,
    int* __restrict__ flags, int* __restrict__ big, int* __restrict__ mov, int movcap,
    int* __restrict__ dL, int* __restrict__ dA, int* __restrict__ headL,
    int* __restrict__ headA, unsigned char* __restrict__ mflag) {
  const int nmov = flags[5];
  // flags[2]: a cell (or the outside bucket) was too big to order by index, so
  // the layout is not a stable sort and cannot be updated incrementally
  const int mode = (flags[0] || flags[6] || flags[2] || nmov > movcap)
                       ? SORT_FULL
                       : (nmov > 0 ? SORT_INCREMENTAL : SORT_NONE);
  if (mode == SORT_NONE) return;
  cg::grid_group grid = cg::this_grid();
  __shared__ int sh[kBigCellMax];
  const int G = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
  const int gt = b * kThreads + tid, gs = G * kThreads;
  const int nbk = g.ncells + 1;
  auto key_of = [&](int e) { return (e >= 0 && e < g.ncells) ? e : g.ncells; };

  if (mode == SORT_INCREMENTAL) {
    const int cur = flags[4];
    const int* __restrict__ po = cur ? perm1 : perm0;
    int* __restrict__ pn = cur ? perm0 : perm1;
    const int* __restrict__ co = cur ? cs1 : cs0;
    int* __restrict__ cn = cur ? cs0 : cs1;
    int* __restrict__ M = mov;
    for (int m = gt; m < nmov; m += gs) {
      const int p = M[kMovRec * m], ko = M[kMovRec * m + 1];
      const int kn = key_of(M[kMovRec * m + 2]);
      const int so = iperm[p];
      M[kMovRec * m + 3] = so;
      mflag[so] = 1;
      M[kMovRec * m + 4] = atomicExch(&headL[ko], m);
      M[kMovRec * m + 5] = atomicExch(&headA[kn], m);
      atomicAdd(&dL[ko], 1);
      atomicAdd(&dA[kn], 1);
    }
    grid.sync();
    // new offsets: cn = co + exclusive scan of (arrivals - leavers)
    const int chunk = (nbk + G - 1) / G;
    const int lo = min(b * chunk, nbk), hi = min(lo + chunk, nbk);
    {
      int sm = 0;
      for (int c = lo + tid; c < hi; c += kThreads) sm += dA[c] - dL[c];
      int tot;
      blk_excl_scan(sm, sh, tot);
      if (tid == 0) bs[b] = tot;
    }
    grid.sync();
    if (b == 0) {
      int carry = 0;
      for (int base = 0; base < G; base += kThreads) {
        const int k = base + tid;
        const int v = k < G ? bs[k] : 0;
        int tot;
        const int ex = blk_excl_scan(v, sh, tot);
        if (k < G) bs[k] = ex + carry;
        carry += tot;
      }
    }
    grid.sync();
    chunk_scan(lo, hi, bs[b], sh, [&](int c) { return dA[c] - dL[c]; }, co, cn);
    if (gt == 0) cn[nbk] = n;
    grid.sync();
    for (int s = gt; s < n; s += gs) {
      if (mflag[s]) continue;
      const int p = po[s];
      const int c = key[p];
      int pos;
      if (dL[c] == 0 && dA[c] == 0) {
        pos = s + (cn[c] - co[c]);
      } else {
        int before = 0, arr = 0;
        for (int m = headL[c]; m >= 0; m = M[kMovRec * m + 4]) before += M[kMovRec * m + 3] < s;
        for (int m = headA[c]; m >= 0; m = M[kMovRec * m + 5]) arr += M[kMovRec * m] < p;
        pos = cn[c] + (s - co[c]) - before + arr;
      }
      pn[pos] = p;
      iperm[p] = pos;
    }
    for (int m = gt; m < nmov; m += gs) {
      const int p = M[kMovRec * m];
      const int kn = key_of(M[kMovRec * m + 2]);
      int l = co[kn], h = co[kn + 1];
      const int first = l;
      while (l < h) {
        const int mid = (l + h) >> 1;
        if (po[mid] < p) l = mid + 1;
        else h = mid;
      }
      int lv = 0, arr = 0;
      for (int q = headL[kn]; q >= 0; q = M[kMovRec * q + 4]) lv += M[kMovRec * q] < p;
      for (int q = headA[kn]; q >= 0; q = M[kMovRec * q + 5]) arr += M[kMovRec * q] < p;
      const int pos = cn[kn] + (l - first) - lv + arr;
      pn[pos] = p;
      iperm[p] = pos;
    }
    grid.sync();
    for (int m = gt; m < nmov; m += gs) {
      const int p = M[kMovRec * m], ko = M[kMovRec * m + 1], e = M[kMovRec * m + 2];
      const int kn = key_of(e);
      eid[p] = e;
      key[p] = kn;
      mflag[M[kMovRec * m + 3]] = 0;
      dL[ko] = 0;
      dA[kn] = 0;
      headL[ko] = -1;
      headA[kn] = -1;
    }
    if (gt == 0) {
      flags[4] = 1 - cur;
      flags[5] = 0;
      flags[1] = 0;
    }
    return;
  }

  // ---- full counting sort into buffer 0 ----
  int* __restrict__ perm = perm0;
  int* __restrict__ cs = cs0;
  for (int c = gt; c < nbk; c += gs) {
    counts[c] = 0;
    fill[c] = 0;
  }
  if (gt == 0) {
    flags[3] = 0;
    flags[1] = 1;
  }
  grid.sync();
  if (gt == 0) flags[2] = 0;  // after the barrier: every block has read it
  for (int p = gt; p < n; p += gs) {
    const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
    eid[p] = e;
    const int k = (e >= 0 && e < g.ncells) ? e : g.ncells;
    key[p] = k;
    atomicAdd(&counts[k], 1);
  }
  grid.sync();
  const int chunk = (nbk + G - 1) / G;
  const int lo = min(b * chunk, nbk), hi = min(lo + chunk, nbk);
  {
    int sm = 0;
    for (int c = lo + tid; c < hi; c += kThreads) sm += counts[c];
    int tot;
    blk_excl_scan(sm, sh, tot);
    if (tid == 0) bs[b] = tot;
  }
  grid.sync();
  if (b == 0) {
    int carry = 0;
    for (int base = 0; base < G; base += kThreads) {
      const int k = base + tid;
      const int v = k < G ? bs[k] : 0;
      int tot;
      const int ex = blk_excl_scan(v, sh, tot);
      if (k < G) bs[k] = ex + carry;
      carry += tot;
    }
  }
  grid.sync();
  {
    int carry = bs[b];
    for (int base = lo; base < hi; base += kThreads) {
      const int c = base + tid;
      const int v = c < hi ? counts[c] : 0;
      int tot;
      const int ex = blk_excl_scan(v, sh, tot);
      if (c < hi) cs[c] = carry + ex;
      carry += tot;
    }
    if (gt == 0) cs[nbk] = n;
  }
  grid.sync();
  for (int p = gt; p < n; p += gs) {
    const int k = key[p];
    perm[cs[k] + atomicAdd(&fill[k], 1)] = p;
  }
  grid.sync();
  for (int c = gt; c <= g.ncells; c += gs) {  // cells and the outside bucket
    const int s0 = cs[c], s1 = cs[c + 1];
    if (s1 - s0 > kSmallCell) {
      big[atomicAdd(&flags[3], 1)] = c;
      continue;
    }
    for (int i = s0 + 1; i < s1; ++i) {
      const int v = perm[i];
      int j = i - 1;
      while (j >= s0 && perm[j] > v) {
        perm[j + 1] = perm[j];
        --j;
      }
      perm[j + 1] = v;
    }
  }
  grid.sync();
  const int nbig = flags[3];
  for (int q = b; q < nbig; q += G) {
    const int c = big[q];
    const int s0 = cs[c], m = cs[c + 1] - s0;
    if (m > kBigCellMax) {
      if (tid == 0) flags[2] = 1;
      continue;
    }
    for (int i = tid; i < m; i += kThreads) sh[i] = perm[s0 + i];
    __syncthreads();
    for (int i = tid; i < m; i += kThreads) {
      const int v = sh[i];
      int rk = 0;
      for (int j = 0; j < m; ++j) rk += sh[j] < v;
      perm[s0 + rk] = v;
    }
    __syncthreads();
  }
  grid.sync();
  int ident = 1;
  for (int s = gt; s < n; s += gs) {
    iperm[perm[s]] = s;
    if (perm[s] != s) ident = 0;
  }
  if (!ident) flags[1] = 0;
  grid.sync();
  if (gt == 0) {
    flags[4] = 0;
    flags[5] = 0;
    flags[6] = 0;
    flags[0] = 0;
  }
}

#undef PR
#undef NR

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

inline unsigned nblk(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int coop_grid_size(const void* kernel, int& out) {
  int dev = 0, ncu = 0, per = 0;
  DMPM_TRY_H(hipGetDevice(&dev));
  DMPM_TRY_H(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  DMPM_TRY_H(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0));
  if (per < 1) return fail(DMPM_EHIP, "cooperative sort kernel cannot be resident");
  out = std::min(ncu * std::min(per, 4), 4096);
  return DMPM_OK;
}

template <typename T, int TXN>
int tile_steps(dmpm_plan* P, const Grid2<T>& g, T* pb, int32_t* eid, T* nb, int scheme, int vu,
               const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
               hipStream_t s, T* alt, T* vkin, int* bs, int G) {
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  const bool split = usf && P->usf_split;
  TileShape tl;
  tl.ty = P->tile_ty > 0 ? P->tile_ty : 8;
  tl.ntx = (g.nx + 1 + TXN - 1) / TXN;
  tl.nty = (g.ny + 1 + tl.ty - 1) / tl.ty;
  const unsigned ntiles = (unsigned)(tl.ntx * tl.nty);
  Mat2 mat;
  mat.kind = P->d.material;
  mat.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) mat.de[k] = P->d.de[k];
  FnTable fns;
  fns.kind = P->fn_kind;
  fns.off = P->fn_off;
  fns.cnt = P->fn_cnt;
  fns.xs = P->fn_x;
  fns.fs = P->fn_f;
  const T gx = T(grav[0]), gy = T(grav[1]);
  const T tdt = T(dt);
  const SortRef sr = sort_ref(P);

  DB<T> cur_is_pb, cur_is_alt;
  for (int k = 0; k < kDB; ++k) {
    T* a = pb + (int64_t)db_row(k) * ld;
    T* w = alt ? alt + (int64_t)k * ld : a;
    cur_is_pb.cur[k] = a;
    cur_is_pb.nxt[k] = w;
    cur_is_alt.cur[k] = w;
    cur_is_alt.nxt[k] = a;
  }
  int parity = 0;
  for (int64_t step = step0; step < step0 + nsteps; ++step) {
    {
      ProfScopeT ps_(P, KC_SORT, s);
      void* args[] = {(void*)&g,        (void*)&pb,       (void*)&ld,        (void*)&n,
                      (void*)&eid,      (void*)&P->key,   (void*)&P->counts, (void*)&P->fill,
                      (void*)&P->perm,  (void*)&P->perm2, (void*)&P->cs,     (void*)&P->cs2,
                      (void*)&P->iperm, (void*)&bs,       (void*)&P->flags,  (void*)&P->big,
                      (void*)&P->mov,   (void*)&P->movcap, (void*)&P->dL,    (void*)&P->dA,
                      (void*)&P->headL, (void*)&P->headA, (void*)&P->mflag};
      DMPM_TRY(hip_check(hipLaunchCooperativeKernel((const void*)k_sort_coop<T>, dim3(G),
                                                    dim3(kThreads), args, 0, s),
                         "hipLaunchCooperativeKernel"));
    }
    const DB<T>& db = parity ? cur_is_alt : cur_is_pb;
    const T t = T(step) * tdt;
    {
      ProfScopeT ps_(P, KC_TILE_KIN, s);
      k_tile<T, MODE_KIN, TXN><<<ntiles, 2 * TXN, 0, s>>>(
          g, tl, pb, ld, sr, nb, nld, vkin, P->cidx, P->cmask, P->cval, db, mat, gx, gy, tdt, t,
          P->nfd, P->nf_off, P->nf_fn, P->nf_val, fns, temps);
      DMPM_TRY(launch_check("tile_kin"));
    }
    if (split) {
      ProfScopeT ps_(P, KC_STRESS, s);
      k_stress_tile<T><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, vkin, mat, tdt, temps);
      DMPM_TRY(launch_check("tile_stress"));
    }
    {
      ProfScopeT ps_(P, KC_TILE_FORCE, s);
      if (usf && !split)
        k_tile<T, MODE_FORCE_USF, TXN><<<ntiles, 2 * TXN, 0, s>>>(
            g, tl, pb, ld, sr, nb, nld, vkin, P->cidx, P->cmask, P->cval, db, mat, gx, gy, tdt,
            t, P->nfd, P->nf_off, P->nf_fn, P->nf_val, fns, temps);
      else
        k_tile<T, MODE_FORCE_USL, TXN><<<ntiles, 2 * TXN, 0, s>>>(
            g, tl, pb, ld, sr, nb, nld, vkin, P->cidx, P->cmask, P->cval, db, mat, gx, gy, tdt,
            t, P->nfd, P->nf_off, P->nf_fn, P->nf_val, fns, temps);
      DMPM_TRY(launch_check("tile_force"));
    }
    {
      ProfScopeT ps_(P, KC_TILE_G2P, s);
      if (usf)
        k_g2p_tile<T, false><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, nb, nld, P->flags, P->mov,
                                                          P->movcap, db, mat, tdt, vu, temps,
                                                          split ? 0 : 1);
      else
        k_g2p_tile<T, true><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, nb, nld, P->flags, P->mov,
                                                         P->movcap, db, mat, tdt, vu, temps, 0);
      DMPM_TRY(launch_check("tile_g2p"));
    }
    if (usf && !split) parity ^= 1;
  }
  if (parity) {  // the latest double-buffered fields are in the workspace
    for (int k = 0; k < kDB; ++k)
      DMPM_TRY(hip_check(hipMemcpyAsync(pb + (int64_t)db_row(k) * ld, alt + (int64_t)k * ld,
                                        (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, s),
                         "hipMemcpyAsync"));
  }
  return DMPM_OK;
}

template <typename T>
int tile_run_t(dmpm_plan* P, T* pb, int32_t* eid, T* nb, int scheme, int vu, const double* grav,
               double dt, int64_t step0, int64_t nsteps, int temps, hipStream_t s) {
  const Grid2<T> g = grid_of<T>(P);
  const int64_t ld = P->d.ld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  const bool dbuf = usf && !P->usf_split;
  T* alt = dbuf ? (T*)plan_workspace(P, 0, (size_t)kDB * ld * sizeof(T)) : nullptr;
  T* vkin = (T*)plan_workspace(P, 1, (size_t)2 * P->nnodes * sizeof(T));
  if ((dbuf && !alt) || !vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  int G = 0;
  DMPM_TRY(coop_grid_size((const void*)k_sort_coop<T>, G));
  int* bs = (int*)plan_workspace(P, 2, (size_t)G * sizeof(int));
  if (!bs) return fail(DMPM_EHIP, "workspace allocation failed");
  k_tile_check<T><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, eid, P->key, P->flags);
  DMPM_TRY(launch_check("tile_check"));
  if (P->tile_tx == 64)
    return tile_steps<T, 64>(P, g, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                             alt, vkin, bs, G);
  return tile_steps<T, 128>(P, g, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                            alt, vkin, bs, G);
}

}  // namespace

namespace dmpm {

bool tile_eligible(const dmpm_plan* P, int pfext) {
  return P->dim == 2 && !P->has_tractions && !pfext && P->n > 0;
}

int tile_run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
             const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
             hipStream_t s) {
  if (P->d.dtype == DMPM_F32)
    return tile_run_t<float>(P, (float*)pbuf, eid, (float*)nbuf, scheme, vu, grav, dt, step0,
                             nsteps, temps, s);
  return tile_run_t<double>(P, (double*)pbuf, eid, (double*)nbuf, scheme, vu, grav, dt, step0,
                            nsteps, temps, s);
}

}  // namespace dmpm
