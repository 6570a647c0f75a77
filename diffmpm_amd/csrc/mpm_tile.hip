// Fused 2D explicit step on node tiles (gfx950).
//
// One step is a re-sort and three (split USF: four) step launches:
//   k_sort_s1..s4   sort by cell: nothing when no particle changed cell, an
//                   O(N) update of the stable order from the recorded cell
//                   changes, a full counting sort when the layout is unknown
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

#include <algorithm>

#include "mpm_common.h"
#include "mpm_plan.h"

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

// Recorded cell changes: the lists and counters the incremental re-sort
// reads (all zero / -1 between sorts).
struct MoveLog {
  int* flags;
  int* mov;
  int* mrec;
  int* dL;
  int* dA;
  int* headL;
  int* headA;
  int* chunkD;
  int movcap;
  int chunk_shift;  // cells per scan chunk = 1 << chunk_shift
};

MoveLog move_log(dmpm_plan* P) {
  MoveLog L;
  L.flags = P->flags;
  L.mov = P->mov;
  L.mrec = P->mrec;
  L.dL = P->dL;
  L.dA = P->dA;
  L.headL = P->headL;
  L.headA = P->headA;
  L.chunkD = P->chunkD;
  L.movcap = P->movcap;
  L.chunk_shift = P->chunk_shift;
  return L;
}

__device__ __forceinline__ int key_of(int e, int ncells) {
  return (e >= 0 && e < ncells) ? e : ncells;
}

// Records that particle p moves from cell (key) ko to element id enew: the
// record joins the leaver list of ko and the arrival list of key(enew), and
// the per-cell and per-chunk counters.  Past movcap records, or a list longer
// than kListMax, a full re-sort is requested instead.
__device__ __forceinline__ void record_move(const MoveLog& L, int ncells, int p, int ko, int enew) {
  const int m = atomicAdd(&L.flags[5], 1);
  if (m >= L.movcap) {
    L.flags[6] = 1;
    return;
  }
  const int kn = key_of(enew, ncells);
  int* r = L.mov + (int64_t)kMovRec * m;
  r[0] = p;
  r[1] = ko;
  r[2] = enew;
  r[3] = atomicExch(&L.headL[ko], m);
  r[4] = atomicExch(&L.headA[kn], m);
  L.mrec[p] = m + 1;
  const int cl = atomicAdd(&L.dL[ko], 1);
  const int ca = atomicAdd(&L.dA[kn], 1);
  atomicSub(&L.chunkD[ko >> L.chunk_shift], 1);
  atomicAdd(&L.chunkD[kn >> L.chunk_shift], 1);
  if (cl >= kListMax || ca >= kListMax) L.flags[6] = 1;
}

// G2P (particle.py:208-247); USL also updates strain/volume/stress with the
// step-start natural coordinates and the updated nodal velocities
// (scheme.py:127-133).  Records the particles that change cell for the next
// re-sort.
template <typename T, bool USL>
__global__ __launch_bounds__(kThreads) void k_g2p_tile(Grid2<T> g, T* __restrict__ pb, int64_t ld,
                                                       int n, const T* __restrict__ nb,
                                                       int64_t nld, MoveLog mlog,
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
    record_move(mlog, g.ncells, p, g.ncells, 0);
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
  if (k1 != e) record_move(mlog, g.ncells, p, e, e1);
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

// Element ids from the positions at the start of a run.  A particle whose cell
// differs from its sort key (or, for a particle with a pending recorded cell
// change, from the recorded new cell) was moved outside the fused step: the
// next re-sort is then a full one.
template <typename T>
__global__ void k_tile_check(Grid2<T> g, const T* __restrict__ pb, int64_t ld, int n,
                             int* __restrict__ eid, const int* __restrict__ key,
                             int* __restrict__ flags, const int* __restrict__ mrec,
                             int* __restrict__ mov) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
  eid[p] = e;
  const int k = key_of(e, g.ncells);
  const int m = mrec[p] - 1;
  if (m >= 0) {
    int* r = mov + (int64_t)kMovRec * m;
    if (key_of(r[2], g.ncells) == k) r[2] = e;
    else flags[0] = 1;
  } else if (key[p] != k) {
    flags[0] = 1;
  }
}

// ---- re-sort by cell ----------------------------------------------------------
//
// Four ordinary launches per step (a cooperative launch and its grid barriers
// cost more than the whole incremental update).  Every workgroup of every
// stage derives the same mode from the flags, which only the last workgroup
// of the final stage of a mode rewrites ("last block" completion count):
//
//   mode     stage 1           stage 2             stage 3            stage 4
//   NONE     -                 -                   -                  -
//   INC      new cell offsets  relayout (stable)   keys, list reset   -
//   FULL     keys, histogram   offsets, big cells  scatter            order fix-up
//
// INC (the recorded cell changes, linked per old and new cell by G2P): the
// new offset of a cell is its old one plus the net arrivals of all cells
// before it; a particle of an untouched cell keeps its rank, one of a touched
// cell loses the leavers before it and gains the arrivals with a smaller
// index, an arriving particle goes before the first particle of its new cell
// with a larger index.  The result equals the stable counting sort of the new
// keys.  FULL is a counting sort with a per-cell fix-up to ascending index.
enum { SORT_NONE = 0, SORT_INCREMENTAL = 1, SORT_FULL = 2 };

struct SortArgs {
  int* eid;
  int* key;
  int* counts;
  int* fill;
  int* perm0;
  int* perm1;
  int* cs0;
  int* cs1;
  int* big;
  int* chunkC;
  MoveLog L;
  int n, ncells, nchunks;
};

__device__ __forceinline__ int sort_mode(const SortArgs& a) {
  const int* f = a.L.flags;
  const int nmov = f[5];
  if (f[0] || f[2] || f[6] || f[10] || nmov > a.L.movcap) return SORT_FULL;
  return nmov > 0 ? SORT_INCREMENTAL : SORT_NONE;
}

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

// out[c] = base(c) + (sum of chunk totals before this chunk) + exclusive
// prefix of val over the chunk, for the cells of chunk blockIdx.x (a run of
// chunk/256 consecutive cells per thread); the last chunk also writes
// out[nbk] = total.
template <typename V, typename B>
__device__ __forceinline__ void chunk_scan(int nbk, int nchunks, int shift,
                                           const int* __restrict__ tot, V val, B base,
                                           int* __restrict__ out, int total, int* sh) {
  const int b = blockIdx.x, tid = threadIdx.x;
  int pre = 0;
  for (int k = tid; k < b; k += kThreads) pre += tot[k];
  int red;
  blk_excl_scan(pre, sh, red);
  const int per = (1 << shift) / kThreads;
  const int c0 = (b << shift) + tid * per, c1 = min(c0 + per, nbk);
  int sum = 0;
  for (int c = c0; c < c1; ++c) sum += val(c);
  int bt;
  int ex = red + blk_excl_scan(sum, sh, bt);
  for (int c = c0; c < c1; ++c) {
    out[c] = base(c) + ex;
    ex += val(c);
  }
  if (b == nchunks - 1 && tid == 0) out[nbk] = total;
}

// Counts finished workgroups; true in the last one (after every other
// workgroup's writes are visible), which also resets the counter.
__device__ __forceinline__ bool last_block(int* counter) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(counter, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    __threadfence();
    if (threadIdx.x == 0) *counter = 0;
  }
  return s_last;
}

// Stage 1.  INC: new cell offsets (one workgroup per chunk).  FULL: element
// ids and keys from the positions, cell histogram, chunk totals (LDS
// histogram per workgroup), release of the pending change lists.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sort_s1(Grid2<T> g, const T* __restrict__ pb,
                                                      int64_t ld, SortArgs a) {
  const int mode = sort_mode(a);
  if (mode == SORT_NONE) return;
  __shared__ int sh[kMaxChunks];
  const int nbk = a.ncells + 1;
  if (mode == SORT_INCREMENTAL) {
    if ((int)blockIdx.x >= a.nchunks) return;
    const int cur = a.L.flags[4];
    const int* __restrict__ co = cur ? a.cs1 : a.cs0;
    int* __restrict__ cn = cur ? a.cs0 : a.cs1;
    const int* __restrict__ dL = a.L.dL;
    const int* __restrict__ dA = a.L.dA;
    chunk_scan(
        nbk, a.nchunks, a.L.chunk_shift, a.L.chunkD, [&](int c) { return dA[c] - dL[c]; },
        [&](int c) { return co[c]; }, cn, a.n, sh);
    return;
  }
  const int tid = threadIdx.x;
  const int gt = blockIdx.x * kThreads + tid, gs = gridDim.x * kThreads;
  for (int k = tid; k < a.nchunks; k += kThreads) sh[k] = 0;
  __syncthreads();
  for (int p = gt; p < a.n; p += gs) {
    const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
    a.eid[p] = e;
    const int k = key_of(e, a.ncells);
    a.key[p] = k;
    atomicAdd(&a.counts[k], 1);
    atomicAdd(&sh[k >> a.L.chunk_shift], 1);
  }
  __syncthreads();
  for (int k = tid; k < a.nchunks; k += kThreads)
    if (sh[k]) atomicAdd(&a.chunkC[k], sh[k]);
  const int nrec = min(a.L.flags[5], a.L.movcap);
  for (int m = gt; m < nrec; m += gs) {
    const int* r = a.L.mov + (int64_t)kMovRec * m;
    const int ko = r[1], kn = key_of(r[2], a.ncells);
    a.L.mrec[r[0]] = 0;
    a.L.dL[ko] = 0;
    a.L.dA[kn] = 0;
    a.L.headL[ko] = -1;
    a.L.headA[kn] = -1;
  }
  for (int k = gt; k < a.nchunks; k += gs) a.L.chunkD[k] = 0;
  if (gt == 0) a.L.flags[1] = 1;
}

// Stage 2.  INC: relayout of every slot into the other buffer, placement of
// the recorded movers, their new element ids.  FULL: cell offsets (one
// workgroup per chunk) and the queue of cells too big for one thread.
__global__ __launch_bounds__(kThreads) void k_sort_s2(SortArgs a) {
  const int mode = sort_mode(a);
  if (mode == SORT_NONE) return;
  __shared__ int sh[kThreads];
  const int nbk = a.ncells + 1;
  if (mode == SORT_FULL) {
    if ((int)blockIdx.x >= a.nchunks) return;
    const int* __restrict__ counts = a.counts;
    chunk_scan(
        nbk, a.nchunks, a.L.chunk_shift, a.chunkC, [&](int c) { return counts[c]; },
        [](int) { return 0; }, a.cs0, a.n, sh);
    const int c0 = blockIdx.x << a.L.chunk_shift;
    for (int c = c0 + threadIdx.x; c < min(c0 + (1 << a.L.chunk_shift), nbk); c += kThreads)
      if (counts[c] > kSmallCell) a.big[atomicAdd(&a.L.flags[3], 1)] = c;
    return;
  }
  const int cur = a.L.flags[4];
  const int* __restrict__ po = cur ? a.perm1 : a.perm0;
  int* __restrict__ pn = cur ? a.perm0 : a.perm1;
  const int* __restrict__ co = cur ? a.cs1 : a.cs0;
  const int* __restrict__ cn = cur ? a.cs0 : a.cs1;
  const int* __restrict__ M = a.L.mov;
  const int* __restrict__ dL = a.L.dL;
  const int* __restrict__ dA = a.L.dA;
  const int gt = blockIdx.x * kThreads + threadIdx.x, gs = gridDim.x * kThreads;
  for (int s = gt; s < a.n; s += gs) {
    const int p = po[s];
    const int c = a.key[p];
    int pos;
    if ((dL[c] | dA[c]) == 0) {
      pos = s + (cn[c] - co[c]);
    } else {
      if (a.L.mrec[p]) continue;  // placed by the mover loop below
      int before = 0, arr = 0;
      for (int m = a.L.headL[c]; m >= 0; m = M[kMovRec * m + 3])
        before += M[kMovRec * m] < p;  // old segment ascends in particle index
      for (int m = a.L.headA[c]; m >= 0; m = M[kMovRec * m + 4]) arr += M[kMovRec * m] < p;
      pos = cn[c] + (s - co[c]) - before + arr;
    }
    pn[pos] = p;
  }
  const int nmov = a.L.flags[5];
  for (int m = gt; m < nmov; m += gs) {
    const int p = M[kMovRec * m], e = M[kMovRec * m + 2];
    const int kn = key_of(e, a.ncells);
    int l = co[kn], h = co[kn + 1];
    const int first = l;
    while (l < h) {
      const int mid = (l + h) >> 1;
      if (po[mid] < p) l = mid + 1;
      else h = mid;
    }
    int lv = 0, arr = 0;
    for (int q = a.L.headL[kn]; q >= 0; q = M[kMovRec * q + 3]) lv += M[kMovRec * q] < p;
    for (int q = a.L.headA[kn]; q >= 0; q = M[kMovRec * q + 4]) arr += M[kMovRec * q] < p;
    pn[cn[kn] + (l - first) - lv + arr] = p;
    a.eid[p] = e;
  }
}

// Stage 3.  INC: new keys of the movers, release of the change lists.
// FULL: scatter into buffer 0.
__global__ __launch_bounds__(kThreads) void k_sort_s3(SortArgs a) {
  const int mode = sort_mode(a);
  if (mode == SORT_NONE) return;
  const int gt = blockIdx.x * kThreads + threadIdx.x, gs = gridDim.x * kThreads;
  if (mode == SORT_FULL) {
    for (int p = gt; p < a.n; p += gs) {
      const int k = a.key[p];
      a.perm0[a.cs0[k] + atomicAdd(&a.fill[k], 1)] = p;
    }
    return;
  }
  const int nmov = a.L.flags[5];
  for (int m = gt; m < nmov; m += gs) {
    const int* r = a.L.mov + (int64_t)kMovRec * m;
    const int p = r[0], ko = r[1], kn = key_of(r[2], a.ncells);
    a.key[p] = kn;
    a.L.mrec[p] = 0;
    a.L.dL[ko] = 0;
    a.L.dA[kn] = 0;
    a.L.headL[ko] = -1;
    a.L.headA[kn] = -1;
    a.L.chunkD[ko >> a.L.chunk_shift] = 0;
    a.L.chunkD[kn >> a.L.chunk_shift] = 0;
  }
}

// Stage 4.  INC: the first workgroup switches buffers and clears the change
// count (any other workgroup then reads NONE, and INC leaves it no work).
// FULL: ascending particle index within each cell — cells of up to
// kSmallCell particles by one thread (insertion sort), bigger ones by one
// workgroup (rank sort in LDS up to kBigCellMax; bigger cells keep the
// scatter order and mark the layout unsorted); restores the zero counters;
// the last workgroup publishes the new state.
__global__ __launch_bounds__(kThreads) void k_sort_s4(SortArgs a) {
  const int mode = sort_mode(a);
  if (mode == SORT_NONE) return;
  if (mode == SORT_INCREMENTAL) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      int* f = a.L.flags;
      f[4] = 1 - f[4];
      f[5] = 0;
      f[1] = 0;
    }
    return;
  }
  __shared__ int sh[kBigCellMax];
  int* __restrict__ perm = a.perm0;
  const int* __restrict__ cs = a.cs0;
  const int nbig = a.L.flags[3];
  for (int q = blockIdx.x; q < nbig; q += gridDim.x) {
    const int c = a.big[q];
    const int s0 = cs[c], m = cs[c + 1] - s0;
    if (m > kBigCellMax) {
      if (threadIdx.x == 0) a.L.flags[7] = 1;
      if (threadIdx.x == 0) a.L.flags[1] = 0;
      continue;
    }
    for (int i = threadIdx.x; i < m; i += kThreads) sh[i] = perm[s0 + i];
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kThreads) {
      const int v = sh[i];
      int rk = 0;
      for (int j = 0; j < m; ++j) rk += sh[j] < v;
      perm[s0 + rk] = v;
      if (v != s0 + rk) a.L.flags[1] = 0;
    }
    __syncthreads();
  }
  const int gt = blockIdx.x * kThreads + threadIdx.x, gs = gridDim.x * kThreads;
  for (int c = gt; c <= a.ncells; c += gs) {  // cells and the outside bucket
    a.counts[c] = 0;
    a.fill[c] = 0;
    const int s0 = cs[c], s1 = cs[c + 1];
    if (s1 - s0 > kSmallCell) continue;
    bool ident = true;
    for (int i = s0 + 1; i < s1; ++i) {
      const int v = perm[i];
      int j = i - 1;
      while (j >= s0 && perm[j] > v) {
        perm[j + 1] = perm[j];
        --j;
      }
      perm[j + 1] = v;
    }
    for (int i = s0; i < s1; ++i) ident = ident && perm[i] == i;
    if (!ident) a.L.flags[1] = 0;
  }
  if (last_block(&a.L.flags[8])) {
    for (int k = threadIdx.x; k < a.nchunks; k += kThreads) a.chunkC[k] = 0;
    if (threadIdx.x == 0) {
      int* f = a.L.flags;
      f[2] = f[7];
      f[7] = 0;
      f[3] = 0;
      f[4] = 0;
      f[5] = 0;
      f[6] = 0;
      f[10] = 0;
      f[0] = 0;
    }
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

constexpr int kSortGrid = 2048;  // workgroups of the grid-stride re-sort stages

template <typename T>
int enqueue_resort(dmpm_plan* P, const Grid2<T>& g, const T* pb, int32_t* eid, hipStream_t s) {
  SortArgs a;
  a.eid = eid;
  a.key = P->key;
  a.counts = P->counts;
  a.fill = P->fill;
  a.perm0 = P->perm;
  a.perm1 = P->perm2;
  a.cs0 = P->cs;
  a.cs1 = P->cs2;
  a.big = P->big;
  a.chunkC = P->chunkC;
  a.L = move_log(P);
  a.n = (int)P->n;
  a.ncells = P->ncells;
  a.nchunks = P->nchunks;
  const int g12 = std::max(P->nchunks, kSortGrid);
  k_sort_s1<T><<<g12, kThreads, 0, s>>>(g, pb, P->d.ld, a);
  k_sort_s2<<<g12, kThreads, 0, s>>>(a);
  k_sort_s3<<<kSortGrid, kThreads, 0, s>>>(a);
  k_sort_s4<<<kSortGrid / 4, kThreads, 0, s>>>(a);
  return launch_check("resort");
}

template <typename T, int TXN>
int tile_steps(dmpm_plan* P, const Grid2<T>& g, T* pb, int32_t* eid, T* nb, int scheme, int vu,
               const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
               hipStream_t s, T* alt, T* vkin) {
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
      DMPM_TRY(enqueue_resort<T>(P, g, pb, eid, s));
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
        k_g2p_tile<T, false><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, nb, nld, move_log(P), db,
                                                          mat, tdt, vu, temps, split ? 0 : 1);
      else
        k_g2p_tile<T, true><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, nb, nld, move_log(P), db,
                                                         mat, tdt, vu, temps, 0);
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
  k_tile_check<T><<<nblk(n), kThreads, 0, s>>>(g, pb, ld, n, eid, P->key, P->flags, P->mrec,
                                               P->mov);
  DMPM_TRY(launch_check("tile_check"));
  if (P->tile_tx == 64)
    return tile_steps<T, 64>(P, g, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                             alt, vkin);
  return tile_steps<T, 128>(P, g, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                            alt, vkin);
}

}  // namespace

namespace dmpm {

bool tile_eligible(const dmpm_plan* P, int pfext) {
  return P->dim == 2 && !P->has_tractions && !pfext && P->n > 0 && P->nchunks <= kMaxChunks;
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
