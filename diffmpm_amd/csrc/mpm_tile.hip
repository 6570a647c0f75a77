// Fused 2D explicit step (gfx950): re-sort by cell, P2G as a cell pass and a
// node pass, G2P.
//
// One step is:
//   k_sort_s1..s4          sort by cell: nothing when no particle changed
//                          cell, an O(N) update of the stable order from the
//                          recorded cell changes, a full counting sort when
//                          the layout is unknown
//   k_cell_kin/k_node_kin  P2G of mass and momentum, nodal velocity,
//                          constraints
//   k_stress_tile          USF strain/volume/stress update, streamed (or, as
//                          an option, inside k_cell_force)
//   k_cell_force/k_node_force  P2G of body and internal forces, nodal forces,
//                          nodal acceleration/velocity update, constraints
//   k_g2p_tile             G2P velocity/position (USL: + strain/volume/stress
//                          update), records cell changes for the next sort
//
// Nodal sums are formed without atomics and in a fixed order (per-cell sums
// in ascending particle index, then the four cells of a node in the order
// (i-1,j-1), (i,j-1), (i-1,j), (i,j)); the per-phase kernels use the same
// order and the same per-term arithmetic, so both paths are bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mpm_common.h"
#include "mpm_plan.h"
#include "mpm_step.h"

using namespace dmpm;

namespace {

constexpr int kSmallCell = 64;
constexpr int kBigCellMax = 8192;

#define PR(r) (pb + (int64_t)(r) * ld)
#define NR(r) (nb + (int64_t)(r) * nld)

// ---- P2G: cell pass + node pass ------------------------------------------------
//
// Cell pass: a workgroup of kThreads threads owns CB consecutive cells and the
// particles the sort put in them (slots cs[c0] .. cs[c0+CB], ~4 per cell).
// The slots are loaded in rounds of kSlots (one round at <= kSlots/CB
// particles per cell), kSlots/kThreads independent coalesced slot runs per
// thread, all issued before any is used.  Each particle's inputs (natural
// coordinates and the quantities it deposits) go to LDS; then thread t < CB
// walks the slots of cell c0+t in order (ascending particle index), forms the
// terms of the four corner nodes and adds them to running per-corner sums,
// written to a plan workspace ("partials", one row per corner and quantity).
// In USF the slot threads also update strain, volume and stress in place
// from the kinematic nodal velocities of the cell, staged in LDS.
//
// Node pass: node (i, j) adds the partials of cells (i-1,j-1), (i,j-1),
// (i-1,j), (i,j) in that order to a zero start and runs the nodal epilogue.
// The per-phase kernels accumulate the same terms in the same order, so both
// paths give bit-identical nodal sums without atomics.
enum { NQ_KIN = 3, NQ_FORCE = 4 };  // quantities per corner
constexpr int kSlots = 1024;        // slots per load round of a cell-pass workgroup
constexpr int kPer = kSlots / kThreads;

// KIN cell pass: terms m*N_k and N_k*(m*v) of every corner k (element.py:140,
// 172); with temps also the natural coordinates (particle.py:180-206).
template <typename T, int CB>
__global__ __launch_bounds__(kThreads) void k_cell_kin(Mesh2<T> M, T* __restrict__ pb,
                                                       int64_t ld, SortRef sr,
                                                       T* __restrict__ part, int temps) {
  constexpr int NIN = 5;  // xi, eta, m, m*vx, m*vy
  __shared__ T st[NIN][kSlots];
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  const int ident = sr.flags[1];
  const int c0 = blockIdx.x * CB, t = threadIdx.x;
  const int cend = min(c0 + CB, M.ncells);
  const int mycell = c0 + t;
  const bool celler = t < CB && mycell < cend;
  const int ms0 = celler ? cs[mycell] : 0, ms1 = celler ? cs[mycell + 1] : 0;
  const int s0 = cs[c0], s1 = cs[cend];
  T acc[4 * NQ_KIN];
#pragma unroll
  for (int v = 0; v < 4 * NQ_KIN; ++v) acc[v] = T(0);
  for (int k0 = s0; k0 < s1; k0 += kSlots) {
    const int kend = min(k0 + kSlots, s1);
    int pp[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int s = k0 + t + j * kThreads;
      pp[j] = s < kend ? (ident ? s : perm[s]) : -1;
    }
    T x[kPer], y[kPer], m[kPer], vx[kPer], vy[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (pp[j] < 0) continue;
      const int p = pp[j];
      x[j] = PR(DMPM_P_X)[p];
      y[j] = PR(DMPM_P_X + 1)[p];
      m[j] = PR(DMPM_P_MASS)[p];
      vx[j] = PR(DMPM_P_V)[p];
      vy[j] = PR(DMPM_P_V + 1)[p];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (pp[j] < 0) continue;
      const int q = t + j * kThreads;
      const Grid2<T> g = M.grid(M.sim_p(pp[j]));
      const int e = element_id<T, 2>(g.geo(), x[j], y[j]);
      const int I = e % g.nx, J = e / g.nx;
      const T xi = natural(x[j], g.X(I), g.X(I + 1));
      const T eta = natural(y[j], g.Y(J), g.Y(J + 1));
      st[0][q] = xi;
      st[1][q] = eta;
      st[2][q] = m[j];
      st[3][q] = m[j] * vx[j];
      st[4][q] = m[j] * vy[j];
      if (temps) {
        PR(DMPM_P_XI)[pp[j]] = xi;
        PR(DMPM_P_XI + 1)[pp[j]] = eta;
      }
    }
    __syncthreads();
    if (celler) {
      const int lo = max(ms0, k0) - k0, hi = min(ms1, kend) - k0;
      for (int q = lo; q < hi; ++q) {
        T N[4];
        q4_N4(st[0][q], st[1][q], N);
        const T mq = st[2][q], mvx = st[3][q], mvy = st[4][q];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          acc[NQ_KIN * k + 0] = acc[NQ_KIN * k + 0] + mq * N[k];
          acc[NQ_KIN * k + 1] = acc[NQ_KIN * k + 1] + N[k] * mvx;
          acc[NQ_KIN * k + 2] = acc[NQ_KIN * k + 2] + N[k] * mvy;
        }
      }
    }
    __syncthreads();
  }
  if (celler)
#pragma unroll
    for (int v = 0; v < 4 * NQ_KIN; ++v) part[(int64_t)v * M.ncells + mycell] = acc[v];
}

// FORCE cell pass: terms N_k*(m*g) (element.py:254-256) and
// -V*(dN_k/dx . sigma) (element.py:868-878) of every corner k.  USF first
// updates strain, volume and stress in place (particle.py:249-327).
template <typename T, int CB, bool USF>
__global__ __launch_bounds__(kThreads) void k_cell_force(Mesh2<T> M, T* __restrict__ pb,
                                                         int64_t ld, SortRef sr,
                                                         T* __restrict__ part,
                                                         const T* __restrict__ vkin, Mat2 mat,
                                                         T gx, T gy, T dt, int temps) {
  constexpr int NIN = 7;  // xi, eta, m, -V, sxx, syy, sxy
  __shared__ T st[NIN][kSlots];
  __shared__ T sv[USF ? 8 : 1][USF ? CB : 1];
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  const int ident = sr.flags[1];
  const int c0 = blockIdx.x * CB, t = threadIdx.x;
  const int cend = min(c0 + CB, M.ncells);
  const int mycell = c0 + t;
  const bool celler = t < CB && mycell < cend;
  const int ms0 = celler ? cs[mycell] : 0, ms1 = celler ? cs[mycell + 1] : 0;
  const int s0 = cs[c0], s1 = cs[cend];
  const int csim = celler ? M.sim_c(mycell) : 0;
  const Grid2<T> gc = M.grid(csim);
  const int lcell = mycell - M.coff(csim);
  if (USF && celler) {
    int ids[4];
    cell_nodes(gc.nx, lcell % gc.nx, lcell / gc.nx, ids, M.noff(csim));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      sv[k][t] = vkin[ids[k]];
      sv[4 + k][t] = vkin[M.nnodes + ids[k]];
    }
  }
  T acc[4 * NQ_FORCE];
#pragma unroll
  for (int v = 0; v < 4 * NQ_FORCE; ++v) acc[v] = T(0);
  for (int k0 = s0; k0 < s1; k0 += kSlots) {
    const int kend = min(k0 + kSlots, s1);
    int pp[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int s = k0 + t + j * kThreads;
      pp[j] = s < kend ? (ident ? s : perm[s]) : -1;
    }
    if (USF && k0 == s0) __syncthreads();  // nodal velocities staged
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (pp[j] < 0) continue;
      const int p = pp[j];
      const int q = t + j * kThreads;
      const T x = PR(DMPM_P_X)[p], y = PR(DMPM_P_X + 1)[p];
      const int ps = M.sim_p(p);
      const Grid2<T> g = M.grid(ps);
      const int e = element_id<T, 2>(g.geo(), x, y);
      const int I = e % g.nx, J = e / g.nx;
      const T X0 = g.X(I), X1 = g.X(I + 1), Y0 = g.Y(J), Y1 = g.Y(J + 1);
      const T xi = natural(x, X0, X1);
      const T eta = natural(y, Y0, Y1);
      T V, sxx, syy, sxy;
      if (USF) {
        T ix, iy;
        q4_invJ(xi, eta, X0, X1, Y0, Y1, ix, iy);
        const int lc = min(max(M.coff(ps) + e - c0, 0), CB - 1);
        const T vxn[4] = {sv[0][lc], sv[1][lc], sv[2][lc], sv[3][lc]};
        const T vyn[4] = {sv[4][lc], sv[5][lc], sv[6][lc], sv[7][lc]};
        T rr[6], rc[6], d[6];
        rate2s(ix, iy, xi, eta, vxn, vyn, rr);
        rate2s(ix, iy, T(0), T(0), vxn, vyn, rc);
        for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
        const T dvol = dt * (rc[0] + rc[1]);
        T ds[4];
        stress_inc(M.mat(mat, ps), mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
        const T f = T(1) + dvol;
        PR(DMPM_P_STRAIN + 0)[p] = PR(DMPM_P_STRAIN + 0)[p] + d[0];
        PR(DMPM_P_STRAIN + 1)[p] = PR(DMPM_P_STRAIN + 1)[p] + d[1];
        PR(DMPM_P_STRAIN + 3)[p] = PR(DMPM_P_STRAIN + 3)[p] + d[3];
        sxx = PR(DMPM_P_STRESS + 0)[p] + ds[0];
        syy = PR(DMPM_P_STRESS + 1)[p] + ds[1];
        PR(DMPM_P_STRESS + 2)[p] = PR(DMPM_P_STRESS + 2)[p] + ds[2];
        sxy = PR(DMPM_P_STRESS + 3)[p] + ds[3];
        PR(DMPM_P_STRESS + 0)[p] = sxx;
        PR(DMPM_P_STRESS + 1)[p] = syy;
        PR(DMPM_P_STRESS + 3)[p] = sxy;
        V = PR(DMPM_P_VOLUME)[p] * f;
        PR(DMPM_P_VOLUME)[p] = V;
        PR(DMPM_P_DENSITY)[p] = PR(DMPM_P_DENSITY)[p] / f;
        PR(DMPM_P_VOLSTRAIN_C)[p] = PR(DMPM_P_VOLSTRAIN_C)[p] + dvol;
        if (temps) {
          for (int k = 0; k < 6; ++k) {
            PR(DMPM_P_STRAIN_RATE + k)[p] = rr[k];
            PR(DMPM_P_DSTRAIN + k)[p] = d[k];
          }
          PR(DMPM_P_DVOLSTRAIN)[p] = dvol;
        }
      } else {
        V = PR(DMPM_P_VOLUME)[p];
        sxx = PR(DMPM_P_STRESS + 0)[p];
        syy = PR(DMPM_P_STRESS + 1)[p];
        sxy = PR(DMPM_P_STRESS + 3)[p];
      }
      st[0][q] = xi;
      st[1][q] = eta;
      st[2][q] = PR(DMPM_P_MASS)[p];
      st[3][q] = -V;
      st[4][q] = sxx;
      st[5][q] = syy;
      st[6][q] = sxy;
    }
    __syncthreads();
    if (celler) {
      const int I = lcell % gc.nx, J = lcell / gc.nx;
      const T X0 = gc.X(I), X1 = gc.X(I + 1), Y0 = gc.Y(J), Y1 = gc.Y(J + 1);
      const int lo = max(ms0, k0) - k0, hi = min(ms1, kend) - k0;
      for (int q = lo; q < hi; ++q) {
        const T xi = st[0][q], eta = st[1][q];
        const T bx = st[2][q] * gx, by = st[2][q] * gy;
        const T nV = st[3][q], sxx = st[4][q], syy = st[5][q], sxy = st[6][q];
        T ix, iy;
        q4_invJ(xi, eta, X0, X1, Y0, Y1, ix, iy);
        T N[4];
        q4_N4(xi, eta, N);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          T dx, dy;
          q4_dNdx(k, xi, eta, ix, iy, dx, dy);
          acc[NQ_FORCE * k + 0] = acc[NQ_FORCE * k + 0] + N[k] * bx;
          acc[NQ_FORCE * k + 1] = acc[NQ_FORCE * k + 1] + N[k] * by;
          acc[NQ_FORCE * k + 2] = acc[NQ_FORCE * k + 2] + nV * (dx * sxx + dy * sxy);
          acc[NQ_FORCE * k + 3] = acc[NQ_FORCE * k + 3] + nV * (dy * syy + dx * sxy);
        }
      }
    }
    __syncthreads();
  }
  if (celler)
#pragma unroll
    for (int v = 0; v < 4 * NQ_FORCE; ++v) part[(int64_t)v * M.ncells + mycell] = acc[v];
}

// KIN node pass: nodal mass, momentum (snapped), velocity (element.py:122-202)
// and velocity constraints (constraint.py:29-44); the kinematic velocity also
// goes to vkin for the force phase.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_node_kin(Mesh2<T> M, const T* __restrict__ part,
                                                       T* __restrict__ nb, int64_t nld,
                                                       T* __restrict__ vkin,
                                                       const int* __restrict__ cidx,
                                                       const int* __restrict__ cmask,
                                                       const double* __restrict__ cval) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= M.nnodes) return;
  const int sm = M.sim_n(nd);
  const Grid2<T> g = M.grid(sm);
  const int nl = nd - M.noff(sm);
  const int i = nl % (g.nx + 1), j = nl / (g.nx + 1);
  T acc[NQ_KIN];
  node_gather<T, NQ_KIN>(g, M.coff(sm), M.ncells, part, i, j, acc);
  const int cr = cidx ? cidx[nd] : -1;
  const int mask = cr >= 0 ? cmask[cr] : 0;
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
    vkin[(int64_t)d * M.nnodes + nd] = v;
  }
}

// FORCE node pass: external force (body + concentrated nodal forces at t,
// element.py:204-284), internal force, acceleration/velocity/momentum update
// (element.py:305-329) and constraints.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_node_force(
    Mesh2<T> M, const T* __restrict__ part, T* __restrict__ nb, int64_t nld,
    const T* __restrict__ vkin, const int* __restrict__ cidx, const int* __restrict__ cmask,
    const double* __restrict__ cval, T dt, T t, const int* __restrict__ nfd,
    const int* __restrict__ nf_off, const int* __restrict__ nf_fn,
    const double* __restrict__ nf_val, FnTable fns) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= M.nnodes) return;
  const int sm = M.sim_n(nd);
  const Grid2<T> g = M.grid(sm);
  const int nl = nd - M.noff(sm);
  const int i = nl % (g.nx + 1), j = nl / (g.nx + 1);
  T acc[NQ_FORCE];
  node_gather<T, NQ_FORCE>(g, M.coff(sm), M.ncells, part, i, j, acc);
  const int cr = cidx ? cidx[nd] : -1;
  const int mask = cr >= 0 ? cmask[cr] : 0;
  const T m = NR(DMPM_N_MASS)[nd];
  for (int d = 0; d < 2; ++d) {
    T fe = acc[d];
    const int qg = nfd ? nfd[(int64_t)d * M.nnodes + nd] : -1;
    if (qg >= 0)
      for (int k = nf_off[qg]; k < nf_off[qg + 1]; ++k)
        fe = fe + fn_value<T>(fns, nf_fn[k], t) * T(nf_val[k]);
    const T fi = acc[2 + d];
    NR(DMPM_N_FEXT + d)[nd] = fe;
    NR(DMPM_N_FINT + d)[nd] = fi;
    const T F = (fi + fe) + NR(DMPM_N_FDAMP + d)[nd];
    T a = nan_to_num(F / m);
    T v = vkin[(int64_t)d * M.nnodes + nd] + a * dt;
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

// Records that particle p moves from cell (key) ko to cell kn, element id
// enew of its own mesh: the record joins the leaver list of ko and the
// arrival list of kn, and the per-cell and per-chunk counters.  Past movcap
// records, or a list longer than kListMax, a full re-sort is requested
// instead.
__device__ __forceinline__ void record_move(const MoveLog& L, int p, int ko, int enew, int kn) {
  const int m = atomicAdd(&L.flags[5], 1);
  if (m >= L.movcap) {
    L.flags[6] = 1;
    return;
  }
  int* r = L.mov + (int64_t)kMovRec * m;
  r[0] = p;
  r[1] = ko;
  r[2] = enew;
  r[3] = atomicExch(&L.headL[ko], m);
  r[4] = atomicExch(&L.headA[kn], m);
  r[5] = kn;
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
__global__ __launch_bounds__(kThreads) void k_g2p_tile(Mesh2<T> M, T* __restrict__ pb, int64_t ld,
                                                       int n, const T* __restrict__ nb,
                                                       int64_t nld, MoveLog mlog, Mat2 mat,
                                                       T dt, int vu, int temps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ps = M.sim_p(p);
  const Grid2<T> g = M.grid(ps);
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
    for (int k = 0; k < 6; ++k) {
      PR(DMPM_P_STRAIN + k)[p] = nan;
      PR(DMPM_P_STRESS + k)[p] = nan;
    }
    PR(DMPM_P_VOLUME)[p] = nan;
    PR(DMPM_P_DENSITY)[p] = nan;
    PR(DMPM_P_VOLSTRAIN_C)[p] = nan;
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
    record_move(mlog, p, M.ncells, 0, M.key(ps, g, 0));
    return;
  }
  const int I = e % g.nx, J = e / g.nx;
  const T xi = natural(x, g.X(I), g.X(I + 1));
  const T eta = natural(y, g.Y(J), g.Y(J + 1));
  int ids[4];
  cell_nodes(g.nx, I, J, ids, M.noff(ps));
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
    stress_inc(M.mat(mat, ps), mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
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
  const int k1 = M.key(ps, g, e1);
  const int k0 = M.key(ps, g, e);
  if (k1 != k0) record_move(mlog, p, k0, e1, k1);
}

// USF strain / volume / stress update (particle.py:249-327) streamed over the
// particles in storage order, in place, from the kinematic nodal velocities;
// used before k_tile<FORCE_USL> when USF runs split (DMPM_OPT_USF_SPLIT).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_stress_tile(Mesh2<T> M, T* __restrict__ pb,
                                                          int64_t ld, int n,
                                                          const T* __restrict__ vkin, Mat2 mat,
                                                          T dt, int temps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ps = M.sim_p(p);
  const Grid2<T> g = M.grid(ps);
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
  int ids[4];
  cell_nodes(g.nx, I, J, ids, M.noff(ps));
  T vx[4], vy[4];
  for (int k = 0; k < 4; ++k) {
    vx[k] = vkin[ids[k]];
    vy[k] = vkin[M.nnodes + ids[k]];
  }
  T rr[6], rc[6], d[6];
  rate2(g, I, J, xi, eta, vx, vy, rr);
  rate2(g, I, J, T(0), T(0), vx, vy, rc);
  for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
  const T dvol = dt * (rc[0] + rc[1]);
  T ds[4];
  stress_inc(M.mat(mat, ps), mat.per_E ? PR(DMPM_P_YOUNGS)[p] : T(0), d, ds);
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
// differs from its sort key was moved outside the fused step (positions set
// by the caller, a recomputed step): the change is recorded like one from
// G2P, for the incremental re-sort.  A particle with a pending recorded
// change whose recorded cell no longer holds, or one never sorted, makes the
// next re-sort a full one.
template <typename T>
__global__ void k_tile_check(Mesh2<T> M, const T* __restrict__ pb, int64_t ld, int n,
                             int* __restrict__ eid, const int* __restrict__ key, MoveLog L) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int ps = M.sim_p(p);
  const Grid2<T> g = M.grid(ps);
  const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
  const int k = M.key(ps, g, e);
  const int m = L.mrec[p] - 1;
  if (m >= 0) {
    int* r = L.mov + (int64_t)kMovRec * m;
    if (r[5] == k) r[2] = e;
    else L.flags[0] = 1;
    return;
  }
  const int ko = key[p];
  if (ko == k) {
    eid[p] = e;
  } else if (ko < 0 || ko > M.ncells) {
    L.flags[0] = 1;
  } else {
    record_move(L, p, ko, e, k);
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
__global__ __launch_bounds__(kThreads) void k_sort_s1(Mesh2<T> Mg, const T* __restrict__ pb,
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
    const int ps = Mg.sim_p(p);
    const Grid2<T> g = Mg.grid(ps);
    const int e = element_id<T, 2>(g.geo(), PR(DMPM_P_X)[p], PR(DMPM_P_X + 1)[p]);
    a.eid[p] = e;
    const int k = Mg.key(ps, g, e);
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
    const int ko = r[1], kn = r[5];
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
    const int kn = M[kMovRec * m + 5];
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
    const int p = r[0], ko = r[1], kn = r[5];
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


constexpr int kSortGrid = 2048;  // workgroups of the grid-stride re-sort stages

template <typename T>
int enqueue_resort(dmpm_plan* P, const Mesh2<T>& M, const T* pb, int32_t* eid, hipStream_t s) {
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
  k_sort_s1<T><<<g12, kThreads, 0, s>>>(M, pb, P->d.ld, a);
  k_sort_s2<<<g12, kThreads, 0, s>>>(a);
  k_sort_s3<<<kSortGrid, kThreads, 0, s>>>(a);
  k_sort_s4<<<kSortGrid / 4, kThreads, 0, s>>>(a);
  return launch_check("resort");
}

template <typename T, int CB>
int tile_steps(dmpm_plan* P, const Mesh2<T>& M, T* pb, int32_t* eid, T* nb, int scheme, int vu,
               const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
               hipStream_t s, T* part, T* vkin) {
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  const bool split = usf && P->usf_split;
  const unsigned ncb = (unsigned)((M.ncells + CB - 1) / CB);
  const unsigned nnb = nblk(M.nnodes);
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
  for (int64_t step = step0; step < step0 + nsteps; ++step) {
    {
      ProfScopeT ps_(P, KC_SORT, s);
      DMPM_TRY(enqueue_resort<T>(P, M, pb, eid, s));
    }
    const T t = T(step) * tdt;
    {
      ProfScopeT ps_(P, KC_TILE_KIN, s);
      k_cell_kin<T, CB><<<ncb, kThreads, 0, s>>>(M, pb, ld, sr, part, temps);
      k_node_kin<T><<<nnb, kThreads, 0, s>>>(M, part, nb, nld, vkin, P->cidx, P->cmask,
                                             P->cval);
      DMPM_TRY(launch_check("p2g_kin"));
    }
    if (split) {
      ProfScopeT ps_(P, KC_STRESS, s);
      k_stress_tile<T><<<nblk(n), kThreads, 0, s>>>(M, pb, ld, n, vkin, mat, tdt, temps);
      DMPM_TRY(launch_check("stress"));
    }
    {
      ProfScopeT ps_(P, KC_TILE_FORCE, s);
      if (usf && !split)
        k_cell_force<T, CB, true><<<ncb, kThreads, 0, s>>>(M, pb, ld, sr, part, vkin, mat, gx,
                                                           gy, tdt, temps);
      else
        k_cell_force<T, CB, false><<<ncb, kThreads, 0, s>>>(M, pb, ld, sr, part, vkin, mat, gx,
                                                            gy, tdt, temps);
      k_node_force<T><<<nnb, kThreads, 0, s>>>(M, part, nb, nld, vkin, P->cidx, P->cmask,
                                               P->cval, tdt, t, P->nfd, P->nf_off, P->nf_fn,
                                               P->nf_val, fns);
      DMPM_TRY(launch_check("p2g_force"));
    }
    {
      ProfScopeT ps_(P, KC_TILE_G2P, s);
      if (usf)
        k_g2p_tile<T, false><<<nblk(n), kThreads, 0, s>>>(M, pb, ld, n, nb, nld, move_log(P), mat,
                                                          tdt, vu, temps);
      else
        k_g2p_tile<T, true><<<nblk(n), kThreads, 0, s>>>(M, pb, ld, n, nb, nld, move_log(P), mat,
                                                         tdt, vu, temps);
      DMPM_TRY(launch_check("g2p"));
    }
  }
  return DMPM_OK;
}

template <typename T>
int tile_run_t(dmpm_plan* P, T* pb, int32_t* eid, T* nb, int scheme, int vu, const double* grav,
               double dt, int64_t step0, int64_t nsteps, int temps, hipStream_t s) {
  const Mesh2<T> M = mesh_of<T>(P);
  const int64_t ld = P->d.ld;
  const int n = (int)P->n;
  T* part = (T*)plan_workspace(P, 0, (size_t)4 * NQ_FORCE * M.ncells * sizeof(T));
  T* vkin = (T*)plan_workspace(P, 1, (size_t)2 * P->nnodes * sizeof(T));
  if (!part || !vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  // apply the cell changes still pending from the last step of the previous
  // run, then record the particles the caller moved since (both incremental
  // unless the layout is unknown)
  {
    ProfScopeT ps_(P, KC_SORT, s);
    DMPM_TRY(enqueue_resort<T>(P, M, pb, eid, s));
  }
  k_tile_check<T><<<nblk(n), kThreads, 0, s>>>(M, pb, ld, n, eid, P->key, move_log(P));
  DMPM_TRY(launch_check("tile_check"));
  switch (P->cell_block) {
    case 64:
      return tile_steps<T, 64>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                               part, vkin);
    default:
      return tile_steps<T, 128>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                                part, vkin);
    case 256:
      return tile_steps<T, 256>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                                part, vkin);
  }
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
