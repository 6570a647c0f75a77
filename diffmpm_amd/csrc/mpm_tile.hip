// Fused 2D explicit step (gfx950): re-sort by cell, then either one launch
// for the whole step (single-simulation plans, the default) or one launch per
// phase (P2G as a cell pass and a node pass, stress update, G2P).
//
// One step is:
//   k_sort_s1..s4          sort by cell: nothing when no particle changed
//                          cell, an O(N) update of the stable order from the
//                          recorded cell changes, a full counting sort when
//                          the layout is unknown
// then, one launch:
//   k_step_fast            per tile of cells, with a recomputed halo: P2G of
//                          mass and momentum, nodal velocity, USF stress
//                          update, P2G of forces, nodal update, G2P (USL:
//                          stress update), cell changes for the next sort
//   k_step_rounds          tiles that do not fit k_step_fast, and particles
//                          outside the mesh
// or, one launch per phase:
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
// (i-1,j-1), (i,j-1), (i-1,j), (i,j)); every path uses the same order and the
// same per-term arithmetic, so all of them are bit-identical.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "mpm_common.h"
#include "mpm_plan.h"
#include "mpm_step.h"
#include "mpm_tiles.h"

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

// Time of a step: t, or (clock given) T(*clock + kofs) * dt with the step
// index read on the device (the same rounding as T(step) * dt on the host).
template <typename T>
__device__ __forceinline__ T step_time(const int64_t* clock, int kofs, T t, T dt) {
  return clock ? T(*clock + kofs) * dt : t;
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
    const double* __restrict__ nf_val, FnTable fns, const int64_t* __restrict__ clock,
    int kofs) {
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
    if (qg >= 0) {
      const T tt = step_time(clock, kofs, t, dt);
      for (int k = nf_off[qg]; k < nf_off[qg + 1]; ++k)
        fe = fe + fn_value<T>(fns, nf_fn[k], tt) * T(nf_val[k]);
    }
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

// ---- single-launch step ----------------------------------------------------------
//
// One workgroup per tile of kTW x kTH cells runs a whole step for the
// particles of its cells, recomputing what it needs from rings of halo cells
// (R_k: the tile grown by k cells on each side, clipped to the mesh):
//   K  P2G of mass and momentum over R_H (H = 2 in USF, 1 in USL): kinematic
//      nodal velocities at the nodes of R_{H-1};
//   S  (USF) strain, volume and stress update of the particles of R_1 (only
//      the tile's own results are written);
//   F  P2G of body and internal force over R_1: nodal acceleration and
//      velocity at the nodes of the tile;
//   G  G2P of the tile's particles (USL: then their stress update), with the
//      cell changes recorded for the next re-sort.
// Cell sums run over each cell's slots in order and node sums over the four
// cells in the fixed order of node_gather, with the same expressions as the
// multi-kernel path, so both give bit-identical results.  Workgroups never
// wait for each other: the step reads the particle state from `src` and
// writes it to `dst` (the run alternates between the caller's block and a
// shadow block), and the previous nodal velocity that a massless node keeps
// is read from a copy (`vold`) that the step does not write.  Particles
// outside the mesh (slots from cs[ncells]) are handled by kTailBlocks extra
// workgroups.
//
// Fast path (FP32, at most kPPT slots of R_H per thread): every slot's
// particle data is loaded once, at the start, into registers; the phases
// exchange per-particle values through LDS only.  Otherwise (crowded tiles,
// FP64) the phases stage the slots in rounds and reload what they need.
constexpr int kTW = kStepTile, kTH = kStepTile;  // cells per tile
constexpr int kFT = 512;           // threads per workgroup
constexpr int kPPT = 4;            // fast path: slots of R_H per thread
constexpr int kKCAP = kFT * kPPT;  // fast path: slots of R_H
constexpr int kFCAP = 1408;        // fast path: slots of R_1
constexpr int kRound = 1024;       // staged slots per round of the other path
constexpr int kTailBlocks = 8;
constexpr int kSegRows = kTH + 4;  // rows of R_2
static_assert(kRound % kFT == 0, "round size");
// particle rows the step updates (alternating blocks) and the strain/stress
// rows it leaves alone (only set to NaN outside the mesh, in the caller's block)
constexpr int kDynRows = 14, kFixRows = 5;
__constant__ const int kDyn[kDynRows] = {
    DMPM_P_X,          DMPM_P_X + 1,      DMPM_P_V,          DMPM_P_V + 1,
    DMPM_P_VOLUME,     DMPM_P_DENSITY,    DMPM_P_VOLSTRAIN_C, DMPM_P_STRAIN + 0,
    DMPM_P_STRAIN + 1, DMPM_P_STRAIN + 3, DMPM_P_STRESS + 0, DMPM_P_STRESS + 1,
    DMPM_P_STRESS + 2, DMPM_P_STRESS + 3};
__constant__ const int kFix[kFixRows] = {DMPM_P_STRAIN + 2, DMPM_P_STRAIN + 4, DMPM_P_STRAIN + 5,
                                         DMPM_P_STRESS + 4, DMPM_P_STRESS + 5};

// Element i of row r of a particle or node block, through buffer
// instructions: one resource per block (base, size in bytes), the row offset
// a scalar, the element offset a 32-bit vector register (the host checks that
// every block is below 4 GiB).  Accesses past a block's size read 0 and write
// nothing.  Dynamic particle rows are read from src and written to dst.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const T* base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)nbytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(const T* base, uint32_t nbytes, uint32_t row_elems,
                                 uint32_t i) {
  const __amdgpu_buffer_rsrc_t rs = brsrc(base, nbytes);
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(
        T, __builtin_amdgcn_raw_buffer_load_b32(rs, i * 4u, row_elems * 4u, 0));
  else
    return __builtin_bit_cast(
        T, __builtin_amdgcn_raw_buffer_load_b64(rs, i * 8u, row_elems * 8u, 0));
}
template <typename T>
__device__ __forceinline__ void bst(T* base, uint32_t nbytes, uint32_t row_elems, uint32_t i,
                                    T v) {
  const __amdgpu_buffer_rsrc_t rs = brsrc(base, nbytes);
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs, i * 4u,
                                          row_elems * 4u, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(__attribute__((ext_vector_type(2))) uint32_t, v), rs, i * 8u,
        row_elems * 8u, 0);
}
#define SR(r, i) dyn_read(a, (r), (uint32_t)(i))
#define PB(r, i) bld<T>(a.pb, a.pbytes, (uint32_t)(r) * a.ld32, (uint32_t)(i))
#define NB(r, i) bld<T>(a.nb, a.nbytes, (uint32_t)(r) * a.nld32, (uint32_t)(i))
#define DR_W(r, i, v) dyn_write(a, (r), (uint32_t)(i), (T)(v))
#define PB_W(r, i, v) bst(a.pb, a.pbytes, (uint32_t)(r) * a.ld32, (uint32_t)(i), (T)(v))
#define NB_W(r, i, v) bst(a.nb, a.nbytes, (uint32_t)(r) * a.nld32, (uint32_t)(i), (T)(v))

// Step parameters the kernels read only in rare branches (cell changes,
// constrained nodes, concentrated nodal forces), kept in device memory so
// that they take no registers elsewhere.
struct StepAux {
  MoveLog L;
  const double* cval;
  const int* nfd;
  const int* nf_off;
  const int* nf_fn;
  const double* nf_val;
  FnTable fns;
};

template <typename T> struct StepArgs {
  Grid2<T> g;
  const T* src;
  T* dst;
  T* pb;  // static rows (mass, E_p), rows the step leaves alone, temporaries
  uint32_t ld32;
  T* nb;
  uint32_t nld32;
  uint32_t pbytes;  // bytes of a particle block (the shadow block is as large)
  uint32_t nbytes;  // bytes of the node block
  const T* vold;  // [2][vld] nodal velocity at the start of the step
  int64_t vld;    // row stride of vold
  T* vnew;        // [2][nnodes] nodal velocity after the step
  T* vkin;        // [2][nnodes] kinematic nodal velocity (for the adjoint's export)
  SortRef sr;
  const int* cidx;
  const int* cmask;
  const StepAux* aux;  // tables used in rare branches
  int has_nf;          // concentrated nodal forces present
  Mat2 mat;
  T gx, gy, dt, t;
  int vu, temps, tiles_x, ntiles, n;
  int* ovf;  // [2] overflow counts by step parity, then [2][ntiles] tile lists
  int par;   // step parity
  int last;  // the run's last step: the node block and vkin are written (earlier steps
             // only pass the nodal velocity on through vnew)
  const T* fdsrc;  // node block whose damping-force rows the step reads (nb or another
                   // block of the same shape)
  int wfd;         // on the last step, also write those rows to nb
  int nrb;   // workgroups of k_step_rounds that process tiles
  const SimRec* sims;   // batch plans: the simulation table (tile_off, tiles_x, ...)
  const int* tile_sim;  // batch plans: the simulation of each tile
  int nc_all, nn_all, nsims;  // cells, nodes and simulations of the plan
  const int64_t* clock;  // optional: device step counter (time = (*clock + kofs) * dt)
  int kofs;
  T* sout;   // optional: the state rows x, y, vx, vy, sxx, syy, sxy, V after the step
  const T* sin;  // optional: the state rows before the step (then read from here, and
                 // written only to sout)
  uint32_t sld32, sbytes;
};

// Row of the state copy (x, y, vx, vy, sxx, syy, sxy, V) that particle row r
// goes to, or -1.
__device__ __forceinline__ int state_row(int r) {
  switch (r) {
    case DMPM_P_X: return 0;
    case DMPM_P_X + 1: return 1;
    case DMPM_P_V: return 2;
    case DMPM_P_V + 1: return 3;
    case DMPM_P_STRESS + 0: return 4;
    case DMPM_P_STRESS + 1: return 5;
    case DMPM_P_STRESS + 3: return 6;
    case DMPM_P_VOLUME: return 7;
    default: return -1;
  }
}
// A dynamic row's value before the step: from the state tensor sin (state
// rows, when given) or from src.
template <typename T>
__device__ __forceinline__ T dyn_read(const StepArgs<T>& a, int r, uint32_t i) {
  if (a.sin) {
    const int q = state_row(r);
    if (q >= 0) return bld(a.sin, a.sbytes, (uint32_t)q * a.sld32, i);
  }
  return bld(a.src, a.pbytes, (uint32_t)r * a.ld32, i);
}
// A dynamic row's value after the step: to dst, and to the state tensor sout
// (state rows; only there when sin is given).
template <typename T>
__device__ __forceinline__ void dyn_write(const StepArgs<T>& a, int r, uint32_t i, T v) {
  const int q = a.sout ? state_row(r) : -1;
  if (q >= 0) bst(a.sout, a.sbytes, (uint32_t)q * a.sld32, i, v);
  if (q < 0 || !a.sin) bst(a.dst, a.pbytes, (uint32_t)r * a.ld32, i, v);
}

// Cell (I, J) of element_id<T, 2>(x, y) without the integer division: the
// element search's quirks map to element 0 (below/left of the mesh, NaN) and
// to the first cell of the next row (x at or beyond the last node).
template <typename T>
__device__ __forceinline__ void cell_of(const Grid2<T>& g, T x, T y, int& I, int& J) {
  I = last_le(x, g.nx, g.hx, g.ihx);
  J = last_le(y, g.ny, g.hy, g.ihy);
  if (I < 0 || J < 0) {
    I = 0;
    J = 0;
  } else if (I == g.nx) {
    I = 0;
    J = J + 1;
  }
}

// Row of virtual index v of a region whose rows hold P[r+1] - P[r] slots.
__device__ __forceinline__ int seg_row(const int* P, int h, int v) {
  int lo = 0, hi = h - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Context of global tile position gid of a step (B: batch plan).
template <typename T, bool B>
__device__ __forceinline__ TileCtx<T> tile_ctx(const StepArgs<T>& a, int gid, bool strip) {
  return tile_ctx_of<T, B>(TileTab<T>{a.sims, a.tile_sim, a.tiles_x, a.ntiles, a.g, a.mat}, gid,
                           strip);
}

// KIN cell sum over staged slots [lo, hi) (k_cell_kin's expressions).
template <typename T>
__device__ __forceinline__ void kin_cell(const T* st, int stride, int lo, int hi,
                                         T acc[4 * NQ_KIN]) {
  for (int q = lo; q < hi; ++q) {
    T N[4];
    q4_N4(st[q], st[stride + q], N);
    const T mq = st[2 * stride + q], mvx = st[3 * stride + q], mvy = st[4 * stride + q];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[NQ_KIN * k + 0] = acc[NQ_KIN * k + 0] + mq * N[k];
      acc[NQ_KIN * k + 1] = acc[NQ_KIN * k + 1] + N[k] * mvx;
      acc[NQ_KIN * k + 2] = acc[NQ_KIN * k + 2] + N[k] * mvy;
    }
  }
}

// FORCE cell sum over staged slots [lo, hi) of cell (X0..X1, Y0..Y1)
// (k_cell_force's expressions).
template <typename T>
__device__ __forceinline__ void force_cell(const T* st, int stride, int lo, int hi, T X0, T X1,
                                           T Y0, T Y1, T gx, T gy, T acc[4 * NQ_FORCE]) {
  // q4_invJ depends on the cell only
  T ix, iy;
  q4_invJ(T(0), T(0), X0, X1, Y0, Y1, ix, iy);
  for (int q = lo; q < hi; ++q) {
    const T xi = st[q], eta = st[stride + q];
    const T bx = st[2 * stride + q] * gx, by = st[2 * stride + q] * gy;
    const T nV = st[3 * stride + q], sxx = st[4 * stride + q], syy = st[5 * stride + q],
            sxy = st[6 * stride + q];
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

// Node sums (node_gather's order) for the nodes of region RN from the cell
// partials of region RC (which holds every cell of those nodes), through LDS
// buffer sl (4 * NQ * cells of RC).  Called by every thread; ends with a
// barrier.
template <typename T, int NQ, int C, int NPT>
__device__ __forceinline__ void node_sums(T* sl, const T (&acc)[C][4 * NQ], const Rgn& RC,
                                          const Rgn& RN, int nx, int ny, T (&nacc)[NPT][NQ]) {
  const int t = threadIdx.x;
  const int wc = RC.i1 - RC.i0, nc = wc * (RC.j1 - RC.j0);
  const int wn = RN.i1 - RN.i0 + 1, nn = wn * (RN.j1 - RN.j0 + 1);
  const int dI[4] = {-1, 0, -1, 0}, dJ[4] = {-1, -1, 0, 0}, corner[4] = {2, 3, 1, 0};
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int lc = t + u * kFT;
    if (lc < nc)
#pragma unroll
      for (int v = 0; v < 4 * NQ; ++v) sl[v * nc + lc] = acc[u][v];
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) nacc[w][q] = T(0);
    const int ln = t + w * kFT;
    if (ln >= nn) continue;
    const int jr = sdiv(dwn, ln);
    const int i = RN.i0 + (ln - jr * wn), j = RN.j0 + jr;
    // every corner's partials are read first (cells outside the mesh read
    // cell 0 of RC and are not added), so the reads overlap
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

// Node-table values of this thread's nodes of region RN (w-th node: index
// t + w * kFT), loaded ahead of the epilogues: constraint row and mask,
// damping force.
template <int NPT> struct NodePre {
  int cr[NPT], mask[NPT];
};
template <typename T, int NPT>
__device__ __forceinline__ void node_pre(const StepArgs<T>& a, const TileCtx<T>& c, const Rgn& RN,
                                         NodePre<NPT>& q, T (*fd)[2]) {
  const int t = threadIdx.x, nx = c.g.nx;
  const int wn = RN.i1 - RN.i0 + 1, nn = wn * (RN.j1 - RN.j0 + 1);
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = min(t + w * kFT, nn - 1);
    const int jr = sdiv(dwn, ln);
    const int nd = c.noff + (RN.j0 + jr) * (nx + 1) + RN.i0 + (ln - jr * wn);
    q.cr[w] = a.cidx ? a.cidx[nd] : -1;
    q.mask[w] = 0;
    if (fd) {
      fd[w][0] = bld<T>(a.fdsrc, a.nbytes, (uint32_t)DMPM_N_FDAMP * a.nld32, (uint32_t)nd);
      fd[w][1] = bld<T>(a.fdsrc, a.nbytes, (uint32_t)(DMPM_N_FDAMP + 1) * a.nld32, (uint32_t)nd);
    }
  }
}
template <typename T, int NPT>
__device__ __forceinline__ void node_pre_mask(const StepArgs<T>& a, NodePre<NPT>& q) {
#pragma unroll
  for (int w = 0; w < NPT; ++w) q.mask[w] = q.cr[w] >= 0 ? a.cmask[q.cr[w]] : 0;
}

// Kinematic nodal epilogue (k_node_kin) of the nodes of RN: nodal mass and
// velocity into LDS, the velocity of the tile's own nodes to a.vkin.
template <typename T, int NPT>
__device__ __forceinline__ void kin_epilogue(const StepArgs<T>& a, const TileCtx<T>& c,
                                             const T (&nacc)[NPT][NQ_KIN],
                                             const NodePre<NPT>& pre, const Rgn& RN,
                                             const Rgn& R0, T* vk, T* nm, int nkn) {
  const int t = threadIdx.x, nx = c.g.nx, ny = c.g.ny, nnodes = a.nn_all;
  const int wn = RN.i1 - RN.i0 + 1, nn = wn * (RN.j1 - RN.j0 + 1);
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = t + w * kFT;
    if (ln >= nn) continue;
    const int jr = sdiv(dwn, ln);
    const int i = RN.i0 + (ln - jr * wn), j = RN.j0 + jr;
    const int nd = c.noff + j * (nx + 1) + i;
    const int cr = pre.cr[w], mask = pre.mask[w];
    const T m = nacc[w][0];
    const T p[2] = {snap(nacc[w][1]), snap(nacc[w][2])};
    const bool own = i >= R0.i0 && j >= R0.j0 && (i < R0.i1 || i == nx) && (j < R0.j1 || j == ny);
    nm[ln] = m;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      T v = (m == T(0)) ? a.vold[(int64_t)d * a.vld + nd] : p[d] / m;
      v = snap(v);
      if (mask & (1 << d)) v = T(a.aux->cval[2 * cr + d]);
      vk[d * nkn + ln] = v;
      if (own && a.last) a.vkin[(int64_t)d * nnodes + nd] = v;
    }
  }
}

// Force nodal epilogue (k_node_force) of the tile's nodes: acceleration and
// velocity into LDS; the tile's own nodes are written to the node block.
template <typename T, int NPT>
__device__ __forceinline__ void force_epilogue(const StepArgs<T>& a, const TileCtx<T>& c,
                                               const T (&nacc)[NPT][NQ_FORCE],
                                               const NodePre<NPT>& pre, const T (*fd)[2],
                                               const Rgn& R0, const Rgn& RN, const T* vk,
                                               const T* nm, int nkn, T* nv, T* na, int ntn) {
  const int t = threadIdx.x, nx = c.g.nx, ny = c.g.ny, nnodes = a.nn_all;
  const int wt = R0.i1 - R0.i0 + 1, nt = wt * (R0.j1 - R0.j0 + 1);
  const int wkn = RN.i1 - RN.i0 + 1;
  const SmallDiv dwt = small_div(wt);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = t + w * kFT;
    if (ln >= nt) continue;
    const int jr = sdiv(dwt, ln);
    const int i = R0.i0 + (ln - jr * wt), j = R0.j0 + jr;
    const int nd = c.noff + j * (nx + 1) + i;
    const int lk = (i - RN.i0) + (j - RN.j0) * wkn;
    const bool own = (i < R0.i1 || i == nx) && (j < R0.j1 || j == ny);
    const int cr = pre.cr[w], mask = pre.mask[w];
    const T m = nm[lk];
    if (own && a.last) NB_W(DMPM_N_MASS, nd, m);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      T fe = nacc[w][d];
      if (a.has_nf) {
        const StepAux& x = *a.aux;
        const int qg = x.nfd[(int64_t)d * nnodes + nd];
        if (qg >= 0)
          for (int k = x.nf_off[qg]; k < x.nf_off[qg + 1]; ++k)
            fe = fe + fn_value<T>(x.fns, x.nf_fn[k], step_time(a.clock, a.kofs, a.t, a.dt)) *
                          T(x.nf_val[k]);
      }
      const T fi = nacc[w][2 + d];
      const T F = (fi + fe) + fd[w][d];
      T ac = nan_to_num(F / m);
      T v = vk[d * nkn + lk] + ac * a.dt;
      T pm = m * v;
      v = snap(v);
      ac = snap(ac);
      if (mask & (1 << d)) {
        const T c = T(a.aux->cval[2 * cr + d]);
        v = snap(c);
        pm = m * c;
        ac = T(0);
      }
      nv[d * ntn + ln] = v;
      na[d * ntn + ln] = ac;
      if (own && a.last) {
        NB_W(DMPM_N_FEXT + d, nd, fe);
        NB_W(DMPM_N_FINT + d, nd, fi);
        NB_W(DMPM_N_V + d, nd, v);
        NB_W(DMPM_N_A + d, nd, ac);
        NB_W(DMPM_N_P + d, nd, pm);
        if (a.wfd) NB_W(DMPM_N_FDAMP + d, nd, fd[w][d]);
      }
      if (own) a.vnew[(int64_t)d * nnodes + nd] = v;
    }
  }
}

// Virtual slot range [v0, v1) of each of this thread's cells of region R
// (table q of the slot segments).
template <int C>
__device__ __forceinline__ void cell_ranges(const Rgn& R, int nx, int coff, const int* cs,
                                            const int* A, const int* P, int (&v0)[C],
                                            int (&v1)[C]) {
  const int w = R.i1 - R.i0, nc = w * (R.j1 - R.j0);
  const SmallDiv dw = small_div(w);
#pragma unroll
  for (int u = 0; u < C; ++u) {
    const int lc = threadIdx.x + u * kFT;
    v0[u] = v1[u] = 0;
    if (lc < nc) {
      const int r = sdiv(dw, lc);
      const int k = coff + (R.j0 + r) * nx + R.i0 + (lc - r * w);
      const int s0 = cs[k];
      v0[u] = P[r] + (s0 - A[r]);
      v1[u] = v0[u] + (cs[k + 1] - s0);
    }
  }
}

// The strain/volume/stress update of one particle (particle.py:249-327, as in
// k_stress_tile) from the four nodal velocities of its cell.
template <typename T> struct StressUpd {
  T rr[6], d[6], ds[4], dvol, f;
};
template <typename T>
__device__ __forceinline__ void stress_update(const Mat2& mat, T E_p, T ix, T iy, T xi, T eta,
                                              const T vx[4], const T vy[4], T dt,
                                              StressUpd<T>& u) {
  T rc[6];
  rate2s(ix, iy, xi, eta, vx, vy, u.rr);
  rate2s(ix, iy, T(0), T(0), vx, vy, rc);
  for (int k = 0; k < 6; ++k) u.d[k] = u.rr[k] * dt;
  u.dvol = dt * (rc[0] + rc[1]);
  stress_inc(mat, E_p, u.d, u.ds);
  u.f = T(1) + u.dvol;
}

// The step for the tiles whose slots fit the fast path (FP32); a tile that
// does not fit appends itself to the overflow list of this step for
// k_step_rounds.
template <typename T, bool USF, bool B>
__global__ __launch_bounds__(kFT) __attribute__((amdgpu_waves_per_eu(4))) void k_step_fast(
    StepArgs<T> a) {
  constexpr int H = USF ? 2 : 1;
  constexpr int CK = ((kTW + 2 * H) * (kTH + 2 * H) + kFT - 1) / kFT;  // cells/thread, K
  constexpr int CF = ((kTW + 2) * (kTH + 2) + kFT - 1) / kFT;          // cells/thread, F
  constexpr int NKN = (kTW + 2 * H - 1) * (kTH + 2 * H - 1);            // nodes of R_{H-1}
  constexpr int NTN = (kTW + 1) * (kTH + 1);                           // nodes of the tile
  constexpr int NK = (NKN + kFT - 1) / kFT, NF = (NTN + kFT - 1) / kFT;
  constexpr int NCOL = kTW + 2 * H, NROW = kTH + 2 * H;
  static_assert(CK * kFT <= 1024 && CF * kFT <= 1024 && NK * kFT <= 1024 &&
                    NF * kFT <= 1024 && NCOL + 1 <= 64,
                "small_div range");
  __shared__ T st[5 * kKCAP];  // staged particle inputs; node-pass slices alias it
  static_assert(7 * kFCAP <= 5 * kKCAP, "F staging");
  __shared__ T s1f[5 * kFCAP];  // sxx, syy, sxy, V, E_p of the R_1 slots
  // nodal arrays; until the K epilogue they hold the slot table
  __shared__ T ndl[3 * NKN + 4 * NTN];
  static_assert(sizeof(T) * (3 * NKN + 4 * NTN) >= sizeof(uint16_t) * kKCAP, "slot table");
  T* const vk = ndl;
  T* const nm = ndl + 2 * NKN;
  T* const nv = ndl + 3 * NKN;
  T* const na = ndl + 3 * NKN + 2 * NTN;
  uint16_t* const tab = reinterpret_cast<uint16_t*>(ndl);
  __shared__ T colC[NCOL], colS[NCOL], rowC[NROW], rowS[NROW];
  __shared__ int segA[3][kSegRows], segP[3][kSegRows + 1];
  __shared__ int rowOf[kKCAP / 32];  // row of R_H holding virtual index 32 c
  const TileCtx<T> c = tile_ctx<T, B>(a, xcd_pos(blockIdx.x, a.ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx, ny = g.ny;
  const int t = threadIdx.x;
  const int* __restrict__ cs = a.sr.cs();
  const int* __restrict__ perm = a.sr.perm();
  const int ident = a.sr.flags[1];

  const int tile = c.tile, ti = c.ti, tj = c.tj;
  const Rgn R0{ti * kTW, min(ti * kTW + kTW, nx), tj * kTH, min(tj * kTH + kTH, ny)};
  const Rgn R1 = grow(R0, 1, nx, ny);
  const Rgn RH = grow(R0, H, nx, ny);
  const Rgn RN = USF ? R1 : R0;  // cells whose nodes get kinematic velocities
  const int h0 = R0.j1 - R0.j0, h1 = R1.j1 - R1.j0, hH = RH.j1 - RH.j0;

  // row, column and slot range of this thread's cells of RH (the slot table)
  int hr[CK], hcol[CK], hs0[CK], hs1[CK];
  cell_slots_rc<CK, kFT>(RH, nx, c.coff, cs, hr, hcol, hs0, hs1);
  // slot segments of the rows of R0 (table 0), R1 (1) and RH (2): one wave
  // per table, lane r = row r, prefix by shuffles
  {
    const int q = t >> 6, r = t & 63;
    if (q < 3) {
      const Rgn R = q == 0 ? R0 : (q == 1 ? R1 : RH);
      int len = 0;
      if (r < R.j1 - R.j0) {
        const int k = c.coff + (R.j0 + r) * nx + R.i0;
        const int s0 = cs[k];
        segA[q][r] = s0;
        len = cs[k + (R.i1 - R.i0)] - s0;
      }
      const int own_len = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(len, o, 64);
        if (r >= o) len += y;
      }
      if (r < kSegRows) segP[q][r + 1] = len;
      if (r == 0) segP[q][0] = 0;
      // R_H: the row of every 32nd virtual index (a row's indices are
      // [len - own_len, len))
      if (q == 2 && r < R.j1 - R.j0)
        for (int cq = (len - own_len + 31) >> 5; cq < min((len + 31) >> 5, kKCAP / 32); ++cq)
          rowOf[cq] = r;
    } else if (q == 3) {
      // cell centres and scales 2/(X1 - X0) of the columns and rows of RH
      // (natural coordinates and inverse Jacobians, particle.py:180-206)
      if (r < RH.i1 - RH.i0) {
        const T X0 = g.X(RH.i0 + r), X1 = g.X(RH.i0 + r + 1);
        colC[r] = (X0 + X1) / T(2);
        colS[r] = T(2) / (X1 - X0);
      }
      if (r < hH) {
        const T Y0 = g.Y(RH.j0 + r), Y1 = g.Y(RH.j0 + r + 1);
        rowC[r] = (Y0 + Y1) / T(2);
        rowS[r] = T(2) / (Y1 - Y0);
      }
    }
  }
  __syncthreads();
  const int SH = segP[2][hH], S1 = segP[1][h1];
  const int wK = RH.i1 - RH.i0;
  const int wKN = RN.i1 - RN.i0 + 1, wTN = R0.i1 - R0.i0 + 1;

  if (SH > kKCAP || S1 > kFCAP) {
    if (t == 0) {
      const int i = atomicAdd(&a.ovf[a.par], 1);
      a.ovf[2 + a.par * a.ntiles + i] = tile;
    }
    return;
  }
  constexpr int PPT = kPPT;
  // slot and particle of each virtual index t + j*kFT of the R_H list
  // (lanes past the end take particle 0): the row from the row table, so
  // that the permutation loads are in flight while the slot table is built
  int pp[PPT], sl[PPT];
  unsigned live = 0;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int v = t + j * kFT;
    const bool lv = v < SH;
    int r = 0;
    if (lv) {
      r = rowOf[v >> 5];
      while (segP[2][r + 1] <= v) ++r;
      live |= 1u << j;
    }
    sl[j] = lv ? segA[2][r] + (v - segP[2][r]) : 0;
    pp[j] = ident ? sl[j] : perm[sl[j]];
  }
  // node tables of the epilogues, loaded ahead
  NodePre<NK> kpre;
  NodePre<NF> fpre;
  T fd[NF][2];
  node_pre<T, NK>(a, c, RN, kpre, nullptr);
  node_pre<T, NF>(a, c, R0, fpre, fd);
  // the slot table of the R_H list: every slot's cell (row << 8 | column
  // of RH), written by the threads that own the cells
  int kv0[CK], kv1[CK];
  cell_ranges_of<CK>(hr, hs0, hs1, segA[2], segP[2], kv0, kv1);
  fill_slot_cells<CK>(tab, hr, hcol, kv0, kv1, 0, SH);
  // the fields every slot needs, issued before the barrier
  T x[PPT], y[PPT], vx[PPT], vy[PPT], m[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int p = pp[j];
    x[j] = SR(DMPM_P_X, p);
    y[j] = SR(DMPM_P_X + 1, p);
    vx[j] = SR(DMPM_P_V, p);
    vy[j] = SR(DMPM_P_V + 1, p);
    m[j] = PB(DMPM_P_MASS, p);
  }
  __syncthreads();
  node_pre_mask<T, NK>(a, kpre);
  node_pre_mask<T, NF>(a, fpre);
  // cell and membership of each slot
  int v1i[PPT], Ii[PPT], Ji[PPT];
  unsigned own = 0;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int v = t + j * kFT;
    const bool lv = (live >> j) & 1u;
    const int code = lv ? (int)tab[v] : 0;
    // the slot's cell is the element search's result for this position (the
    // re-sort keyed the slot by it)
    const int I = RH.i0 + (code & 255), J = RH.j0 + (code >> 8);
    Ii[j] = I;
    Ji[j] = J;
    v1i[j] = -1;
    if (lv && in_rgn(R1, I, J)) {
      const int r1 = J - R1.j0;
      v1i[j] = segP[1][r1] + (sl[j] - segA[1][r1]);
    }
    if (lv && in_rgn(R0, I, J)) own |= 1u << j;
  }
  // the other particle loads of the step, issued together (lanes without a
  // field read particle 0's)
  T szz[PPT], ex[PPT], ey[PPT], gxy[PPT], rho[PPT], evc[PPT];
  {
    T sxx[PPT], syy[PPT], sxy[PPT], V[PPT], E[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int p = pp[j];
      const int p1 = v1i[j] >= 0 ? p : 0;
      const int po = (own & (1u << j)) ? p : 0;
      sxx[j] = SR(DMPM_P_STRESS + 0, p1);
      syy[j] = SR(DMPM_P_STRESS + 1, p1);
      sxy[j] = SR(DMPM_P_STRESS + 3, p1);
      V[j] = SR(DMPM_P_VOLUME, p1);
      if constexpr (USF) {
        // (USL loads these in its G phase, where it uses them: held from
        // here, they spill)
        szz[j] = SR(DMPM_P_STRESS + 2, po);
        ex[j] = SR(DMPM_P_STRAIN + 0, po);
        ey[j] = SR(DMPM_P_STRAIN + 1, po);
        gxy[j] = SR(DMPM_P_STRAIN + 3, po);
        rho[j] = SR(DMPM_P_DENSITY, po);
        evc[j] = SR(DMPM_P_VOLSTRAIN_C, po);
      }
    }
    if (a.mat.per_E)
#pragma unroll
      for (int j = 0; j < PPT; ++j) E[j] = PB(DMPM_P_YOUNGS, v1i[j] >= 0 ? pp[j] : 0);
    else
#pragma unroll
      for (int j = 0; j < PPT; ++j) E[j] = T(0);
    // the R_1 fields wait in LDS until the S phase
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      if (v1i[j] < 0) continue;
      const int q = v1i[j];
      s1f[q] = sxx[j];
      s1f[kFCAP + q] = syy[j];
      s1f[2 * kFCAP + q] = sxy[j];
      s1f[3 * kFCAP + q] = V[j];
      s1f[4 * kFCAP + q] = E[j];
    }
  }
  // ---- K ----
  T xi[PPT], eta[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int v = t + j * kFT;
    const int I = Ii[j], J = Ji[j];
    const int ci = min(max(I - RH.i0, 0), NCOL - 1), rj = min(max(J - RH.j0, 0), NROW - 1);
    xi[j] = (x[j] - colC[ci]) * colS[ci];
    eta[j] = (y[j] - rowC[rj]) * rowS[rj];
    st[v] = xi[j];
    st[kKCAP + v] = eta[j];
    st[2 * kKCAP + v] = m[j];
    st[3 * kKCAP + v] = m[j] * vx[j];
    st[4 * kKCAP + v] = m[j] * vy[j];
    if (a.temps && (own & (1u << j))) {
      PB_W(DMPM_P_XI, pp[j], xi[j]);
      PB_W(DMPM_P_XI + 1, pp[j], eta[j]);
    }
  }
  __syncthreads();
  {
    T acc[CK][4 * NQ_KIN];
#pragma unroll
    for (int u = 0; u < CK; ++u) {
#pragma unroll
      for (int v = 0; v < 4 * NQ_KIN; ++v) acc[u][v] = T(0);
      kin_cell(st, kKCAP, kv0[u], kv1[u], acc[u]);
    }
    __syncthreads();
    T nacc[NK][NQ_KIN];
    node_sums<T, NQ_KIN, CK, NK>(st, acc, RH, RN, nx, ny, nacc);
    kin_epilogue<T, NK>(a, c, nacc, kpre, RN, R0, vk, nm, NKN);
  }
  int fv0[CF], fv1[CF];
  cell_ranges<CF>(R1, nx, c.coff, cs, segA[1], segP[1], fv0, fv1);
  __syncthreads();
  // ---- S: stresses (USF); F inputs staged by R_1 slot ----
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (v1i[j] < 0) continue;
    const int q = v1i[j];
    const T sxx = s1f[q], syy = s1f[kFCAP + q], sxy = s1f[2 * kFCAP + q],
            V = s1f[3 * kFCAP + q], E = s1f[4 * kFCAP + q];
    T Vn = V, sx = sxx, sy = syy, sq = sxy;
    if (USF) {
      const int I = Ii[j], J = Ji[j];
      const int Ic = min(max(I, RN.i0), RN.i1 - 1), Jc = min(max(J, RN.j0), RN.j1 - 1);
      const int ll = (Ic - RN.i0) + (Jc - RN.j0) * wKN;
      const int ids[4] = {ll, ll + 1, ll + wKN + 1, ll + wKN};
      T nvx[4], nvy[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        nvx[k] = vk[ids[k]];
        nvy[k] = vk[NKN + ids[k]];
      }
      const int ci = min(max(I - RH.i0, 0), NCOL - 1), rj = min(max(J - RH.j0, 0), NROW - 1);
      StressUpd<T> u;
      stress_update(c.mat, E, colS[ci], rowS[rj], xi[j], eta[j], nvx, nvy, a.dt, u);
      sx = sxx + u.ds[0];
      sy = syy + u.ds[1];
      sq = sxy + u.ds[3];
      Vn = V * u.f;
      if (own & (1u << j)) {
        const int p = pp[j];
        DR_W(DMPM_P_STRAIN + 0, p, ex[j] + u.d[0]);
        DR_W(DMPM_P_STRAIN + 1, p, ey[j] + u.d[1]);
        DR_W(DMPM_P_STRAIN + 3, p, gxy[j] + u.d[3]);
        DR_W(DMPM_P_STRESS + 0, p, sx);
        DR_W(DMPM_P_STRESS + 1, p, sy);
        DR_W(DMPM_P_STRESS + 2, p, szz[j] + u.ds[2]);
        DR_W(DMPM_P_STRESS + 3, p, sq);
        DR_W(DMPM_P_VOLUME, p, Vn);
        DR_W(DMPM_P_DENSITY, p, rho[j] / u.f);
        DR_W(DMPM_P_VOLSTRAIN_C, p, evc[j] + u.dvol);
        if (a.temps) {
          for (int k = 0; k < 6; ++k) {
            PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
            PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
          }
          PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
        }
      }
    }
    st[q] = xi[j];
    st[kFCAP + q] = eta[j];
    st[2 * kFCAP + q] = m[j];
    st[3 * kFCAP + q] = -Vn;
    st[4 * kFCAP + q] = sx;
    st[5 * kFCAP + q] = sy;
    st[6 * kFCAP + q] = sq;
  }
  __syncthreads();
  // ---- F ----
  {
    T acc[CF][4 * NQ_FORCE];
    const int wF = R1.i1 - R1.i0;
    const SmallDiv dwF = small_div(wF);
#pragma unroll
    for (int u = 0; u < CF; ++u) {
#pragma unroll
      for (int v = 0; v < 4 * NQ_FORCE; ++v) acc[u][v] = T(0);
      const int lc = t + u * kFT;
      if (lc < wF * h1) {
        const int jr = sdiv(dwF, lc);
        const int I = R1.i0 + (lc - jr * wF), J = R1.j0 + jr;
        force_cell(st, kFCAP, fv0[u], fv1[u], g.X(I), g.X(I + 1), g.Y(J), g.Y(J + 1), a.gx,
                   a.gy, acc[u]);
      }
    }
    __syncthreads();
    T nacc[NF][NQ_FORCE];
    node_sums<T, NQ_FORCE, CF, NF>(st, acc, R1, R0, nx, ny, nacc);
    force_epilogue<T, NF>(a, c, nacc, fpre, fd, R0, RN, vk, nm, NKN, nv, na, NTN);
  }
  __syncthreads();
  // ---- G ----
  if constexpr (!USF) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int po = (own & (1u << j)) ? pp[j] : 0;
      szz[j] = SR(DMPM_P_STRESS + 2, po);
      ex[j] = SR(DMPM_P_STRAIN + 0, po);
      ey[j] = SR(DMPM_P_STRAIN + 1, po);
      gxy[j] = SR(DMPM_P_STRAIN + 3, po);
      rho[j] = SR(DMPM_P_DENSITY, po);
      evc[j] = SR(DMPM_P_VOLSTRAIN_C, po);
    }
  }
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (!(own & (1u << j))) continue;
    const int p = pp[j];
    const int I = Ii[j], J = Ji[j];
    const int Ic = min(max(I, R0.i0), R0.i1 - 1), Jc = min(max(J, R0.j0), R0.j1 - 1);
    const int ll = (Ic - R0.i0) + (Jc - R0.j0) * wTN;
    const int ids[4] = {ll, ll + 1, ll + wTN + 1, ll + wTN};
    T N[4];
    q4_N4(xi[j], eta[j], N);
    T gvx[4], gvy[4];
    T nvx = T(0), nvy = T(0), nax = T(0), nay = T(0);
    for (int k = 0; k < 4; ++k) {
      gvx[k] = nv[ids[k]];
      gvy[k] = nv[NTN + ids[k]];
      nvx = nvx + N[k] * gvx[k];
      nvy = nvy + N[k] * gvy[k];
      nax = nax + N[k] * na[ids[k]];
      nay = nay + N[k] * na[NTN + ids[k]];
    }
    const T pvx = a.vu ? nvx : vx[j] + nax * a.dt;
    const T pvy = a.vu ? nvy : vy[j] + nay * a.dt;
    const T x1 = x[j] + nvx * a.dt, y1 = y[j] + nvy * a.dt;
    DR_W(DMPM_P_V, p, pvx);
    DR_W(DMPM_P_V + 1, p, pvy);
    DR_W(DMPM_P_X, p, x1);
    DR_W(DMPM_P_X + 1, p, y1);
    if (a.temps) {
      PB_W(DMPM_P_MOMENTUM, p, m[j] * pvx);
      PB_W(DMPM_P_MOMENTUM + 1, p, m[j] * pvy);
    }
    if (!USF) {
      const int ci = min(max(I - RH.i0, 0), NCOL - 1), rj = min(max(J - RH.j0, 0), NROW - 1);
      const int q = v1i[j];
      StressUpd<T> u;
      stress_update(c.mat, s1f[4 * kFCAP + q], colS[ci], rowS[rj], xi[j], eta[j], gvx, gvy,
                    a.dt, u);
      DR_W(DMPM_P_STRAIN + 0, p, ex[j] + u.d[0]);
      DR_W(DMPM_P_STRAIN + 1, p, ey[j] + u.d[1]);
      DR_W(DMPM_P_STRAIN + 3, p, gxy[j] + u.d[3]);
      DR_W(DMPM_P_STRESS + 0, p, s1f[q] + u.ds[0]);
      DR_W(DMPM_P_STRESS + 1, p, s1f[kFCAP + q] + u.ds[1]);
      DR_W(DMPM_P_STRESS + 2, p, szz[j] + u.ds[2]);
      DR_W(DMPM_P_STRESS + 3, p, s1f[2 * kFCAP + q] + u.ds[3]);
      DR_W(DMPM_P_VOLUME, p, s1f[3 * kFCAP + q] * u.f);
      DR_W(DMPM_P_DENSITY, p, rho[j] / u.f);
      DR_W(DMPM_P_VOLSTRAIN_C, p, evc[j] + u.dvol);
      if (a.temps) {
        for (int k = 0; k < 6; ++k) {
          PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
          PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
        }
        PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
      }
    }
    const int k0 = c.coff + J * nx + I;
    // still inside cell (I, J): no element search (the same element id)
    if (x1 >= g.X(I) && x1 < g.X(I + 1) && y1 >= g.Y(J) && y1 < g.Y(J + 1)) continue;
    const int e1 = element_id<T, 2>(g.geo(), x1, y1);
    const int k1 = (e1 >= 0 && e1 < g.ncells) ? c.coff + e1 : a.nc_all;
    if (k1 != k0) record_move(a.aux->L, p, k0, e1, k1);
  }
}

// The step for the tiles of the overflow list of this step (all tiles when
// ovf is null: FP64), slots staged in rounds; workgroups from nrb on handle
// the particles outside the mesh.
template <typename T, bool USF, bool B>
__global__ __launch_bounds__(kFT) void k_step_rounds(StepArgs<T> a) {
  constexpr int H = USF ? 2 : 1;
  constexpr int CK = ((kTW + 2 * H) * (kTH + 2 * H) + kFT - 1) / kFT;  // cells/thread, K
  constexpr int CF = ((kTW + 2) * (kTH + 2) + kFT - 1) / kFT;          // cells/thread, F
  constexpr int NKN = (kTW + 2 * H - 1) * (kTH + 2 * H - 1);            // nodes of R_{H-1}
  constexpr int NTN = (kTW + 1) * (kTH + 1);                           // nodes of the tile
  constexpr int NK = (NKN + kFT - 1) / kFT, NF = (NTN + kFT - 1) / kFT;
  constexpr int NCOL = kTW + 2 * H, NROW = kTH + 2 * H;
  static_assert(CK * kFT <= 1024 && CF * kFT <= 1024 && NK * kFT <= 1024 &&
                    NF * kFT <= 1024 && NCOL + 1 <= 64,
                "small_div range");
  __shared__ T st[7 * kRound];  // staged particle inputs; node-pass slices alias it
  __shared__ T vk[2 * NKN];
  __shared__ T nm[NKN];
  __shared__ T nv[2 * NTN];
  __shared__ T na[2 * NTN];
  __shared__ T colC[NCOL], colS[NCOL], rowC[NROW], rowS[NROW];
  __shared__ int segA[3][kSegRows], segP[3][kSegRows + 1];
  const int t = threadIdx.x;
  const int* __restrict__ cs = a.sr.cs();
  const int* __restrict__ perm = a.sr.perm();
  const int ident = a.sr.flags[1];

  if ((int)blockIdx.x >= a.nrb) {
    // particles outside the mesh: the state becomes NaN (as in k_g2p_tile)
    const T nan = qnan<T>();
    const int gs = kTailBlocks * kFT;
    for (int s = cs[a.nc_all] + ((int)blockIdx.x - a.nrb) * kFT + t; s < a.n; s += gs) {
      const int p = ident ? s : perm[s];
      for (int k = 0; k < kDynRows; ++k) DR_W(kDyn[k], p, nan);
      for (int k = 0; k < kFixRows; ++k) PB_W(kFix[k], p, nan);
      if (a.temps) {
        PB_W(DMPM_P_XI, p, nan);
        PB_W(DMPM_P_XI + 1, p, nan);
        PB_W(DMPM_P_MOMENTUM, p, nan);
        PB_W(DMPM_P_MOMENTUM + 1, p, nan);
        for (int k = 0; k < 6; ++k) {
          PB_W(DMPM_P_STRAIN_RATE + k, p, nan);
          PB_W(DMPM_P_DSTRAIN + k, p, nan);
        }
        PB_W(DMPM_P_DVOLSTRAIN, p, nan);
      }
      // NaN positions search to element 0 (element.py:827-837) of the
      // particle's own mesh
      int c0 = 0;
      if constexpr (B) {
        int lo = 0, hi = a.nsims - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (a.sims[mid].part_off <= p) lo = mid;
          else hi = mid - 1;
        }
        c0 = a.sims[lo].cell_off;
      }
      record_move(a.aux->L, p, a.nc_all, 0, c0);
    }
    return;
  }

  const int cnt = a.ovf ? a.ovf[a.par] : a.ntiles;
  for (int qt = blockIdx.x; qt < cnt; qt += a.nrb) {
    const int tile = a.ovf ? a.ovf[2 + a.par * a.ntiles + qt] : qt;
    __syncthreads();  // the previous tile's LDS
    const TileCtx<T> c = tile_ctx<T, B>(a, tile, false);
    const Grid2<T> g = c.g;
    const int nx = g.nx, ny = g.ny;
    const int ti = c.ti, tj = c.tj;
    const Rgn R0{ti * kTW, min(ti * kTW + kTW, nx), tj * kTH, min(tj * kTH + kTH, ny)};
    const Rgn R1 = grow(R0, 1, nx, ny);
    const Rgn RH = grow(R0, H, nx, ny);
    const Rgn RN = USF ? R1 : R0;  // cells whose nodes get kinematic velocities
    const int h0 = R0.j1 - R0.j0, h1 = R1.j1 - R1.j0, hH = RH.j1 - RH.j0;

    // slot segments of the rows of R0 (table 0), R1 (1) and RH (2): one wave
    // per table, lane r = row r, prefix by shuffles
    {
      const int q = t >> 6, r = t & 63;
      if (q < 3) {
        const Rgn R = q == 0 ? R0 : (q == 1 ? R1 : RH);
        int len = 0;
        if (r < R.j1 - R.j0) {
          const int k = c.coff + (R.j0 + r) * nx + R.i0;
          const int s0 = cs[k];
          segA[q][r] = s0;
          len = cs[k + (R.i1 - R.i0)] - s0;
        }
  #pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(len, o, 64);
          if (r >= o) len += y;
        }
        if (r < kSegRows) segP[q][r + 1] = len;
        if (r == 0) segP[q][0] = 0;
      } else if (q == 3) {
        // cell centres and scales 2/(X1 - X0) of the columns and rows of RH
        // (natural coordinates and inverse Jacobians, particle.py:180-206)
        if (r < RH.i1 - RH.i0) {
          const T X0 = g.X(RH.i0 + r), X1 = g.X(RH.i0 + r + 1);
          colC[r] = (X0 + X1) / T(2);
          colS[r] = T(2) / (X1 - X0);
        }
        if (r < hH) {
          const T Y0 = g.Y(RH.j0 + r), Y1 = g.Y(RH.j0 + r + 1);
          rowC[r] = (Y0 + Y1) / T(2);
          rowS[r] = T(2) / (Y1 - Y0);
        }
      }
    }
    __syncthreads();
    const int SH = segP[2][hH], S1 = segP[1][h1];
    const int wK = RH.i1 - RH.i0;
    const int wKN = RN.i1 - RN.i0 + 1, wTN = R0.i1 - R0.i0 + 1;

    // ================= rounds: stage slots, reload per phase =====================
    constexpr int RPT = kRound / kFT;  // slots per thread per round
    // ---- K ----
    {
      int kv0[CK], kv1[CK];
      cell_ranges<CK>(RH, nx, c.coff, cs, segA[2], segP[2], kv0, kv1);
      T acc[CK][4 * NQ_KIN];
  #pragma unroll
      for (int u = 0; u < CK; ++u)
  #pragma unroll
        for (int v = 0; v < 4 * NQ_KIN; ++v) acc[u][v] = T(0);
      for (int v0 = 0; v0 < SH; v0 += kRound) {
        const int vend = min(v0 + kRound, SH);
        int pp[RPT];
  #pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int v = v0 + t + j * kFT;
          int p = -1;
          if (v < vend) {
            const int r = seg_row(segP[2], hH, v);
            const int s = segA[2][r] + (v - segP[2][r]);
            p = ident ? s : perm[s];
          }
          pp[j] = p;
        }
        T x[RPT], y[RPT], m[RPT], vx[RPT], vy[RPT];
  #pragma unroll
        for (int j = 0; j < RPT; ++j) {
          if (pp[j] < 0) continue;
          const int p = pp[j];
          x[j] = SR(DMPM_P_X, p);
          y[j] = SR(DMPM_P_X + 1, p);
          m[j] = PB(DMPM_P_MASS, p);
          vx[j] = SR(DMPM_P_V, p);
          vy[j] = SR(DMPM_P_V + 1, p);
        }
  #pragma unroll
        for (int j = 0; j < RPT; ++j) {
          if (pp[j] < 0) continue;
          const int q = t + j * kFT;
          const int e = element_id<T, 2>(g.geo(), x[j], y[j]);
          const int I = e % nx, J = e / nx;
          const T xi = natural(x[j], g.X(I), g.X(I + 1));
          const T eta = natural(y[j], g.Y(J), g.Y(J + 1));
          st[q] = xi;
          st[kRound + q] = eta;
          st[2 * kRound + q] = m[j];
          st[3 * kRound + q] = m[j] * vx[j];
          st[4 * kRound + q] = m[j] * vy[j];
          if (a.temps && in_rgn(R0, I, J)) {
            PB_W(DMPM_P_XI, pp[j], xi);
            PB_W(DMPM_P_XI + 1, pp[j], eta);
          }
        }
        __syncthreads();
  #pragma unroll
        for (int u = 0; u < CK; ++u)
          kin_cell(st, kRound, max(kv0[u], v0) - v0, min(kv1[u], vend) - v0, acc[u]);
        __syncthreads();
      }
      NodePre<NK> kpre;
      node_pre<T, NK>(a, c, RN, kpre, nullptr);
      node_pre_mask<T, NK>(a, kpre);
      T nacc[NK][NQ_KIN];
      node_sums<T, NQ_KIN, CK, NK>(st, acc, RH, RN, nx, ny, nacc);
      kin_epilogue<T, NK>(a, c, nacc, kpre, RN, R0, vk, nm, NKN);
      __syncthreads();
    }
    // ---- S + F ----
    {
      int fv0[CF], fv1[CF];
      cell_ranges<CF>(R1, nx, c.coff, cs, segA[1], segP[1], fv0, fv1);
      T acc[CF][4 * NQ_FORCE];
  #pragma unroll
      for (int u = 0; u < CF; ++u)
  #pragma unroll
        for (int v = 0; v < 4 * NQ_FORCE; ++v) acc[u][v] = T(0);
      const int wF = R1.i1 - R1.i0;
    const SmallDiv dwF = small_div(wF);
      for (int v0 = 0; v0 < S1; v0 += kRound) {
        const int vend = min(v0 + kRound, S1);
        int pp[RPT];
  #pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int v = v0 + t + j * kFT;
          int p = -1;
          if (v < vend) {
            const int r = seg_row(segP[1], h1, v);
            const int s = segA[1][r] + (v - segP[1][r]);
            p = ident ? s : perm[s];
          }
          pp[j] = p;
        }
  #pragma unroll
        for (int j = 0; j < RPT; ++j) {
          if (pp[j] < 0) continue;
          const int p = pp[j];
          const int q = t + j * kFT;
          const T x = SR(DMPM_P_X, p), y = SR(DMPM_P_X + 1, p);
          const int e = element_id<T, 2>(g.geo(), x, y);
          const int I = e % nx, J = e / nx;
          const T X0 = g.X(I), X1 = g.X(I + 1), Y0 = g.Y(J), Y1 = g.Y(J + 1);
          const T xi = natural(x, X0, X1);
          const T eta = natural(y, Y0, Y1);
          T V, sxx, syy, sxy;
          if (USF) {
            const int Ic = min(max(I, RN.i0), RN.i1 - 1), Jc = min(max(J, RN.j0), RN.j1 - 1);
            const int ll = (Ic - RN.i0) + (Jc - RN.j0) * wKN;
            const int ids[4] = {ll, ll + 1, ll + wKN + 1, ll + wKN};
            T nvx[4], nvy[4];
  #pragma unroll
            for (int k = 0; k < 4; ++k) {
              nvx[k] = vk[ids[k]];
              nvy[k] = vk[NKN + ids[k]];
            }
            T ix, iy;
            q4_invJ(xi, eta, X0, X1, Y0, Y1, ix, iy);
            StressUpd<T> u;
            stress_update(c.mat, a.mat.per_E ? PB(DMPM_P_YOUNGS, p) : T(0), ix, iy, xi, eta, nvx,
                          nvy, a.dt, u);
            sxx = SR(DMPM_P_STRESS + 0, p) + u.ds[0];
            syy = SR(DMPM_P_STRESS + 1, p) + u.ds[1];
            sxy = SR(DMPM_P_STRESS + 3, p) + u.ds[3];
            V = SR(DMPM_P_VOLUME, p) * u.f;
            if (in_rgn(R0, I, J)) {
              DR_W(DMPM_P_STRAIN + 0, p, SR(DMPM_P_STRAIN + 0, p) + u.d[0]);
              DR_W(DMPM_P_STRAIN + 1, p, SR(DMPM_P_STRAIN + 1, p) + u.d[1]);
              DR_W(DMPM_P_STRAIN + 3, p, SR(DMPM_P_STRAIN + 3, p) + u.d[3]);
              DR_W(DMPM_P_STRESS + 0, p, sxx);
              DR_W(DMPM_P_STRESS + 1, p, syy);
              DR_W(DMPM_P_STRESS + 2, p, SR(DMPM_P_STRESS + 2, p) + u.ds[2]);
              DR_W(DMPM_P_STRESS + 3, p, sxy);
              DR_W(DMPM_P_VOLUME, p, V);
              DR_W(DMPM_P_DENSITY, p, SR(DMPM_P_DENSITY, p) / u.f);
              DR_W(DMPM_P_VOLSTRAIN_C, p, SR(DMPM_P_VOLSTRAIN_C, p) + u.dvol);
              if (a.temps) {
                for (int k = 0; k < 6; ++k) {
                  PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
                  PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
                }
                PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
              }
            }
          } else {
            V = SR(DMPM_P_VOLUME, p);
            sxx = SR(DMPM_P_STRESS + 0, p);
            syy = SR(DMPM_P_STRESS + 1, p);
            sxy = SR(DMPM_P_STRESS + 3, p);
          }
          st[q] = xi;
          st[kRound + q] = eta;
          st[2 * kRound + q] = PB(DMPM_P_MASS, p);
          st[3 * kRound + q] = -V;
          st[4 * kRound + q] = sxx;
          st[5 * kRound + q] = syy;
          st[6 * kRound + q] = sxy;
        }
        __syncthreads();
  #pragma unroll
        for (int u = 0; u < CF; ++u) {
          const int lc = t + u * kFT;
          if (lc >= wF * h1) continue;
          const int jr = sdiv(dwF, lc);
        const int I = R1.i0 + (lc - jr * wF), J = R1.j0 + jr;
          force_cell(st, kRound, max(fv0[u], v0) - v0, min(fv1[u], vend) - v0, g.X(I), g.X(I + 1),
                     g.Y(J), g.Y(J + 1), a.gx, a.gy, acc[u]);
        }
        __syncthreads();
      }
      NodePre<NF> fpre;
      T fd[NF][2];
      node_pre<T, NF>(a, c, R0, fpre, fd);
      node_pre_mask<T, NF>(a, fpre);
      T nacc[NF][NQ_FORCE];
      node_sums<T, NQ_FORCE, CF, NF>(st, acc, R1, R0, nx, ny, nacc);
      force_epilogue<T, NF>(a, c, nacc, fpre, fd, R0, RN, vk, nm, NKN, nv, na, NTN);
      __syncthreads();
    }
    // ---- G ----
    const int S0 = segP[0][h0];
    for (int v = t; v < S0; v += kFT) {
      const int r = seg_row(segP[0], h0, v);
      const int s = segA[0][r] + (v - segP[0][r]);
      const int p = ident ? s : perm[s];
      const T x = SR(DMPM_P_X, p), y = SR(DMPM_P_X + 1, p);
      const int e = element_id<T, 2>(g.geo(), x, y);
      const int I = e % nx, J = e / nx;
      const T X0 = g.X(I), X1 = g.X(I + 1), Y0 = g.Y(J), Y1 = g.Y(J + 1);
      const T xi = natural(x, X0, X1);
      const T eta = natural(y, Y0, Y1);
      const int Ic = min(max(I, R0.i0), R0.i1 - 1), Jc = min(max(J, R0.j0), R0.j1 - 1);
      const int ll = (Ic - R0.i0) + (Jc - R0.j0) * wTN;
      const int ids[4] = {ll, ll + 1, ll + wTN + 1, ll + wTN};
      T N[4];
      q4_N4(xi, eta, N);
      T gvx[4], gvy[4];
      T nvx = T(0), nvy = T(0), nax = T(0), nay = T(0);
      for (int k = 0; k < 4; ++k) {
        gvx[k] = nv[ids[k]];
        gvy[k] = nv[NTN + ids[k]];
        nvx = nvx + N[k] * gvx[k];
        nvy = nvy + N[k] * gvy[k];
        nax = nax + N[k] * na[ids[k]];
        nay = nay + N[k] * na[NTN + ids[k]];
      }
      const T pvx = a.vu ? nvx : SR(DMPM_P_V, p) + nax * a.dt;
      const T pvy = a.vu ? nvy : SR(DMPM_P_V + 1, p) + nay * a.dt;
      const T x1 = x + nvx * a.dt, y1 = y + nvy * a.dt;
      DR_W(DMPM_P_V, p, pvx);
      DR_W(DMPM_P_V + 1, p, pvy);
      DR_W(DMPM_P_X, p, x1);
      DR_W(DMPM_P_X + 1, p, y1);
      if (a.temps) {
        const T m = PB(DMPM_P_MASS, p);
        PB_W(DMPM_P_MOMENTUM, p, m * pvx);
        PB_W(DMPM_P_MOMENTUM + 1, p, m * pvy);
      }
      if (!USF) {
        T ix, iy;
        q4_invJ(xi, eta, X0, X1, Y0, Y1, ix, iy);
        StressUpd<T> u;
        stress_update(c.mat, a.mat.per_E ? PB(DMPM_P_YOUNGS, p) : T(0), ix, iy, xi, eta, gvx, gvy,
                      a.dt, u);
        DR_W(DMPM_P_STRAIN + 0, p, SR(DMPM_P_STRAIN + 0, p) + u.d[0]);
        DR_W(DMPM_P_STRAIN + 1, p, SR(DMPM_P_STRAIN + 1, p) + u.d[1]);
        DR_W(DMPM_P_STRAIN + 3, p, SR(DMPM_P_STRAIN + 3, p) + u.d[3]);
        for (int k = 0; k < 4; ++k) DR_W(DMPM_P_STRESS + k, p, SR(DMPM_P_STRESS + k, p) + u.ds[k]);
        DR_W(DMPM_P_VOLUME, p, SR(DMPM_P_VOLUME, p) * u.f);
        DR_W(DMPM_P_DENSITY, p, SR(DMPM_P_DENSITY, p) / u.f);
        DR_W(DMPM_P_VOLSTRAIN_C, p, SR(DMPM_P_VOLSTRAIN_C, p) + u.dvol);
        if (a.temps) {
          for (int k = 0; k < 6; ++k) {
            PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
            PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
          }
          PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
        }
      }
      const int e1 = element_id<T, 2>(g.geo(), x1, y1);
      const int k1 = (e1 >= 0 && e1 < g.ncells) ? c.coff + e1 : a.nc_all;
      if (k1 != c.coff + e) record_move(a.aux->L, p, c.coff + e, e1, k1);
    }
  }
}

// ---- split step: three launches per step --------------------------------------
//
// The step of k_step_fast in three launches.  No workgroup holds a tile's
// whole particle state, and only the column and the row of cells left of and
// below a tile are recomputed (1.13x the tile's particles in K, S and F,
// against 1.56x in K and 1.27x in S and F with the two-ring halo):
//   k_split_kin    node tile: P2G of mass and momentum over the cells of the
//                  tile's nodes; nodal mass and kinematic velocity of the
//                  nodes the tile owns
//   k_split_force  node tile: USF stress update of the particles of the same
//                  cells (the tile's own cells' results written), P2G of body
//                  and internal forces, nodal update of the owned nodes
//   k_split_g2p    cell tile: G2P of the tile's particles (USL: then their
//                  stress update) with the cell changes recorded for the next
//                  re-sort; kTailBlocks more workgroups handle the particles
//                  outside the mesh
// Tile (ti, tj) owns the cells R0 = [16 ti, 16 ti + 16) x [16 tj, 16 tj + 16)
// (clipped to the mesh) and the nodes (i, j) with i in [R0.i0, R0.i1), plus
// i = nx in the last tile column, and likewise j.  The cells of those nodes
// are RC = R0 grown by one cell to the left and below.  Every launch reads
// the particle state from src and writes dst, so halo reads see the step's
// input.  Slots are staged in rounds of NT * RPT (one round unless a tile is
// crowded).  In each round the threads that own the region's cells write
// every slot's cell into an LDS table, so a slot finds its cell with one LDS
// read instead of a search (the re-sort keyed the slot by the element search
// of its position, so this is the cell the search gives).  Cell sums run in
// slot order and node sums in node_gather's order with the expressions of
// the other paths, so the results are bit-identical to theirs.

// Node-table values (constraint row, damping force) of this thread's nodes of
// the half-open node region NR (w-th node: index t + w * NT).
template <typename T, int NPT, int NT>
__device__ __forceinline__ void node_pre_nt(const StepArgs<T>& a, const TileCtx<T>& c,
                                            const Rgn& NR, NodePre<NPT>& q, T (*fd)[2],
                                            T* mass = nullptr) {
  const int t = threadIdx.x, nx = c.g.nx;
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = min(t + w * NT, nn - 1);
    const int jr = sdiv(dwn, ln);
    const int nd = c.noff + (NR.j0 + jr) * (nx + 1) + NR.i0 + (ln - jr * wn);
    q.cr[w] = a.cidx ? a.cidx[nd] : -1;
    if (fd) {
      fd[w][0] = bld<T>(a.fdsrc, a.nbytes, (uint32_t)DMPM_N_FDAMP * a.nld32, (uint32_t)nd);
      fd[w][1] = bld<T>(a.fdsrc, a.nbytes, (uint32_t)(DMPM_N_FDAMP + 1) * a.nld32, (uint32_t)nd);
    }
    if (mass) mass[w] = NB(DMPM_N_MASS, nd);
  }
#pragma unroll
  for (int w = 0; w < NPT; ++w) q.mask[w] = q.cr[w] >= 0 ? a.cmask[q.cr[w]] : 0;
}

// Workgroup size and slots per thread and round of the split kernels.
constexpr int kSKinT = 512, kSKinR = 3;
constexpr int kSForT = 512, kSForR = 3;
constexpr int kSG2pT = 256, kSG2pR = 5;
constexpr int kSCells = (kTW + 1) * (kTH + 1);  // cells of RC (at most)
constexpr int kSNodes = (kTW + 1) * (kTH + 1);  // owned nodes (at most)
constexpr int kSVN = (kTW + 2) * (kTH + 2);     // nodes of RC's cells (at most)

// Kinematic launch: P2G of mass and momentum over RC (k_cell_kin's
// expressions), then nodal mass, momentum and velocity of the owned nodes
// (k_node_kin's), written to the node block's mass row and to vkin.
template <typename T, bool B, int NT, int RPT, int WPE = 1>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_split_kin(
    StepArgs<T> a) {
  constexpr int CAP = NT * RPT;
  constexpr int CC = (kSCells + NT - 1) / NT;
  constexpr int NPT = (kSNodes + NT - 1) / NT;
  static_assert(4 * NQ_KIN * kSCells <= 5 * CAP, "node-sum slices");
  static_assert(NT >= 128, "two waves for the tables");
  __shared__ T st[5 * CAP];  // xi, eta, m, m vx, m vy; the node sums alias it
  __shared__ uint16_t tab[CAP];
  __shared__ T colC[kTW + 1], colS[kTW + 1], rowC[kTH + 1], rowS[kTH + 1];
  __shared__ int segA[kTH + 1], segP[kTH + 2];
  const TileCtx<T> c = tile_ctx<T, B>(a, xcd_pos(blockIdx.x, a.ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx, ny = g.ny, t = threadIdx.x;
  const int* __restrict__ cs = a.sr.cs();
  const int* __restrict__ perm = a.sr.perm();
  const int ident = a.sr.flags[1];
  const int ti = c.ti, tj = c.tj;
  const Rgn R0{ti * kTW, min(ti * kTW + kTW, nx), tj * kTH, min(tj * kTH + kTH, ny)};
  const Rgn RC{max(R0.i0 - 1, 0), R0.i1, max(R0.j0 - 1, 0), R0.j1};
  const Rgn NR = owned_nodes(R0, nx, ny);
  const int hC = RC.j1 - RC.j0;
  int cr_[CC], ccol[CC], cs0[CC], cs1[CC];
  cell_slots_rc<CC, NT>(RC, nx, c.coff, cs, cr_, ccol, cs0, cs1);
  if (t < 64) row_segments(RC, nx, c.coff, cs, segA, segP);
  else if (t < 128) cell_tables(g, RC, colC, colS, rowC, rowS);
  NodePre<NPT> pre;
  node_pre_nt<T, NPT, NT>(a, c, NR, pre, nullptr);
  __syncthreads();
  const int S = segP[hC];
  int cv0[CC], cv1[CC];
  cell_ranges_of<CC>(cr_, cs0, cs1, segA, segP, cv0, cv1);
  T acc[CC][4 * NQ_KIN];
#pragma unroll
  for (int u = 0; u < CC; ++u)
#pragma unroll
    for (int v = 0; v < 4 * NQ_KIN; ++v) acc[u][v] = T(0);
  for (int v0 = 0; v0 < S; v0 += CAP) {
    const int vend = min(v0 + CAP, S);
    fill_slot_cells<CC>(tab, cr_, ccol, cv0, cv1, v0, vend);
    __syncthreads();
    int pp[RPT], code[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      const bool live = v0 + q < vend;
      code[j] = live ? (int)tab[q] : 0;
      const int r = code[j] >> 8;
      const int s = live ? segA[r] + (v0 + q - segP[r]) : 0;
      pp[j] = ident ? s : perm[s];
    }
    T x[RPT], y[RPT], vx[RPT], vy[RPT], m[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int p = pp[j];
      x[j] = SR(DMPM_P_X, p);
      y[j] = SR(DMPM_P_X + 1, p);
      vx[j] = SR(DMPM_P_V, p);
      vy[j] = SR(DMPM_P_V + 1, p);
      m[j] = PB(DMPM_P_MASS, p);
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      if (v0 + q >= vend) continue;
      // the slot's cell is the element search's result for this position
      // (the re-sort keyed the slot by it)
      const int r = code[j] >> 8, ci = code[j] & 255;
      st[q] = (x[j] - colC[ci]) * colS[ci];
      st[CAP + q] = (y[j] - rowC[r]) * rowS[r];
      st[2 * CAP + q] = m[j];
      st[3 * CAP + q] = m[j] * vx[j];
      st[4 * CAP + q] = m[j] * vy[j];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CC; ++u)
      kin_cell(st, CAP, max(cv0[u], v0) - v0, min(cv1[u], vend) - v0, acc[u]);
    __syncthreads();
  }
  T nacc[NPT][NQ_KIN];
  node_sums_nt<T, NQ_KIN, CC, NPT, NT>(st, acc, RC, NR, nx, ny, nacc);
  // kinematic nodal epilogue (k_node_kin / kin_epilogue)
  const int nnodes = a.nn_all;
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = t + w * NT;
    if (ln >= nn) continue;
    const int jr = sdiv(dwn, ln);
    const int nd = c.noff + (NR.j0 + jr) * (nx + 1) + NR.i0 + (ln - jr * wn);
    const int cr = pre.cr[w], mask = pre.mask[w];
    const T mm = nacc[w][0];
    const T p[2] = {snap(nacc[w][1]), snap(nacc[w][2])};
    NB_W(DMPM_N_MASS, nd, mm);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      T v;
      if (mm == T(0)) v = a.vold[(int64_t)d * a.vld + nd];
      else v = p[d] / mm;
      v = snap(v);
      if (mask & (1 << d)) v = T(a.aux->cval[2 * cr + d]);
      a.vkin[(int64_t)d * nnodes + nd] = v;
    }
  }
}

// Force launch: (USF) strain, volume and stress update of RC's particles from
// the kinematic nodal velocities (own cells' results written), P2G of body
// and internal forces over RC (k_cell_force's expressions), and the nodal
// update of the owned nodes (k_node_force's): node block rows, vnew.
template <typename T, bool USF, bool B, int NT, int RPT, int WPE = 1>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_split_force(
    StepArgs<T> a) {
  constexpr int CAP = NT * RPT;
  constexpr int CC = (kSCells + NT - 1) / NT;
  constexpr int NPT = (kSNodes + NT - 1) / NT;
  static_assert(4 * NQ_FORCE * kSCells <= 7 * CAP, "node-sum slices");
  static_assert(NT >= 128, "two waves for the tables");
  __shared__ T st[7 * CAP];  // xi, eta, m, -V, sxx, syy, sxy; the node sums alias it
  __shared__ uint16_t tab[CAP];
  __shared__ T vk[2 * kSVN];  // kinematic velocity at the nodes of RC's cells
  __shared__ T colC[kTW + 1], colS[kTW + 1], rowC[kTH + 1], rowS[kTH + 1];
  __shared__ int segA[kTH + 1], segP[kTH + 2];
  const TileCtx<T> c = tile_ctx<T, B>(a, xcd_pos(blockIdx.x, a.ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx, ny = g.ny, t = threadIdx.x;
  const int* __restrict__ cs = a.sr.cs();
  const int* __restrict__ perm = a.sr.perm();
  const int ident = a.sr.flags[1];
  const int ti = c.ti, tj = c.tj;
  const Rgn R0{ti * kTW, min(ti * kTW + kTW, nx), tj * kTH, min(tj * kTH + kTH, ny)};
  const Rgn RC{max(R0.i0 - 1, 0), R0.i1, max(R0.j0 - 1, 0), R0.j1};
  const Rgn NR = owned_nodes(R0, nx, ny);
  const int hC = RC.j1 - RC.j0, wC = RC.i1 - RC.i0;
  const int wV = wC + 1, nV = wV * (hC + 1);  // nodes of RC's cells
  const int oi = R0.i0 - RC.i0, oj = R0.j0 - RC.j0;  // first own column and row of RC
  const int nnodes = a.nn_all;
  int cr_[CC], ccol[CC], cs0[CC], cs1[CC];
  cell_slots_rc<CC, NT>(RC, nx, c.coff, cs, cr_, ccol, cs0, cs1);
  if (t < 64) row_segments(RC, nx, c.coff, cs, segA, segP);
  else if (t < 128) cell_tables(g, RC, colC, colS, rowC, rowS);
  if (USF) {
    const SmallDiv dwv = small_div(wV);
    for (int ln = t; ln < nV; ln += NT) {
      const int jr = sdiv(dwv, ln);
      const int nd = c.noff + (RC.j0 + jr) * (nx + 1) + RC.i0 + (ln - jr * wV);
      vk[ln] = a.vkin[nd];
      vk[kSVN + ln] = a.vkin[(int64_t)nnodes + nd];
    }
  }
  NodePre<NPT> pre;
  T fd[NPT][2], nmass[NPT];
  node_pre_nt<T, NPT, NT>(a, c, NR, pre, fd, nmass);
  __syncthreads();
  const int S = segP[hC];
  int cv0[CC], cv1[CC];
  cell_ranges_of<CC>(cr_, cs0, cs1, segA, segP, cv0, cv1);
  T acc[CC][4 * NQ_FORCE];
#pragma unroll
  for (int u = 0; u < CC; ++u)
#pragma unroll
    for (int v = 0; v < 4 * NQ_FORCE; ++v) acc[u][v] = T(0);
  for (int v0 = 0; v0 < S; v0 += CAP) {
    const int vend = min(v0 + CAP, S);
    fill_slot_cells<CC>(tab, cr_, ccol, cv0, cv1, v0, vend);
    __syncthreads();
    int pp[RPT], code[RPT];
    unsigned own = 0;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      const bool live = v0 + q < vend;
      code[j] = live ? (int)tab[q] : 0;
      const int r = code[j] >> 8, ci = code[j] & 255;
      const int s = live ? segA[r] + (v0 + q - segP[r]) : 0;
      if (live && r >= oj && ci >= oi) own |= 1u << j;
      pp[j] = ident ? s : perm[s];
    }
    T x[RPT], y[RPT], m[RPT], sxx[RPT], syy[RPT], sxy[RPT], V[RPT], E[RPT];
    T szz[RPT], ex[RPT], ey[RPT], gxy[RPT], rho[RPT], evc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int p = pp[j];
      x[j] = SR(DMPM_P_X, p);
      y[j] = SR(DMPM_P_X + 1, p);
      m[j] = PB(DMPM_P_MASS, p);
      sxx[j] = SR(DMPM_P_STRESS + 0, p);
      syy[j] = SR(DMPM_P_STRESS + 1, p);
      sxy[j] = SR(DMPM_P_STRESS + 3, p);
      V[j] = SR(DMPM_P_VOLUME, p);
      if (USF) {
        const int po = (own & (1u << j)) ? p : 0;
        szz[j] = SR(DMPM_P_STRESS + 2, po);
        ex[j] = SR(DMPM_P_STRAIN + 0, po);
        ey[j] = SR(DMPM_P_STRAIN + 1, po);
        gxy[j] = SR(DMPM_P_STRAIN + 3, po);
        rho[j] = SR(DMPM_P_DENSITY, po);
        evc[j] = SR(DMPM_P_VOLSTRAIN_C, po);
      }
    }
    if (USF && a.mat.per_E)
#pragma unroll
      for (int j = 0; j < RPT; ++j) E[j] = PB(DMPM_P_YOUNGS, pp[j]);
    else
#pragma unroll
      for (int j = 0; j < RPT; ++j) E[j] = T(0);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      if (v0 + q >= vend) continue;
      const int rj = code[j] >> 8, ci = code[j] & 255;
      const T xi = (x[j] - colC[ci]) * colS[ci];
      const T eta = (y[j] - rowC[rj]) * rowS[rj];
      T Vn = V[j], sx = sxx[j], sy = syy[j], sq = sxy[j];
      if (USF) {
        const int ll = ci + rj * wV;
        const int ids[4] = {ll, ll + 1, ll + wV + 1, ll + wV};
        T nvx[4], nvy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          nvx[k] = vk[ids[k]];
          nvy[k] = vk[kSVN + ids[k]];
        }
        StressUpd<T> u;
        stress_update(c.mat, E[j], colS[ci], rowS[rj], xi, eta, nvx, nvy, a.dt, u);
        sx = sxx[j] + u.ds[0];
        sy = syy[j] + u.ds[1];
        sq = sxy[j] + u.ds[3];
        Vn = V[j] * u.f;
        if (own & (1u << j)) {
          const int p = pp[j];
          DR_W(DMPM_P_STRAIN + 0, p, ex[j] + u.d[0]);
          DR_W(DMPM_P_STRAIN + 1, p, ey[j] + u.d[1]);
          DR_W(DMPM_P_STRAIN + 3, p, gxy[j] + u.d[3]);
          DR_W(DMPM_P_STRESS + 0, p, sx);
          DR_W(DMPM_P_STRESS + 1, p, sy);
          DR_W(DMPM_P_STRESS + 2, p, szz[j] + u.ds[2]);
          DR_W(DMPM_P_STRESS + 3, p, sq);
          DR_W(DMPM_P_VOLUME, p, Vn);
          DR_W(DMPM_P_DENSITY, p, rho[j] / u.f);
          DR_W(DMPM_P_VOLSTRAIN_C, p, evc[j] + u.dvol);
          if (a.temps) {
            for (int k = 0; k < 6; ++k) {
              PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
              PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
            }
            PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
          }
        }
      }
      st[q] = xi;
      st[CAP + q] = eta;
      st[2 * CAP + q] = m[j];
      st[3 * CAP + q] = -Vn;
      st[4 * CAP + q] = sx;
      st[5 * CAP + q] = sy;
      st[6 * CAP + q] = sq;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CC; ++u) {
      if (cr_[u] < 0) continue;
      const int I = RC.i0 + ccol[u], J = RC.j0 + cr_[u];
      force_cell(st, CAP, max(cv0[u], v0) - v0, min(cv1[u], vend) - v0, g.X(I), g.X(I + 1),
                 g.Y(J), g.Y(J + 1), a.gx, a.gy, acc[u]);
    }
    __syncthreads();
  }
  T nacc[NPT][NQ_FORCE];
  node_sums_nt<T, NQ_FORCE, CC, NPT, NT>(st, acc, RC, NR, nx, ny, nacc);
  // force nodal epilogue (k_node_force / force_epilogue) of the owned nodes
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const SmallDiv dwn = small_div(wn);
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = t + w * NT;
    if (ln >= nn) continue;
    const int jr = sdiv(dwn, ln);
    const int i = NR.i0 + (ln - jr * wn), j = NR.j0 + jr;
    const int nd = c.noff + j * (nx + 1) + i;
    const int lk = (i - RC.i0) + (j - RC.j0) * wV;
    const int cr = pre.cr[w], mask = pre.mask[w];
    const T mm = nmass[w];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      T fe = nacc[w][d];
      if (a.has_nf) {
        const StepAux& x = *a.aux;
        const int qg = x.nfd[(int64_t)d * nnodes + nd];
        if (qg >= 0)
          for (int k = x.nf_off[qg]; k < x.nf_off[qg + 1]; ++k)
            fe = fe + fn_value<T>(x.fns, x.nf_fn[k], step_time(a.clock, a.kofs, a.t, a.dt)) *
                          T(x.nf_val[k]);
      }
      const T fi = nacc[w][2 + d];
      const T F = (fi + fe) + fd[w][d];
      T ac = nan_to_num(F / mm);
      const T vkd = USF ? vk[d * kSVN + lk] : a.vkin[(int64_t)d * nnodes + nd];
      T v = vkd + ac * a.dt;
      T pm = mm * v;
      v = snap(v);
      ac = snap(ac);
      if (mask & (1 << d)) {
        const T cv = T(a.aux->cval[2 * cr + d]);
        v = snap(cv);
        pm = mm * cv;
        ac = T(0);
      }
      if (a.last) {
        NB_W(DMPM_N_FEXT + d, nd, fe);
        NB_W(DMPM_N_FINT + d, nd, fi);
        NB_W(DMPM_N_V + d, nd, v);
        NB_W(DMPM_N_P + d, nd, pm);
        if (a.wfd) NB_W(DMPM_N_FDAMP + d, nd, fd[w][d]);
      }
      // (G2P reads the acceleration from the node block)
      NB_W(DMPM_N_A + d, nd, ac);
      a.vnew[(int64_t)d * nnodes + nd] = v;
    }
  }
}

// G2P launch over the tile's own particles (k_step_fast's G phase), with the
// nodal velocity and acceleration of the tile's nodes staged in LDS; USL then
// updates strain, volume and stress from the new nodal velocities.  Records
// the particles that change cell.  Workgroups from ntiles on handle the
// particles outside the mesh (as k_step_rounds does).
template <typename T, bool USF, bool B, int NT, int RPT, int WPE = 1>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_split_g2p(
    StepArgs<T> a) {
  constexpr int NTN = (kTW + 1) * (kTH + 1);
  constexpr int CAP = NT * RPT;
  constexpr int CC = (kTW * kTH + NT - 1) / NT;
  static_assert(NT >= 128, "two waves for the tables");
  __shared__ T nv[2 * NTN], na[2 * NTN];
  __shared__ uint16_t tab[CAP];
  __shared__ T colC[kTW], colS[kTW], rowC[kTH], rowS[kTH], colX[kTW + 1], rowY[kTH + 1];
  __shared__ int segA[kTH], segP[kTH + 1];
  const int t = threadIdx.x;
  const int* __restrict__ cs = a.sr.cs();
  const int* __restrict__ perm = a.sr.perm();
  const int ident = a.sr.flags[1];
  if ((int)blockIdx.x >= a.ntiles) {
    // particles outside the mesh: the state becomes NaN (as in k_g2p_tile)
    const T nan = qnan<T>();
    const int gs = kTailBlocks * NT;
    for (int s = cs[a.nc_all] + ((int)blockIdx.x - a.ntiles) * NT + t; s < a.n; s += gs) {
      const int p = ident ? s : perm[s];
      for (int k = 0; k < kDynRows; ++k) DR_W(kDyn[k], p, nan);
      for (int k = 0; k < kFixRows; ++k) PB_W(kFix[k], p, nan);
      if (a.temps) {
        PB_W(DMPM_P_XI, p, nan);
        PB_W(DMPM_P_XI + 1, p, nan);
        PB_W(DMPM_P_MOMENTUM, p, nan);
        PB_W(DMPM_P_MOMENTUM + 1, p, nan);
        for (int k = 0; k < 6; ++k) {
          PB_W(DMPM_P_STRAIN_RATE + k, p, nan);
          PB_W(DMPM_P_DSTRAIN + k, p, nan);
        }
        PB_W(DMPM_P_DVOLSTRAIN, p, nan);
      }
      // NaN positions search to element 0 (element.py:827-837) of the
      // particle's own mesh
      int c0 = 0;
      if constexpr (B) {
        int lo = 0, hi = a.nsims - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (a.sims[mid].part_off <= p) lo = mid;
          else hi = mid - 1;
        }
        c0 = a.sims[lo].cell_off;
      }
      record_move(a.aux->L, p, a.nc_all, 0, c0);
    }
    return;
  }
  const TileCtx<T> c = tile_ctx<T, B>(a, xcd_pos(blockIdx.x, a.ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx;
  const int ti = c.ti, tj = c.tj;
  const Rgn R0{ti * kTW, min(ti * kTW + kTW, nx), tj * kTH, min(tj * kTH + kTH, g.ny)};
  const int h0 = R0.j1 - R0.j0, w0 = R0.i1 - R0.i0;
  const int wTN = w0 + 1, nTN = wTN * (h0 + 1);
  const int nnodes = a.nn_all;
  int cr_[CC], ccol[CC], cs0[CC], cs1[CC];
  cell_slots_rc<CC, NT>(R0, nx, c.coff, cs, cr_, ccol, cs0, cs1);
  if (t < 64) {
    row_segments(R0, nx, c.coff, cs, segA, segP);
  } else if (t < 128) {
    cell_tables(g, R0, colC, colS, rowC, rowS);
    // node coordinates of the tile's columns and rows (the cell-change test)
    const int r = t - 64;
    if (r <= w0) colX[r] = g.X(R0.i0 + r);
    if (r <= h0) rowY[r] = g.Y(R0.j0 + r);
  }
  {
    const SmallDiv dw = small_div(wTN);
    for (int ln = t; ln < nTN; ln += NT) {
      const int jr = sdiv(dw, ln);
      const int nd = c.noff + (R0.j0 + jr) * (nx + 1) + R0.i0 + (ln - jr * wTN);
      nv[ln] = a.vnew[nd];
      nv[NTN + ln] = a.vnew[(int64_t)nnodes + nd];
      na[ln] = NB(DMPM_N_A, nd);
      na[NTN + ln] = NB(DMPM_N_A + 1, nd);
    }
  }
  __syncthreads();
  const int S = segP[h0];
  int cv0[CC], cv1[CC];
  cell_ranges_of<CC>(cr_, cs0, cs1, segA, segP, cv0, cv1);
  for (int v0 = 0; v0 < S; v0 += CAP) {
    const int vend = min(v0 + CAP, S);
    if (v0 > 0) __syncthreads();  // the previous round's table reads
    fill_slot_cells<CC>(tab, cr_, ccol, cv0, cv1, v0, vend);
    __syncthreads();
    int pp[RPT], code[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      const bool live = v0 + q < vend;
      code[j] = live ? (int)tab[q] : 0;
      const int r = code[j] >> 8;
      const int s = live ? segA[r] + (v0 + q - segP[r]) : 0;
      pp[j] = ident ? s : perm[s];
    }
    T x[RPT], y[RPT], vx[RPT], vy[RPT], m[RPT];
    T sxx[RPT], syy[RPT], szz[RPT], sxy[RPT], V[RPT], E[RPT];
    T ex[RPT], ey[RPT], gxy[RPT], rho[RPT], evc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int p = pp[j];
      x[j] = SR(DMPM_P_X, p);
      y[j] = SR(DMPM_P_X + 1, p);
      vx[j] = SR(DMPM_P_V, p);
      vy[j] = SR(DMPM_P_V + 1, p);
      if (!USF) {
        sxx[j] = SR(DMPM_P_STRESS + 0, p);
        syy[j] = SR(DMPM_P_STRESS + 1, p);
        szz[j] = SR(DMPM_P_STRESS + 2, p);
        sxy[j] = SR(DMPM_P_STRESS + 3, p);
        V[j] = SR(DMPM_P_VOLUME, p);
        ex[j] = SR(DMPM_P_STRAIN + 0, p);
        ey[j] = SR(DMPM_P_STRAIN + 1, p);
        gxy[j] = SR(DMPM_P_STRAIN + 3, p);
        rho[j] = SR(DMPM_P_DENSITY, p);
        evc[j] = SR(DMPM_P_VOLSTRAIN_C, p);
      }
    }
    if (a.temps)
#pragma unroll
      for (int j = 0; j < RPT; ++j) m[j] = PB(DMPM_P_MASS, pp[j]);
    if (!USF) {
      if (a.mat.per_E)
#pragma unroll
        for (int j = 0; j < RPT; ++j) E[j] = PB(DMPM_P_YOUNGS, pp[j]);
      else
#pragma unroll
        for (int j = 0; j < RPT; ++j) E[j] = T(0);
    }
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      if (v0 + t + j * NT >= vend) continue;
      const int p = pp[j];
      const int rj = code[j] >> 8, ci = code[j] & 255;
      const T xi = (x[j] - colC[ci]) * colS[ci];
      const T eta = (y[j] - rowC[rj]) * rowS[rj];
      if (a.temps) {
        PB_W(DMPM_P_XI, p, xi);
        PB_W(DMPM_P_XI + 1, p, eta);
      }
      const int ll = ci + rj * wTN;
      const int ids[4] = {ll, ll + 1, ll + wTN + 1, ll + wTN};
      T N[4];
      q4_N4(xi, eta, N);
      T gvx[4], gvy[4];
      T nvx = T(0), nvy = T(0), nax = T(0), nay = T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gvx[k] = nv[ids[k]];
        gvy[k] = nv[NTN + ids[k]];
        nvx = nvx + N[k] * gvx[k];
        nvy = nvy + N[k] * gvy[k];
        nax = nax + N[k] * na[ids[k]];
        nay = nay + N[k] * na[NTN + ids[k]];
      }
      const T pvx = a.vu ? nvx : vx[j] + nax * a.dt;
      const T pvy = a.vu ? nvy : vy[j] + nay * a.dt;
      const T x1 = x[j] + nvx * a.dt, y1 = y[j] + nvy * a.dt;
      DR_W(DMPM_P_V, p, pvx);
      DR_W(DMPM_P_V + 1, p, pvy);
      DR_W(DMPM_P_X, p, x1);
      DR_W(DMPM_P_X + 1, p, y1);
      if (a.temps) {
        PB_W(DMPM_P_MOMENTUM, p, m[j] * pvx);
        PB_W(DMPM_P_MOMENTUM + 1, p, m[j] * pvy);
      }
      if (!USF) {
        StressUpd<T> u;
        stress_update(c.mat, E[j], colS[ci], rowS[rj], xi, eta, gvx, gvy, a.dt, u);
        DR_W(DMPM_P_STRAIN + 0, p, ex[j] + u.d[0]);
        DR_W(DMPM_P_STRAIN + 1, p, ey[j] + u.d[1]);
        DR_W(DMPM_P_STRAIN + 3, p, gxy[j] + u.d[3]);
        DR_W(DMPM_P_STRESS + 0, p, sxx[j] + u.ds[0]);
        DR_W(DMPM_P_STRESS + 1, p, syy[j] + u.ds[1]);
        DR_W(DMPM_P_STRESS + 2, p, szz[j] + u.ds[2]);
        DR_W(DMPM_P_STRESS + 3, p, sxy[j] + u.ds[3]);
        DR_W(DMPM_P_VOLUME, p, V[j] * u.f);
        DR_W(DMPM_P_DENSITY, p, rho[j] / u.f);
        DR_W(DMPM_P_VOLSTRAIN_C, p, evc[j] + u.dvol);
        if (a.temps) {
          for (int k = 0; k < 6; ++k) {
            PB_W(DMPM_P_STRAIN_RATE + k, p, u.rr[k]);
            PB_W(DMPM_P_DSTRAIN + k, p, u.d[k]);
          }
          PB_W(DMPM_P_DVOLSTRAIN, p, u.dvol);
        }
      }
      // still inside its cell: no element search (the same element id)
      if (x1 >= colX[ci] && x1 < colX[ci + 1] && y1 >= rowY[rj] && y1 < rowY[rj + 1]) continue;
      const int I = R0.i0 + ci, J = R0.j0 + rj;
      const int k0 = c.coff + J * nx + I;
      const int e1 = element_id<T, 2>(g.geo(), x1, y1);
      const int k1 = (e1 >= 0 && e1 < g.ncells) ? c.coff + e1 : a.nc_all;
      if (k1 != k0) record_move(a.aux->L, p, k0, e1, k1);
    }
  }
}

#undef SR
#undef PB
#undef NB
#undef DR_W
#undef PB_W
#undef NB_W

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
  int* ovf;  // optional: the one-launch step's overflow counters, zeroed by stage 4
  int* subT;  // [nsub] sub-chunk of 1 << kSubShift cells has recorded changes (INC)
  int nsub;
};

__device__ __forceinline__ int sort_mode(const SortArgs& a) {
  const int* f = a.L.flags;
  const int nmov = f[5];
  if (f[0] || f[2] || f[6] || f[10] || nmov > a.L.movcap) return SORT_FULL;
  return nmov > 0 ? SORT_INCREMENTAL : SORT_NONE;
}

__device__ __forceinline__ int blk_excl_scan(int v, int* sh, int& total) {
  // inclusive scan within each wave by shuffles, then the wave totals
  static_assert(kThreads % 64 == 0 && kThreads / 64 <= 64, "waves per workgroup");
  constexpr int NW = kThreads / 64;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const int u = sh[i];
    before += i < w ? u : 0;
    all += u;
  }
  __syncthreads();  // sh is reused by the caller
  total = all;
  return before + x - v;
}

// out[c] = base(c) + (sum of chunk totals before this chunk) + exclusive
// prefix of val over the chunk, for the cells of chunk blockIdx.x (a run of
// chunk/256 consecutive cells per thread); the last chunk also writes
// out[nbk] = total.  The loads of a thread's cells, and of the chunk totals,
// are issued together.
template <typename V, typename B>
__device__ __forceinline__ void chunk_scan(int nbk, int nchunks, int shift,
                                           const int* __restrict__ tot, V val, B base,
                                           int* __restrict__ out, int total, int* sh) {
  const int b = blockIdx.x, tid = threadIdx.x;
  int pre = 0;
  {
    int k = tid;
    for (; k + 3 * kThreads < b; k += 4 * kThreads)
      pre += (tot[k] + tot[k + kThreads]) + (tot[k + 2 * kThreads] + tot[k + 3 * kThreads]);
    for (; k < b; k += kThreads) pre += tot[k];
  }
  int red;
  blk_excl_scan(pre, sh, red);
  const int per = (1 << shift) / kThreads;
  const int c0 = (b << shift) + tid * per, c1 = min(c0 + per, nbk);
  constexpr int kPer = (1 << kMinChunkShift) / kThreads;
  if (per == kPer) {
    int v[kPer], bs[kPer], sum = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const bool in = c0 + u < c1;
      v[u] = in ? val(c0 + u) : 0;
      bs[u] = in ? base(c0 + u) : 0;
      sum += v[u];
    }
    int bt;
    int ex = red + blk_excl_scan(sum, sh, bt);
#pragma unroll
    for (int u = 0; u < kPer; ++u)
      if (c0 + u < c1) {
        out[c0 + u] = bs[u] + ex;
        ex += v[u];
      }
  } else {
    int sum = 0;
    for (int c = c0; c < c1; ++c) sum += val(c);
    int bt;
    int ex = red + blk_excl_scan(sum, sh, bt);
    for (int c = c0; c < c1; ++c) {
      out[c] = base(c) + ex;
      ex += val(c);
    }
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
    const int c0 = (int)blockIdx.x << a.L.chunk_shift;
    if (a.L.chunk_shift == kMinChunkShift) {
      // new offsets as chunk_scan computes them, and the chunk's sub-chunks
      // with recorded changes (stage 2 moves the others whole), from one load
      // of each cell's counts: 4 cells per thread, OR over the 16 threads of
      // a sub-chunk
      static_assert((1 << kMinChunkShift) == 4 * kThreads && (1 << kSubShift) == 64,
                    "16 threads of 4 cells per sub-chunk");
      const int b = blockIdx.x, tid = threadIdx.x;
      int pre = 0;
      {
        int k = tid;
        for (; k + 3 * kThreads < b; k += 4 * kThreads)
          pre += (a.L.chunkD[k] + a.L.chunkD[k + kThreads]) +
                 (a.L.chunkD[k + 2 * kThreads] + a.L.chunkD[k + 3 * kThreads]);
        for (; k < b; k += kThreads) pre += a.L.chunkD[k];
      }
      int red;
      blk_excl_scan(pre, sh, red);
      const int cb = c0 + tid * 4;
      int v[4], bs[4], sum = 0, tch = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool in = cb + u < nbk;
        const int la = in ? dA[cb + u] : 0, ll = in ? dL[cb + u] : 0;
        v[u] = la - ll;
        bs[u] = in ? co[cb + u] : 0;
        tch |= la | ll;
        sum += v[u];
      }
      int bt;
      int ex = red + blk_excl_scan(sum, sh, bt);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (cb + u < nbk) {
          cn[cb + u] = bs[u] + ex;
          ex += v[u];
        }
      if (b == a.nchunks - 1 && tid == 0) cn[nbk] = a.n;
      tch = tch != 0;
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) tch |= __shfl_xor(tch, o, 64);
      const int k = (c0 >> kSubShift) + (tid >> 4);
      if ((tid & 15) == 0 && k < a.nsub) a.subT[k] = tch;
    } else {
      chunk_scan(
          nbk, a.nchunks, a.L.chunk_shift, a.L.chunkD, [&](int c) { return dA[c] - dL[c]; },
          [&](int c) { return co[c]; }, cn, a.n, sh);
      const int k1 = min((c0 + (1 << a.L.chunk_shift)) >> kSubShift, a.nsub);
      for (int k = (c0 >> kSubShift) + (int)threadIdx.x; k < k1; k += kThreads) a.subT[k] = 1;
    }
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
  // one sub-chunk of 64 cells per wave.  Without recorded changes its slots
  // all shift by the same offset change (a coalesced copy); otherwise each
  // slot's cell is found by binary search over the sub-chunk's old offsets,
  // held one per lane and read by shuffles (waves take different paths, so
  // there is no workgroup barrier here)
  constexpr int kWaves = kThreads / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = blockIdx.x * kWaves + wv; k < a.nsub; k += gridDim.x * kWaves) {
    const int c_lo = k << kSubShift, ncl = min(64, nbk - c_lo);
    const int s_lo = co[c_lo], s_hi = co[c_lo + ncl];
    if (!a.subT[k]) {
      const int dlt = cn[c_lo] - s_lo;
      for (int s = s_lo + lane; s < s_hi; s += 64) pn[s + dlt] = po[s];
      continue;
    }
    const int cl = c_lo + min(lane, ncl - 1);
    const int my_co = co[cl], my_d = cn[cl] - co[cl], my_t = dL[cl] | dA[cl];
    for (int sb = s_lo; sb < s_hi; sb += 64) {
      const int s = sb + lane;
      int lo = 0, hi = ncl - 1;
#pragma unroll
      for (int it = 0; it < kSubShift; ++it) {
        const int mid = (lo + hi + 1) >> 1;
        const int v = __shfl(my_co, mid, 64);
        if (lo < hi) {
          if (v <= s) lo = mid;
          else hi = mid - 1;
        }
      }
      const int dlt = __shfl(my_d, lo, 64), tch = __shfl(my_t, lo, 64);
      if (s >= s_hi) continue;
      const int p = po[s], c = c_lo + lo;
      int pos = s + dlt;
      if (tch) {
        if (a.L.mrec[p]) continue;  // placed by the mover loop below
        // the old segment ascends in particle index
        for (int m = a.L.headL[c]; m >= 0; m = M[kMovRec * m + 3]) pos -= M[kMovRec * m] < p;
        for (int m = a.L.headA[c]; m >= 0; m = M[kMovRec * m + 4]) pos += M[kMovRec * m] < p;
      }
      pn[pos] = p;
    }
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
// workgroup (rank sort of a copy of the cell, in LDS up to kBigCellMax
// particles, beyond that in buffer 1); restores the zero counters;
// the last workgroup publishes the new state.
__global__ __launch_bounds__(kThreads) void k_sort_s4(SortArgs a) {
  // the overflow counters of the step that follows (read by no stage)
  if (a.ovf && blockIdx.x == 0 && threadIdx.x < 2) a.ovf[threadIdx.x] = 0;
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
    // the rank sort reads a copy of the cell: in LDS, or for cells beyond
    // kBigCellMax in buffer 1 (no live order during a full sort)
    int* src = m > kBigCellMax ? a.perm1 + s0 : sh;
    for (int i = threadIdx.x; i < m; i += kThreads) src[i] = perm[s0 + i];
    __threadfence_block();
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kThreads) {
      const int v = src[i];
      int rk = 0;
      for (int j = 0; j < m; ++j) rk += src[j] < v;
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
      f[2] = 0;
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
int enqueue_resort(dmpm_plan* P, const Mesh2<T>& M, const T* pb, int32_t* eid, hipStream_t s,
                   int64_t ld = -1, int* ovf = nullptr) {
  SortArgs a;
  a.ovf = ovf;
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
  a.subT = P->subT;
  a.nsub = P->nsub;
  // grid-stride stages: up to kSortGrid workgroups, fewer for small plans
  // (a thread per particle, a wave per sub-chunk, a thread per cell), at
  // least one per chunk where a stage takes a chunk per workgroup
  const int np = (int)((P->n + kThreads - 1) / kThreads);
  const int nsubw = (P->nsub + kThreads / 64 - 1) / (kThreads / 64);
  const int ncw = (P->ncells + 1 + kThreads - 1) / kThreads;
  const int g12 = std::max(P->nchunks, std::min(kSortGrid, std::max(np, nsubw)));
  // (positions: rows 0 and 1 of pb, row stride ld)
  k_sort_s1<T><<<g12, kThreads, 0, s>>>(M, pb, ld < 0 ? P->d.ld : ld, a);
  k_sort_s2<<<g12, kThreads, 0, s>>>(a);
  // stages 3 and 4 have little to do in the incremental mode (the movers, a
  // flag): small grids, since scheduling idle workgroups costs more
  k_sort_s3<<<std::max(1, std::min(kSortGrid / 4, np)), kThreads, 0, s>>>(a);
  k_sort_s4<<<std::max(1, std::min(kSortGrid / 8, ncw)), kThreads, 0, s>>>(a);
  return launch_check("resort");
}

template <typename T, int CB>
int tile_steps(dmpm_plan* P, const Mesh2<T>& M, T* pb, int32_t* eid, T* nb, int scheme, int vu,
               const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
               hipStream_t s, T* part, T* vkin, const int64_t* clock, int kbase) {
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
                                               P->nf_val, fns, clock,
                                               kbase + (int)(step - step0));
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

// Row r (blockIdx.y) of a block: dst[r * dld + i] = src[r * sld + i], i < n.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy_rows(T* __restrict__ dst, int64_t dld,
                                                        const T* __restrict__ src, int64_t sld,
                                                        int n) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[blockIdx.y * dld + i] = src[blockIdx.y * sld + i];
}

__global__ void k_clock_add(int64_t* clock, int64_t n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clock += n;
}

// Whether the one-launch step can run this plan: the particle and node blocks
// below 4 GiB (the kernels' 32-bit buffer offsets).
bool step_eligible(const dmpm_plan* P) {
  const uint64_t esz = P->d.dtype == DMPM_F32 ? 4 : 8;
  return P->step_kernel &&
         (uint64_t)DMPM_P_ROWS * (uint64_t)P->d.ld * esz < (1ull << 32) &&
         (uint64_t)DMPM_N_ROWS * (uint64_t)P->d.nld * esz < (1ull << 32);
}

// The kernels of one step (B: batch plan): the one-launch step (kind 1) or
// the split step (kind 2).  With events e0, e1 (timing): k_step_fast's
// dispatch records them at its start and end (FP32, kind 1); otherwise
// they bracket the step's launches.
template <typename T, bool B>
void step_launch(const StepArgs<T>& a, bool usf, int kind, hipStream_t s,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  constexpr bool kFast = sizeof(T) == 4;
  if (e0 && !(kFast && kind == 1)) {
    (void)hipEventRecord(e0, s);
    step_launch<T, B>(a, usf, kind, s);
    (void)hipEventRecord(e1, s);
    return;
  }
  if (e0) {
    if constexpr (kFast) {
      if (usf) {
        hipExtLaunchKernelGGL(k_step_fast<T, true, B>, dim3(a.ntiles), dim3(kFT), 0, s, e0, e1,
                              0, a);
        k_step_rounds<T, true, B><<<a.nrb + kTailBlocks, kFT, 0, s>>>(a);
      } else {
        hipExtLaunchKernelGGL(k_step_fast<T, false, B>, dim3(a.ntiles), dim3(kFT), 0, s, e0, e1,
                              0, a);
        k_step_rounds<T, false, B><<<a.nrb + kTailBlocks, kFT, 0, s>>>(a);
      }
    }
    return;
  }
  if (kind == 2) {
    k_split_kin<T, B, kSKinT, kSKinR><<<a.ntiles, kSKinT, 0, s>>>(a);
    if (usf) {
      k_split_force<T, true, B, kSForT, kSForR><<<a.ntiles, kSForT, 0, s>>>(a);
      k_split_g2p<T, true, B, kSG2pT, kSG2pR><<<a.ntiles + kTailBlocks, kSG2pT, 0, s>>>(a);
    } else {
      k_split_force<T, false, B, kSForT, kSForR><<<a.ntiles, kSForT, 0, s>>>(a);
      k_split_g2p<T, false, B, kSG2pT, kSG2pR><<<a.ntiles + kTailBlocks, kSG2pT, 0, s>>>(a);
    }
    return;
  }
  if (usf) {
    if constexpr (kFast) k_step_fast<T, true, B><<<a.ntiles, kFT, 0, s>>>(a);
    k_step_rounds<T, true, B><<<a.nrb + kTailBlocks, kFT, 0, s>>>(a);
  } else {
    if constexpr (kFast) k_step_fast<T, false, B><<<a.ntiles, kFT, 0, s>>>(a);
    k_step_rounds<T, false, B><<<a.nrb + kTailBlocks, kFT, 0, s>>>(a);
  }
}

// The one-launch step: per step, the re-sort, k_step_fast over every tile
// (FP32; batch plans: the tiles of every simulation) and k_step_rounds over
// the tiles that overflowed the fast path and the particles outside the
// mesh.  The particle state alternates between blocks pa and pb2 (dynamic
// rows only; pb2 null: the plan's shadow block), starting from pa; pstat
// holds the static rows (mass, E) and takes the temporaries.  clock: step
// k's time is read on the device as (*clock + kbase + k) * dt.  prepare:
// only allocate the workspaces and upload the tables.
// Workspaces of step_run supplied by the caller instead of the plan
// (dmpm_step_between_noalloc): the previous/next nodal velocities, the
// overflow lists and the device copy of the step tables.
template <typename T> struct StepScratch {
  T* vbuf;       // [4][nnodes]
  int* ovf;      // [2 + 2 ntiles] (FP32)
  StepAux* aux;  // uploaded by every call (the memory is shared between plans)
};

template <typename T>
int step_run(dmpm_plan* P, const Mesh2<T>& M, T* pa, T* pb2, T* pstat, int32_t* eid, T* nb,
             int scheme, int vu, const double* grav, double dt, int64_t step0, int64_t nsteps,
             int temps, hipStream_t s, T* vkin, T* sout = nullptr, int64_t sld = 0,
             const T* sin = nullptr, const int64_t* clock = nullptr, int kbase = 0,
             bool prepare = false, const StepScratch<T>* ext = nullptr,
             const T* vprev = nullptr, const T* fdprev = nullptr) {
  constexpr bool kFast = sizeof(T) == 4;
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int nn = P->nnodes;
  StepArgs<T> a;
  a.tiles_x = P->sims_h[0].tiles_x;
  a.ntiles = P->ntiles;
  a.sims = P->sims;
  a.tile_sim = P->tile_sim;
  a.nc_all = P->ncells;
  a.nn_all = nn;
  a.nsims = P->nsims;
  T* shadow = pb2;
  T* vbuf;
  int* ovf;
  if (ext) {
    if (!pb2) return fail(DMPM_EINVAL, "caller-supplied workspaces need both particle blocks");
    vbuf = ext->vbuf;
    ovf = kFast ? ext->ovf : nullptr;
  } else {
    if (!shadow) shadow = (T*)plan_workspace(P, 5, (size_t)DMPM_P_ROWS * ld * sizeof(T));
    vbuf = (T*)plan_workspace(P, 6, (size_t)4 * nn * sizeof(T));
    ovf = kFast ? (int*)plan_workspace(P, 7, (size_t)(2 + 2 * a.ntiles) * sizeof(int)) : nullptr;
  }
  if (!shadow || !vbuf || (kFast && !ovf)) return fail(DMPM_EHIP, "workspace allocation failed");
  if (!prepare && !vprev) {
    // the nodal velocity a massless node keeps (the counters of ovf are
    // zeroed by each step's re-sort); with vprev the first step reads it
    // from there
    k_copy_rows<T><<<dim3((unsigned)((nn + kThreads - 1) / kThreads), 2), kThreads, 0, s>>>(
        vbuf, nn, nb + DMPM_N_V * nld, nld, nn);
    DMPM_TRY(launch_check("copy_rows"));
  }
  a.g = M.g0;
  a.pb = pstat;
  a.ld32 = (uint32_t)ld;
  a.nb = nb;
  a.nld32 = (uint32_t)nld;
  // damping-force rows: read from fdprev (another node block of the same
  // shape) and written to nb on the last step, or read from nb
  a.fdsrc = fdprev ? fdprev : nb;
  a.wfd = fdprev ? 1 : 0;
  a.pbytes = (uint32_t)((uint64_t)DMPM_P_ROWS * ld * sizeof(T));
  a.nbytes = (uint32_t)((uint64_t)DMPM_N_ROWS * nld * sizeof(T));
  a.vkin = vkin;
  a.sr = sort_ref(P);
  a.cidx = P->cidx;
  a.cmask = P->cmask;
  StepAux x;
  std::memset(&x, 0, sizeof x);  // padding too: the struct is compared bytewise
  x.L = move_log(P);
  x.cval = P->cval;
  x.nfd = P->nfd;
  x.nf_off = P->nf_off;
  x.nf_fn = P->nf_fn;
  x.nf_val = P->nf_val;
  x.fns.kind = P->fn_kind;
  x.fns.off = P->fn_off;
  x.fns.cnt = P->fn_cnt;
  x.fns.xs = P->fn_x;
  x.fns.fs = P->fn_f;
  a.has_nf = P->nfd != nullptr;
  static_assert(sizeof(StepAux) <= sizeof(P->step_aux), "step_aux size");
  StepAux* aux = ext ? ext->aux : (StepAux*)plan_workspace(P, 8, sizeof(StepAux));
  if (!aux) return fail(DMPM_EHIP, "workspace allocation failed");
  if (ext) {
    // shared memory: upload on every call, from host memory the plan keeps
    std::memcpy(P->step_aux_ext, &x, sizeof(StepAux));
    DMPM_TRY(hip_check(
        hipMemcpyAsync(aux, P->step_aux_ext, sizeof(StepAux), hipMemcpyHostToDevice, s),
        "hipMemcpyAsync"));
  } else if (P->step_aux_bytes != (int)sizeof(StepAux) ||
             std::memcmp(P->step_aux, &x, sizeof(StepAux)) != 0) {
    // upload only when the tables changed (a copy from host memory orders
    // the host behind the stream)
    std::memcpy(P->step_aux, &x, sizeof(StepAux));
    P->step_aux_bytes = (int)sizeof(StepAux);
    DMPM_TRY(hip_check(hipMemcpyAsync(aux, P->step_aux, sizeof(StepAux), hipMemcpyHostToDevice, s),
                       "hipMemcpyAsync"));
  }
  if (prepare) return DMPM_OK;
  a.aux = aux;
  a.clock = clock;
  a.mat.kind = P->d.material;
  a.mat.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) a.mat.de[k] = P->d.de[k];
  a.gx = T(grav[0]);
  a.gy = T(grav[1]);
  a.dt = T(dt);
  a.vu = vu;
  a.temps = temps;
  a.n = (int)P->n;
  a.ovf = ovf;
  a.sout = sout;
  a.sin = sin;
  a.sld32 = (uint32_t)sld;
  a.sbytes = (uint32_t)(8 * (uint64_t)sld * sizeof(T));
  // FP32: overflow tiles are rare, a few workgroups loop over them
  a.nrb = kFast ? std::min(a.ntiles, 64) : a.ntiles;
  for (int64_t k = 0; k < nsteps; ++k) {
    T* src = (k & 1) ? shadow : pa;
    {
      ProfScopeT ps_(P, KC_SORT, s);
      // positions: from the state tensor when the state rows live there
      if (sin && k == 0)
        DMPM_TRY(enqueue_resort<T>(P, M, sin, eid, s, sld, ovf));
      else
        DMPM_TRY(enqueue_resort<T>(P, M, src, eid, s, -1, ovf));
    }
    a.src = src;
    a.dst = (k & 1) ? pa : shadow;
    a.vold = (vprev && k == 0) ? vprev : vbuf + (k & 1) * 2 * (int64_t)nn;
    a.vld = (vprev && k == 0) ? nld : nn;
    a.vnew = vbuf + ((k + 1) & 1) * 2 * (int64_t)nn;
    a.t = T(step0 + k) * a.dt;
    a.kofs = kbase + (int)k;
    a.par = (int)(k & 1);
    a.last = k == nsteps - 1 ? 1 : 0;
    // timing: the events of k_step_fast's own dispatch (the class's time is
    // that kernel's), of the step's launches otherwise
    hipEvent_t pe0 = nullptr, pe1 = nullptr;
    const bool timed = prof_slot(P, KC_STEP, &pe0, &pe1);
    if (P->nsims > 1)
      step_launch<T, true>(a, scheme == DMPM_USF, P->step_kernel, s, pe0, pe1);
    else
      step_launch<T, false>(a, scheme == DMPM_USF, P->step_kernel, s, pe0, pe1);
    if (timed) prof_commit(P, KC_STEP);
    DMPM_TRY(launch_check("step"));
  }
  return DMPM_OK;
}

template <typename T>
int tile_steps_cb(dmpm_plan* P, const Mesh2<T>& M, T* pb, int32_t* eid, T* nb, int scheme,
                  int vu, const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
                  hipStream_t s, T* part, T* vkin, const int64_t* clock = nullptr,
                  int kbase = 0);

template <typename T>
int tile_run_t(dmpm_plan* P, T* pb, int32_t* eid, T* nb, int scheme, int vu, const double* grav,
               double dt, int64_t step0, int64_t nsteps, int temps, hipStream_t s,
               int64_t* clock) {
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
  if (step_eligible(P)) {
    // the one-launch step alternates blocks, so it runs an even number of
    // steps and the state ends in the caller's block; the first step of an
    // odd-length run uses the (in-place) per-phase launches
    const int odd = (int)(nsteps & 1);
    if (odd)
      DMPM_TRY(tile_steps_cb<T>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, 1, temps, s,
                                part, vkin, clock, 0));
    if (nsteps > odd)
      DMPM_TRY(step_run<T>(P, M, pb, nullptr, pb, eid, nb, scheme, vu, grav, dt, step0 + odd,
                           nsteps - odd, temps, s, vkin, nullptr, 0, nullptr, clock, odd));
  } else {
    DMPM_TRY(tile_steps_cb<T>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                              part, vkin, clock, 0));
  }
  if (clock) {
    k_clock_add<<<1, 64, 0, s>>>(clock, nsteps);
    DMPM_TRY(launch_check("clock"));
  }
  return DMPM_OK;
}

// Workspaces of tile_run_t (and its table upload), with no other work.
template <typename T>
int tile_prepare_t(dmpm_plan* P, int scheme, hipStream_t s) {
  const Mesh2<T> M = mesh_of<T>(P);
  T* part = (T*)plan_workspace(P, 0, (size_t)4 * NQ_FORCE * M.ncells * sizeof(T));
  T* vkin = (T*)plan_workspace(P, 1, (size_t)2 * P->nnodes * sizeof(T));
  if (!part || !vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  if (!step_eligible(P)) return DMPM_OK;
  const double grav[2] = {0, 0};
  return step_run<T>(P, M, nullptr, nullptr, nullptr, nullptr, nullptr, scheme, 0, grav, 0, 0,
                     0, 0, s, vkin, nullptr, 0, nullptr, nullptr, 0, true);
}

// One step from block pin to block pout (pin unchanged; pout's static rows
// must hold the particle set's mass and E).  Single-simulation plans run the
// one-launch step between the two blocks; others copy the dynamic rows and
// step pout in place.
template <typename T>
int tile_between_t(dmpm_plan* P, const T* pin, T* pout, int32_t* eid, T* nb, int scheme, int vu,
                   const double* grav, double dt, int64_t step, int temps, int chained,
                   const T* sin, T* sout, int64_t sld, hipStream_t s, const T* nprev = nullptr,
                   T* vkin_out = nullptr) {
  const Mesh2<T> M = mesh_of<T>(P);
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int n = (int)P->n;
  T* part = (T*)plan_workspace(P, 0, (size_t)4 * NQ_FORCE * M.ncells * sizeof(T));
  T* vkin = (T*)plan_workspace(P, 1, (size_t)2 * P->nnodes * sizeof(T));
  if (!part || !vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  if (!chained) {
    // record the particles moved since the plan's last step (as dmpm_run does)
    // positions: rows 0, 1 of the state tensor when the state rows live there
    const T* px = sin ? sin : pin;
    const int64_t pld = sin ? sld : ld;
    {
      ProfScopeT ps_(P, KC_SORT, s);
      DMPM_TRY(enqueue_resort<T>(P, M, px, eid, s, pld));
    }
    k_tile_check<T><<<nblk(n), kThreads, 0, s>>>(M, px, pld, n, eid, P->key, move_log(P));
    DMPM_TRY(launch_check("tile_check"));
  }
  const bool other = nprev && nprev != nb;
  if (step_eligible(P) && (uint64_t)8 * (uint64_t)sld * sizeof(T) < (1ull << 32))
    // a massless node keeps the velocity of the previous node block, and the
    // damping force is that block's (read there directly when it is another
    // block, whose damping rows the step then writes to nb)
    return step_run<T>(P, M, const_cast<T*>(pin), pout, pout, eid, nb, scheme, vu, grav, dt,
                       step, 1, temps, s, vkin_out ? vkin_out : vkin, sout, sld, sin, nullptr,
                       0, false, nullptr, other ? nprev + (int64_t)DMPM_N_V * nld : nullptr,
                       other ? nprev : nullptr);
  if (other)
    // the rows the per-phase step reads from the node block
    for (const int r : {(int)DMPM_N_V, (int)DMPM_N_FDAMP})
      DMPM_TRY(hip_check(hipMemcpy2DAsync(nb + (int64_t)r * nld, nld * sizeof(T),
                                          nprev + (int64_t)r * nld, nld * sizeof(T),
                                          (size_t)P->nnodes * sizeof(T), 2,
                                          hipMemcpyDeviceToDevice, s),
                         "hipMemcpy2DAsync"));
  const int runs[6][2] = {{DMPM_P_X, 4},          {DMPM_P_VOLUME, 2}, {DMPM_P_STRAIN, 2},
                          {DMPM_P_STRAIN + 3, 1}, {DMPM_P_STRESS, 4}, {DMPM_P_VOLSTRAIN_C, 1}};
  for (const auto& r : runs)
    DMPM_TRY(hip_check(hipMemcpy2DAsync(pout + r[0] * ld, ld * sizeof(T), pin + r[0] * ld,
                                        ld * sizeof(T), (size_t)n * sizeof(T), r[1],
                                        hipMemcpyDeviceToDevice, s),
                       "hipMemcpy2DAsync"));
  const int srows[4][3] = {{DMPM_P_X, 0, 4},
                           {DMPM_P_STRESS, 4, 2},
                           {DMPM_P_STRESS + 3, 6, 1},
                           {DMPM_P_VOLUME, 7, 1}};
  if (sin)
    for (const auto& r : srows)
      DMPM_TRY(hip_check(hipMemcpy2DAsync(pout + r[0] * ld, ld * sizeof(T), sin + r[1] * sld,
                                          sld * sizeof(T), (size_t)n * sizeof(T), r[2],
                                          hipMemcpyDeviceToDevice, s),
                         "hipMemcpy2DAsync"));
  DMPM_TRY(tile_steps_cb<T>(P, M, pout, eid, nb, scheme, vu, grav, dt, step, 1, temps, s, part,
                            vkin));
  if (sout) {
    for (const auto& r : srows)
      DMPM_TRY(hip_check(hipMemcpy2DAsync(sout + r[1] * sld, sld * sizeof(T), pout + r[0] * ld,
                                          ld * sizeof(T), (size_t)n * sizeof(T), r[2],
                                          hipMemcpyDeviceToDevice, s),
                         "hipMemcpy2DAsync"));
  }
  if (vkin_out)
    DMPM_TRY(hip_check(hipMemcpyAsync(vkin_out, vkin, (size_t)2 * P->nnodes * sizeof(T),
                                      hipMemcpyDeviceToDevice, s),
                       "hipMemcpyAsync"));
  return DMPM_OK;
}

// tile_steps for the plan's cells per workgroup of the cell passes.
template <typename T>
int tile_steps_cb(dmpm_plan* P, const Mesh2<T>& M, T* pb, int32_t* eid, T* nb, int scheme,
                  int vu, const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
                  hipStream_t s, T* part, T* vkin, const int64_t* clock, int kbase) {
  switch (P->cell_block) {
    case 64:
      return tile_steps<T, 64>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                               part, vkin, clock, kbase);
    default:
      return tile_steps<T, 128>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                                part, vkin, clock, kbase);
    case 256:
      return tile_steps<T, 256>(P, M, pb, eid, nb, scheme, vu, grav, dt, step0, nsteps, temps, s,
                                part, vkin, clock, kbase);
  }
}

// ---- one step without allocation ---------------------------------------------
//
// dmpm_step_between_noalloc keeps its workspaces (kinematic nodal velocity,
// previous and next nodal velocities, overflow lists, step tables) in one
// file-scope device array instead of plan memory.  Host code must not pass
// the array's name to a kernel: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device at plan creation
// (noalloc_scratch_setup) and kept in g_scratch_dev.
constexpr size_t kNoAllocScratchBytes = (size_t)DMPM_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the array that one step of plan P needs (see the layout below).
template <typename T>
size_t noalloc_bytes(const dmpm_plan* P) {
  const size_t nn = (size_t)P->nnodes;
  return align256(2 * nn * sizeof(T)) + align256(4 * nn * sizeof(T)) +
         align256((2 + 2 * (size_t)P->ntiles) * sizeof(int)) + align256(sizeof(StepAux));
}

template <typename T>
int tile_between_noalloc_t(dmpm_plan* P, const T* pin, T* pout, int32_t* eid, T* nb, int scheme,
                           int vu, const double* grav, double dt, int64_t step, int temps,
                           int chained, const T* sin, T* sout, int64_t sld, hipStream_t s) {
  if (P->device < 0 || P->device >= kMaxDevices || !g_scratch_dev[P->device])
    return fail(DMPM_EHIP, "no scratch address for this plan's device");
  if (P->nsims != 1 || !step_eligible(P) ||
      (uint64_t)8 * (uint64_t)sld * sizeof(T) >= (1ull << 32))
    return fail(DMPM_EINVAL,
                "dmpm_step_between_noalloc covers the one-launch step of single-simulation "
                "plans (blocks below 4 GiB, step kernel option on)");
  const size_t need = noalloc_bytes<T>(P);
  if (need > kNoAllocScratchBytes)
    return fail(DMPM_EINVAL, "dmpm_step_between_noalloc: the plan needs " +
                                 std::to_string(need) + " bytes of scratch, above the " +
                                 std::to_string(kNoAllocScratchBytes) +
                                 "-byte array (24 bytes per node in FP32, 48 in FP64)");
  // layout: vkin [2][nn] | vbuf [4][nn] | ovf [2 + 2 ntiles] | aux
  unsigned char* base = (unsigned char*)g_scratch_dev[P->device];
  const size_t nn = (size_t)P->nnodes;
  T* vkin = (T*)base;
  StepScratch<T> ext;
  ext.vbuf = (T*)(base + align256(2 * nn * sizeof(T)));
  ext.ovf = (int*)((unsigned char*)ext.vbuf + align256(4 * nn * sizeof(T)));
  ext.aux = (StepAux*)((unsigned char*)ext.ovf +
                       align256((2 + 2 * (size_t)P->ntiles) * sizeof(int)));
  const Mesh2<T> M = mesh_of<T>(P);
  const int64_t ld = P->d.ld;
  const int n = (int)P->n;
  if (!chained) {
    const T* px = sin ? sin : pin;
    const int64_t pld = sin ? sld : ld;
    {
      ProfScopeT ps_(P, KC_SORT, s);
      DMPM_TRY(enqueue_resort<T>(P, M, px, eid, s, pld));
    }
    k_tile_check<T><<<nblk(n), kThreads, 0, s>>>(M, px, pld, n, eid, P->key, move_log(P));
    DMPM_TRY(launch_check("tile_check"));
  }
  return step_run<T>(P, M, const_cast<T*>(pin), pout, pout, eid, nb, scheme, vu, grav, dt, step,
                     1, temps, s, vkin, sout, sld, sin, nullptr, 0, false, &ext);
}

}  // namespace

namespace dmpm {

bool tile_eligible(const dmpm_plan* P, int pfext) {
  return P->dim == 2 && !P->has_tractions && !pfext && P->n > 0 && P->nchunks <= kMaxChunks;
}

int tile_between(dmpm_plan* P, const void* pin, void* pout, int32_t* eid, void* nbuf,
                 int scheme, int vu, const double* grav, double dt, int64_t step, int temps,
                 int chained, const void* sin, void* sout, int64_t sld, hipStream_t s) {
  if (P->d.dtype == DMPM_F32)
    return tile_between_t<float>(P, (const float*)pin, (float*)pout, eid, (float*)nbuf, scheme,
                                 vu, grav, dt, step, temps, chained, (const float*)sin,
                                 (float*)sout, sld, s);
  return tile_between_t<double>(P, (const double*)pin, (double*)pout, eid, (double*)nbuf, scheme,
                                vu, grav, dt, step, temps, chained, (const double*)sin,
                                (double*)sout, sld, s);
}

int tile_between_saved(dmpm_plan* P, const void* pin, void* pout, int32_t* eid,
                       const void* nprev, void* nout, int scheme, int vu, const double* grav,
                       double dt, int64_t step, int chained, const void* sin, void* sout,
                       int64_t sld, void* vkin_out, int32_t* perm_out, int32_t* cs_out,
                       hipStream_t s) {
  if (P->d.dtype == DMPM_F32)
    DMPM_TRY(tile_between_t<float>(P, (const float*)pin, (float*)pout, eid, (float*)nout,
                                   scheme, vu, grav, dt, step, 0, chained, (const float*)sin,
                                   (float*)sout, sld, s, (const float*)nprev,
                                   (float*)vkin_out));
  else
    DMPM_TRY(tile_between_t<double>(P, (const double*)pin, (double*)pout, eid, (double*)nout,
                                    scheme, vu, grav, dt, step, 0, chained, (const double*)sin,
                                    (double*)sout, sld, s, (const double*)nprev,
                                    (double*)vkin_out));
  return export_sort(P, perm_out, cs_out, s);
}

int tile_run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
             const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
             hipStream_t s, int64_t* clock) {
  if (P->d.dtype == DMPM_F32)
    return tile_run_t<float>(P, (float*)pbuf, eid, (float*)nbuf, scheme, vu, grav, dt, step0,
                             nsteps, temps, s, clock);
  return tile_run_t<double>(P, (double*)pbuf, eid, (double*)nbuf, scheme, vu, grav, dt, step0,
                            nsteps, temps, s, clock);
}

int tile_prepare(dmpm_plan* P, int scheme, hipStream_t s) {
  if (P->d.dtype == DMPM_F32) return tile_prepare_t<float>(P, scheme, s);
  return tile_prepare_t<double>(P, scheme, s);
}

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(DMPM_EINVAL, "device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return DMPM_OK;
  void* p = nullptr;
  DMPM_TRY(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress"));
  if (!p) return fail(DMPM_EHIP, "hipGetSymbolAddress returned null");
  g_scratch_dev[device] = p;
  return DMPM_OK;
}

int tile_between_noalloc(dmpm_plan* P, const void* pin, void* pout, int32_t* eid, void* nbuf,
                         int scheme, int vu, const double* grav, double dt, int64_t step,
                         int temps, int chained, const void* sin, void* sout, int64_t sld,
                         hipStream_t s) {
  if (P->d.dtype == DMPM_F32)
    return tile_between_noalloc_t<float>(P, (const float*)pin, (float*)pout, eid, (float*)nbuf,
                                         scheme, vu, grav, dt, step, temps, chained,
                                         (const float*)sin, (float*)sout, sld, s);
  return tile_between_noalloc_t<double>(P, (const double*)pin, (double*)pout, eid, (double*)nbuf,
                                        scheme, vu, grav, dt, step, temps, chained,
                                        (const double*)sin, (double*)sout, sld, s);
}

}  // namespace dmpm
