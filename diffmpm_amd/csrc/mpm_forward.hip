// diffmpm_amd forward step for MI355X (gfx950): plan, deterministic counting
// sort of particles by cell, per-method kernels and the phase/step drivers.
//
// Determinism: P2G never uses float atomics.  Particles are ordered by a
// stable counting sort on their cell (ties keep particle-index order), and
// every node sums its contributions in one fixed order — adjacent cells
// (i-1,j-1), (i,j-1), (i-1,j), (i,j), and within a cell ascending particle
// index — so results are bit-reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "mpm_common.h"
#include "mpm_plan.h"

using namespace dmpm;

// ----------------------------------------------------------------------------
// error handling
// ----------------------------------------------------------------------------
namespace dmpm {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return DMPM_OK;
  return fail(DMPM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

int launch_check(const char* what) { return hip_check(hipGetLastError(), what); }
}  // namespace dmpm



static inline unsigned nblocks(int64_t n, int b = kBlock) {
  return (unsigned)((n + b - 1) / b);
}

// ----------------------------------------------------------------------------
// Sort kernels.  flags[0] is the "dirty" flag: every sort kernel returns at once
// when it is clear, so the chain can be enqueued every step at the cost of
// launch latency only.
// ----------------------------------------------------------------------------
__global__ void k_keys_from_eid(const int* __restrict__ eid, int n, int ncells,
                                int* __restrict__ key, int* __restrict__ flags) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int e = eid[p];
  int k = (e >= 0 && e < ncells) ? e : ncells;
  if (key[p] != k) {
    key[p] = k;
    flags[0] = 1;
  }
}

__global__ void k_sort_zero(int* __restrict__ counts, int* __restrict__ fill, int nb,
                            int* __restrict__ flags) {
  if (flags[0] == 0) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    counts[i] = 0;
    fill[i] = 0;
  }
}

__global__ void k_sort_count(const int* __restrict__ key, int n, int* __restrict__ counts,
                             const int* __restrict__ flags) {
  if (flags[0] == 0) return;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) atomicAdd(&counts[key[p]], 1);
}

// Exclusive scan of counts[0..nb) into cs[0..nb], cs[nb] = total.  Three
// passes: per-block scan of kScanChunk elements, scan of the block totals by
// one workgroup, add-back.
constexpr int kScanChunk = 1024;  // 256 threads x 4 elements

__device__ __forceinline__ int block_excl_scan(int v, int* sh, int& total) {
  // Hillis-Steele over 256 threads in LDS.
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int a = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  int incl = sh[t];
  total = sh[kBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_blocks(const int* __restrict__ counts, int nb, int* __restrict__ cs,
                              int* __restrict__ bsum, const int* __restrict__ flags) {
  if (flags[0] == 0) return;
  __shared__ int sh[kBlock];
  int base = blockIdx.x * kScanChunk + threadIdx.x * 4;
  int v[4];
  int s = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < nb) ? counts[base + k] : 0;
    s += v[k];
  }
  int total;
  int ex = block_excl_scan(s, sh, total);
  for (int k = 0; k < 4; ++k) {
    if (base + k < nb) cs[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void k_scan_sums(int* __restrict__ bsum, int nbs, const int* __restrict__ flags) {
  if (flags[0] == 0) return;
  __shared__ int sh[kBlock];
  int carry = 0;
  for (int base = 0; base < nbs; base += kBlock) {
    int i = base + threadIdx.x;
    int v = (i < nbs) ? bsum[i] : 0;
    int total;
    int ex = block_excl_scan(v, sh, total);
    if (i < nbs) bsum[i] = ex + carry;
    carry += total;
  }
}

__global__ void k_scan_add(int* __restrict__ cs, int nb, int n, const int* __restrict__ bsum,
                           const int* __restrict__ flags) {
  if (flags[0] == 0) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) cs[i] += bsum[i / kScanChunk];
  if (i == 0) cs[nb] = n;
}

__global__ void k_sort_scatter(const int* __restrict__ key, int n, const int* __restrict__ cs,
                               int* __restrict__ fill, int* __restrict__ perm,
                               const int* __restrict__ flags) {
  if (flags[0] == 0) return;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int k = key[p];
  int slot = cs[k] + atomicAdd(&fill[k], 1);
  perm[slot] = p;
}

// Restores ascending particle index inside each cell (makes the order stable
// and independent of atomic arrival order).  Cells of up to kSmallCell
// particles are insertion-sorted by one thread; larger cells are queued for
// k_sort_fix_big.  The outside bucket is never used and stays unsorted.
constexpr int kSmallCell = 64;
constexpr int kBigCellMax = 8192;  // LDS capacity of the rank sort

__global__ void k_sort_fix(const int* __restrict__ cs, int ncells, int* __restrict__ perm,
                           int* __restrict__ flags, int* __restrict__ big,
                           int* __restrict__ counts, int* __restrict__ fill) {
  if (flags[0] == 0) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;  // cells and the outside bucket
  counts[c] = 0;  // both stay zero between sorts
  fill[c] = 0;
  int b = cs[c], e = cs[c + 1];
  if (e - b > kSmallCell) {
    int k = atomicAdd(&flags[3], 1);
    big[k] = c;
    return;
  }
  for (int i = b + 1; i < e; ++i) {
    int v = perm[i];
    int j = i - 1;
    while (j >= b && perm[j] > v) {
      perm[j + 1] = perm[j];
      --j;
    }
    perm[j + 1] = v;
  }
}

// One workgroup per queued cell: each element's slot is its rank among the
// cell's (distinct) particle indices, O(k^2 / 256) per cell, so the cell
// ends in ascending particle index whatever the scatter's arrival order.
// Cells of up to kBigCellMax particles rank from a copy in LDS; bigger ones
// from a copy in `scratch` (the second permutation buffer, which holds no
// live order during this sort: the sort ends by selecting buffer 0 and
// invalidating the fused path's incremental state).
__global__ void k_sort_fix_big(const int* __restrict__ cs, int* __restrict__ perm,
                               int* __restrict__ flags, const int* __restrict__ big,
                               int* __restrict__ scratch) {
  if (flags[0] == 0) return;
  __shared__ int sh[kBigCellMax];
  const int nbig = flags[3];
  for (int q = blockIdx.x; q < nbig; q += gridDim.x) {
    const int c = big[q];
    const int b = cs[c], n = cs[c + 1] - b;
    int* src = n > kBigCellMax ? scratch + b : sh;
    for (int i = threadIdx.x; i < n; i += blockDim.x) src[i] = perm[b + i];
    __threadfence_block();
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int v = src[i];
      int r = 0;
      for (int j = 0; j < n; ++j) r += src[j] < v;
      perm[b + r] = v;
    }
    __syncthreads();
  }
}

__global__ void k_sort_done(int* flags) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (flags[0]) {
      flags[1] = 0;   // a new permutation: no longer known to be the identity
      flags[4] = 0;   // it is in buffer 0
      flags[10] = 1;  // the fused path's recorded cell changes no longer apply
    }
    flags[0] = 0;
    flags[3] = 0;
  }
}

namespace dmpm {
int enqueue_sort(dmpm_plan* P, hipStream_t s) {
  ProfScope ps_(P, KC_SORT, s);
  const int n = (int)P->n;
  const int nb = P->ncells + 1;
  if (n > 0) {
    k_sort_zero<<<nblocks(nb), kBlock, 0, s>>>(P->counts, P->fill, nb, P->flags);
    k_sort_count<<<nblocks(n), kBlock, 0, s>>>(P->key, n, P->counts, P->flags);
    const int nbs = (nb + kScanChunk - 1) / kScanChunk;
    k_scan_blocks<<<nbs, kBlock, 0, s>>>(P->counts, nb, P->cs, P->bsum, P->flags);
    k_scan_sums<<<1, kBlock, 0, s>>>(P->bsum, nbs, P->flags);
    k_scan_add<<<nblocks(nb), kBlock, 0, s>>>(P->cs, nb, n, P->bsum, P->flags);
    k_sort_scatter<<<nblocks(n), kBlock, 0, s>>>(P->key, n, P->cs, P->fill, P->perm, P->flags);
    k_sort_fix<<<nblocks(P->ncells + 1), kBlock, 0, s>>>(P->cs, P->ncells, P->perm, P->flags,
                                                     P->big, P->counts, P->fill);
    k_sort_fix_big<<<256, kBlock, 0, s>>>(P->cs, P->perm, P->flags, P->big, P->perm2);
  }
  k_sort_done<<<1, 64, 0, s>>>(P->flags);
  return launch_check("sort");
}

int ensure_sorted(dmpm_plan* P, const int32_t* eid, hipStream_t s) {
  if (P->n > 0)
    k_keys_from_eid<<<nblocks(P->n), kBlock, 0, s>>>(eid, (int)P->n, P->ncells, P->key,
                                                     P->flags);
  DMPM_TRY(launch_check("keys"));
  return enqueue_sort(P, s);
}
}  // namespace dmpm

// ----------------------------------------------------------------------------
// Per-particle kernels
// ----------------------------------------------------------------------------
#define PROW(r) (pb + (int64_t)(r) * ld)
#define NROW(r) (nb + (int64_t)(r) * nld)

// Element ids (A1); also refreshes the sort key and, optionally, writes
// reference_loc (A2) for the new ids.
template <typename T, int DIM>
__global__ void k_locate(Geo<T> g, T* __restrict__ pb, int64_t ld, int n,
                         int* __restrict__ eid, int* __restrict__ key, int* __restrict__ flags,
                         int write_xi) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  T x = PROW(DMPM_P_X)[p];
  T y = DIM == 2 ? PROW(DMPM_P_X + 1)[p] : T(0);
  int e = element_id<T, DIM>(g, x, y);
  eid[p] = e;
  int k = (e >= 0 && e < g.ncells) ? e : g.ncells;
  if (key[p] != k) {
    key[p] = k;
    flags[0] = 1;
  }
  if (write_xi) {
    if (k == g.ncells) {
      PROW(DMPM_P_XI)[p] = qnan<T>();
      if (DIM == 2) PROW(DMPM_P_XI + 1)[p] = qnan<T>();
    } else if (DIM == 1) {
      PROW(DMPM_P_XI)[p] = natural(x, g.X(e), g.X(e + 1));
    } else {
      int I = e % g.nx, J = e / g.nx;
      PROW(DMPM_P_XI)[p] = natural(x, g.X(I), g.X(I + 1));
      PROW(DMPM_P_XI + 1)[p] = natural(y, g.Y(J), g.Y(J + 1));
    }
  }
}

template <typename T, int DIM>
__global__ void k_natural_coords(Geo<T> g, T* __restrict__ pb, int64_t ld, int n,
                                 const int* __restrict__ eid) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int e = eid[p];
  if (e < 0 || e >= g.ncells) {
    PROW(DMPM_P_XI)[p] = qnan<T>();
    if (DIM == 2) PROW(DMPM_P_XI + 1)[p] = qnan<T>();
    return;
  }
  if (DIM == 1) {
    PROW(DMPM_P_XI)[p] = natural(PROW(DMPM_P_X)[p], g.X(e), g.X(e + 1));
  } else {
    int I = e % g.nx, J = e / g.nx;
    PROW(DMPM_P_XI)[p] = natural(PROW(DMPM_P_X)[p], g.X(I), g.X(I + 1));
    PROW(DMPM_P_XI + 1)[p] = natural(PROW(DMPM_P_X + 1)[p], g.Y(J), g.Y(J + 1));
  }
}

// particle.py:158-178.  Counts come from the sorted cell offsets.
template <typename T, int DIM>
__global__ void k_particle_volume(Geo<T> g, T* __restrict__ pb, int64_t ld, int n,
                                  const int* __restrict__ eid, SortRef sr, T vol_e) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int* __restrict__ cs = sr.cs();
  int e = eid[p];
  T V;
  if (e < 0 || e >= g.ncells) {
    V = qnan<T>();
  } else {
    V = vol_e / T(cs[e + 1] - cs[e]);
  }
  PROW(DMPM_P_VOLUME)[p] = V;
  T sz = DIM == 2 ? sqrt(V) : V;
  PROW(DMPM_P_SIZE)[p] = sz;
  if (DIM == 2) PROW(DMPM_P_SIZE + 1)[p] = sz;
  PROW(DMPM_P_MASS)[p] = V * PROW(DMPM_P_DENSITY)[p];
}

// ---- strain / volume / stress (A8-A10) --------------------------------------
enum { S_STRAIN = 1, S_VOLUME = 2, S_STRESS = 4, S_TEMPS = 8 };

struct Mat {
  int kind;       // DMPM_LINEAR_ELASTIC / DMPM_SIMPLE_MATERIAL
  int per_E;      // per-particle E from row DMPM_P_YOUNGS
  double de[3];   // (a1, a2, G) or (E, -, -)
};

template <typename T>
__device__ __forceinline__ void stress_increment(const Mat& m, const T* __restrict__ pb,
                                                 int64_t ld, int p, const T d[6], T ds[6]) {
  if (m.kind == DMPM_SIMPLE_MATERIAL) {
    T E = T(m.de[0]);
    for (int k = 0; k < 6; ++k) ds[k] = d[k] * E;  // material.py:130-131
    return;
  }
  T a1 = T(m.de[0]), a2 = T(m.de[1]), G = T(m.de[2]);
  if (m.per_E) {
    T E = PROW(DMPM_P_YOUNGS)[p];
    a1 = a1 * E;
    a2 = a2 * E;
    G = G * E;
  }
  // De @ dstrain with the zero blocks of De dropped (material.py:102-116).
  ds[0] = (a1 * d[0] + a2 * d[1]) + a2 * d[2];
  ds[1] = (a2 * d[0] + a1 * d[1]) + a2 * d[2];
  ds[2] = (a2 * d[0] + a2 * d[1]) + a1 * d[2];
  ds[3] = G * d[3];
  ds[4] = G * d[4];
  ds[5] = G * d[5];
}

// Strain rate at (xi, eta) of the particle's cell from nodal velocities
// (particle.py:281-313), snapped.
template <typename T, int DIM>
__device__ __forceinline__ void strain_rate(const Geo<T>& g, int I, int J, T xi, T eta,
                                            const T vx[4], const T vy[4], T r[6]) {
  for (int k = 0; k < 6; ++k) r[k] = T(0);
  if (DIM == 1) {
    T inv = l1_inv(g.X(I), g.X(I + 1));
    T L = l1_dNdx<T>(0, inv) * vx[0] + l1_dNdx<T>(1, inv) * vx[1];
    r[0] = L;
    r[1] = L;
    r[3] = L + L;
  } else {
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
    r[0] = L00;
    r[1] = L11;
    r[3] = L01 + L10;
  }
  for (int k = 0; k < 6; ++k) r[k] = snap(r[k]);
}

template <typename T, int DIM>
__device__ __forceinline__ void gather_cell_velocity(const Geo<T>& g, const T* __restrict__ nb,
                                                     int64_t nld, int e, int& I, int& J,
                                                     T vx[4], T vy[4]) {
  if (DIM == 1) {
    I = e;
    J = 0;
    vx[0] = NROW(DMPM_N_V)[e];
    vx[1] = NROW(DMPM_N_V)[e + 1];
    vy[0] = vy[1] = T(0);
  } else {
    I = e % g.nx;
    J = e / g.nx;
    int ll = J * (g.nx + 1) + I;
    int ids[4] = {ll, ll + 1, ll + g.nx + 2, ll + g.nx + 1};
    for (int k = 0; k < 4; ++k) {
      vx[k] = NROW(DMPM_N_V)[ids[k]];
      vy[k] = NROW(DMPM_N_V + 1)[ids[k]];
    }
  }
}

template <typename T, int DIM>
__global__ void k_stress(Geo<T> g, Mat mat, T* __restrict__ pb, int64_t ld, int n,
                         const int* __restrict__ eid, const T* __restrict__ nb, int64_t nld,
                         T dt, int mode) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int e = (mode & S_STRAIN) ? eid[p] : 0;
  bool inside = e >= 0 && e < g.ncells;
  T d[6];
  T dvol;
  if (mode & S_STRAIN) {
    T r[6], rc[6];
    if (inside) {
      int I, J;
      T vx[4], vy[4];
      gather_cell_velocity<T, DIM>(g, nb, nld, e, I, J, vx, vy);
      T xi = PROW(DMPM_P_XI)[p];
      T eta = DIM == 2 ? PROW(DMPM_P_XI + 1)[p] : T(0);
      strain_rate<T, DIM>(g, I, J, xi, eta, vx, vy, r);
      strain_rate<T, DIM>(g, I, J, T(0), T(0), vx, vy, rc);
    } else {
      for (int k = 0; k < 6; ++k) r[k] = rc[k] = qnan<T>();
    }
    for (int k = 0; k < 6; ++k) d[k] = r[k] * dt;
    // particle.py:267-279
    T sc = DIM == 2 ? rc[0] + rc[1] : rc[0];
    dvol = dt * sc;
    if (mode & S_TEMPS) {
      for (int k = 0; k < 6; ++k) {
        PROW(DMPM_P_STRAIN_RATE + k)[p] = r[k];
        PROW(DMPM_P_DSTRAIN + k)[p] = d[k];
      }
      PROW(DMPM_P_DVOLSTRAIN)[p] = dvol;
    }
    // strain += dstrain; the rows with a structurally zero increment keep
    // their value (x + 0 == x).
    PROW(DMPM_P_STRAIN + 0)[p] = PROW(DMPM_P_STRAIN + 0)[p] + d[0];
    PROW(DMPM_P_STRAIN + 1)[p] = PROW(DMPM_P_STRAIN + 1)[p] + d[1];
    PROW(DMPM_P_STRAIN + 3)[p] = PROW(DMPM_P_STRAIN + 3)[p] + d[3];
    if (!inside) {
      PROW(DMPM_P_STRAIN + 2)[p] = qnan<T>();
      PROW(DMPM_P_STRAIN + 4)[p] = qnan<T>();
      PROW(DMPM_P_STRAIN + 5)[p] = qnan<T>();
    }
    PROW(DMPM_P_VOLSTRAIN_C)[p] = PROW(DMPM_P_VOLSTRAIN_C)[p] + dvol;
  } else {
    for (int k = 0; k < 6; ++k) d[k] = PROW(DMPM_P_DSTRAIN + k)[p];
    dvol = PROW(DMPM_P_DVOLSTRAIN)[p];
  }
  if (mode & S_VOLUME) {  // particle.py:324-327
    T f = T(1) + dvol;
    PROW(DMPM_P_VOLUME)[p] = PROW(DMPM_P_VOLUME)[p] * f;
    PROW(DMPM_P_DENSITY)[p] = PROW(DMPM_P_DENSITY)[p] / f;
  }
  if (mode & S_STRESS) {  // particle.py:315-322
    T ds[6];
    stress_increment(mat, pb, ld, p, d, ds);
    int nk = (mat.kind == DMPM_SIMPLE_MATERIAL || !(mode & S_STRAIN)) ? 6 : 4;
    for (int k = 0; k < nk; ++k) PROW(DMPM_P_STRESS + k)[p] = PROW(DMPM_P_STRESS + k)[p] + ds[k];
    if (nk == 4 && !inside) {
      PROW(DMPM_P_STRESS + 4)[p] = qnan<T>();
      PROW(DMPM_P_STRESS + 5)[p] = qnan<T>();
    }
  }
}

// ---- G2P (A17) ---------------------------------------------------------------
template <typename T, int DIM>
__global__ void k_g2p(Geo<T> g, T* __restrict__ pb, int64_t ld, int n,
                      const int* __restrict__ eid, const T* __restrict__ nb, int64_t nld,
                      T dt, int velocity_update, int write_momentum) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int e = eid[p];
  T nvx, nvy = T(0), nax, nay = T(0);
  if (e < 0 || e >= g.ncells) {
    nvx = nvy = nax = nay = qnan<T>();
  } else if (DIM == 1) {
    T xi = PROW(DMPM_P_XI)[p];
    T N0 = l1_N<T>(0, xi), N1 = l1_N<T>(1, xi);
    nvx = N0 * NROW(DMPM_N_V)[e] + N1 * NROW(DMPM_N_V)[e + 1];
    nax = N0 * NROW(DMPM_N_A)[e] + N1 * NROW(DMPM_N_A)[e + 1];
  } else {
    int I = e % g.nx, J = e / g.nx;
    int ll = J * (g.nx + 1) + I;
    int ids[4] = {ll, ll + 1, ll + g.nx + 2, ll + g.nx + 1};
    T N[4];
    q4_N4(PROW(DMPM_P_XI)[p], PROW(DMPM_P_XI + 1)[p], N);
    nvx = nvy = nax = nay = T(0);
    for (int k = 0; k < 4; ++k) {
      nvx = nvx + N[k] * NROW(DMPM_N_V)[ids[k]];
      nvy = nvy + N[k] * NROW(DMPM_N_V + 1)[ids[k]];
      nax = nax + N[k] * NROW(DMPM_N_A)[ids[k]];
      nay = nay + N[k] * NROW(DMPM_N_A + 1)[ids[k]];
    }
  }
  T vx = velocity_update ? nvx : PROW(DMPM_P_V)[p] + nax * dt;
  PROW(DMPM_P_V)[p] = vx;
  PROW(DMPM_P_X)[p] = PROW(DMPM_P_X)[p] + nvx * dt;
  if (write_momentum) PROW(DMPM_P_MOMENTUM)[p] = PROW(DMPM_P_MASS)[p] * vx;
  if (DIM == 2) {
    T vy = velocity_update ? nvy : PROW(DMPM_P_V + 1)[p] + nay * dt;
    PROW(DMPM_P_V + 1)[p] = vy;
    PROW(DMPM_P_X + 1)[p] = PROW(DMPM_P_X + 1)[p] + nvy * dt;
    if (write_momentum) PROW(DMPM_P_MOMENTUM + 1)[p] = PROW(DMPM_P_MASS)[p] * vy;
  }
}

// ---- traction (A13, particle side) ------------------------------------------
template <typename T>
__global__ void k_traction_zero(T* __restrict__ pb, int64_t ld, int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  PROW(DMPM_P_TRACTION)[p] = T(0);
  PROW(DMPM_P_TRACTION + 1)[p] = T(0);
}

// One thread per distinct (particle, dir); its terms are applied in list order.
template <typename T>
__global__ void k_traction_assign(T* __restrict__ pb, int64_t ld, int ngroups,
                                  const int* __restrict__ gid, const int* __restrict__ gdir,
                                  const int* __restrict__ goff, const int* __restrict__ tfn,
                                  const double* __restrict__ tval, FnTable fns, T t) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ngroups) return;
  int p = gid[q], dir = gdir[q];
  T* row = PROW(DMPM_P_TRACTION + dir);
  T acc = row[p];
  T V = PROW(DMPM_P_VOLUME)[p];
  T sz = PROW(DMPM_P_SIZE + dir)[p];
  for (int k = goff[q]; k < goff[q + 1]; ++k) {
    T tv = fn_value<T>(fns, tfn[k], t) * T(tval[k]);  // mesh.py:69-70
    acc = acc + (tv * V) / sz;                          // particle.py:341-343
  }
  row[p] = acc;
}

// ----------------------------------------------------------------------------
// Node-pull P2G kernels: one thread per node gathers the particles of its
// adjacent cells from the sorted order.
// ----------------------------------------------------------------------------
enum { K_MASS = 1, K_MOM = 2, K_VEL = 4 };
enum { F_EXT_ZERO = 1, F_PFEXT = 2, F_BODY = 4, F_TRAC = 8, F_INT = 16 };

template <int DIM> struct CellList {
  int cell[4];
  int corner[4];
  int count;
};

// Adjacent cells of node `nd` in the canonical order.
template <typename T, int DIM>
__device__ __forceinline__ CellList<DIM> node_cells(const Geo<T>& g, int nd) {
  CellList<DIM> L;
  L.count = 0;
  if (DIM == 1) {
    int i = nd;
    if (i - 1 >= 0) { L.cell[L.count] = i - 1; L.corner[L.count++] = 1; }
    if (i < g.nx) { L.cell[L.count] = i; L.corner[L.count++] = 0; }
  } else {
    int i = nd % (g.nx + 1), j = nd / (g.nx + 1);
    if (j - 1 >= 0) {
      if (i - 1 >= 0) { L.cell[L.count] = (j - 1) * g.nx + i - 1; L.corner[L.count++] = 2; }
      if (i < g.nx) { L.cell[L.count] = (j - 1) * g.nx + i; L.corner[L.count++] = 3; }
    }
    if (j < g.ny) {
      if (i - 1 >= 0) { L.cell[L.count] = j * g.nx + i - 1; L.corner[L.count++] = 1; }
      if (i < g.nx) { L.cell[L.count] = j * g.nx + i; L.corner[L.count++] = 0; }
    }
  }
  return L;
}

template <typename T, int DIM>
__device__ __forceinline__ T corner_N(int corner, T xi, T eta) {
  return DIM == 1 ? l1_N<T>(corner, xi) : q4_N<T>(corner, xi, eta);
}

template <typename T, int DIM>
__global__ void k_p2g_kin(Geo<T> g, const T* __restrict__ pb, int64_t ld, SortRef sr,
                          T* __restrict__ nb, int64_t nld, int mode) {
  int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= g.nnodes) return;
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  CellList<DIM> L = node_cells<T, DIM>(g, nd);
  T am = T(0), apx = T(0), apy = T(0);
  if (mode & (K_MASS | K_MOM)) {
    for (int c = 0; c < L.count; ++c) {  // per-cell partial sums, then the node sum
      int b = cs[L.cell[c]], e = cs[L.cell[c] + 1];
      T cm = T(0), cpx = T(0), cpy = T(0);
      for (int s = b; s < e; ++s) {
        int p = perm[s];
        T xi = PROW(DMPM_P_XI)[p];
        T eta = DIM == 2 ? PROW(DMPM_P_XI + 1)[p] : T(0);
        T N = corner_N<T, DIM>(L.corner[c], xi, eta);
        T m = PROW(DMPM_P_MASS)[p];
        if (mode & K_MASS) cm = cm + m * N;  // element.py:140
        if (mode & K_MOM) {                   // element.py:172 with pmom = m*v
          cpx = cpx + N * (m * PROW(DMPM_P_V)[p]);
          if (DIM == 2) cpy = cpy + N * (m * PROW(DMPM_P_V + 1)[p]);
        }
      }
      am = am + cm;
      apx = apx + cpx;
      apy = apy + cpy;
    }
  }
  if (mode & K_MASS) NROW(DMPM_N_MASS)[nd] = am;
  if (mode & K_MOM) {
    apx = snap(apx);
    apy = snap(apy);
    NROW(DMPM_N_P)[nd] = apx;
    if (DIM == 2) NROW(DMPM_N_P + 1)[nd] = apy;
  }
  if (mode & K_VEL) {  // element.py:191-202
    T m = (mode & K_MASS) ? am : NROW(DMPM_N_MASS)[nd];
    T px = (mode & K_MOM) ? apx : NROW(DMPM_N_P)[nd];
    T vx = (m == T(0)) ? NROW(DMPM_N_V)[nd] : px / m;
    NROW(DMPM_N_V)[nd] = snap(vx);
    if (DIM == 2) {
      T py = (mode & K_MOM) ? apy : NROW(DMPM_N_P + 1)[nd];
      T vy = (m == T(0)) ? NROW(DMPM_N_V + 1)[nd] : py / m;
      NROW(DMPM_N_V + 1)[nd] = snap(vy);
    }
  }
}

template <typename T, int DIM>
__global__ void k_p2g_force(Geo<T> g, const T* __restrict__ pb, int64_t ld, SortRef sr,
                            T* __restrict__ nb, int64_t nld, int mode, T gx, T gy) {
  int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= g.nnodes) return;
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  CellList<DIM> L = node_cells<T, DIM>(g, nd);
  if (mode & (F_EXT_ZERO | F_PFEXT | F_BODY | F_TRAC)) {
    T fx = (mode & F_EXT_ZERO) ? T(0) : NROW(DMPM_N_FEXT)[nd];
    T fy = (DIM == 2 && !(mode & F_EXT_ZERO)) ? NROW(DMPM_N_FEXT + 1)[nd] : T(0);
    // Three passes so each source accumulates in the reference's loop order
    // (element.py:204-303: external, then body, then traction).
    for (int pass = 0; pass < 3; ++pass) {
      int bit = pass == 0 ? F_PFEXT : (pass == 1 ? F_BODY : F_TRAC);
      if (!(mode & bit)) continue;
      for (int c = 0; c < L.count; ++c) {
        int b = cs[L.cell[c]], e = cs[L.cell[c] + 1];
        T cx = T(0), cy = T(0);
        for (int s = b; s < e; ++s) {
          int p = perm[s];
          T xi = PROW(DMPM_P_XI)[p];
          T eta = DIM == 2 ? PROW(DMPM_P_XI + 1)[p] : T(0);
          T N = corner_N<T, DIM>(L.corner[c], xi, eta);
          T qx, qy = T(0);
          if (bit == F_PFEXT) {
            qx = PROW(DMPM_P_FEXT)[p];
            if (DIM == 2) qy = PROW(DMPM_P_FEXT + 1)[p];
          } else if (bit == F_BODY) {
            T m = PROW(DMPM_P_MASS)[p];
            qx = m * gx;
            qy = m * gy;
          } else {
            qx = PROW(DMPM_P_TRACTION)[p];
            if (DIM == 2) qy = PROW(DMPM_P_TRACTION + 1)[p];
          }
          cx = cx + N * qx;
          if (DIM == 2) cy = cy + N * qy;
        }
        fx = fx + cx;
        if (DIM == 2) fy = fy + cy;
      }
    }
    NROW(DMPM_N_FEXT)[nd] = fx;
    if (DIM == 2) NROW(DMPM_N_FEXT + 1)[nd] = fy;
  }
  if (mode & F_INT) {  // element.py:548-600 / 842-901
    T fx = T(0), fy = T(0);
    for (int c = 0; c < L.count; ++c) {
      int cell = L.cell[c];
      int b = cs[cell], e = cs[cell + 1];
      int I = DIM == 1 ? cell : cell % g.nx;
      int J = DIM == 1 ? 0 : cell / g.nx;
      T cx = T(0), cy = T(0);
      for (int s = b; s < e; ++s) {
        int p = perm[s];
        T xi = PROW(DMPM_P_XI)[p];
        T nV = -PROW(DMPM_P_VOLUME)[p];
        T sxx = PROW(DMPM_P_STRESS + 0)[p];
        if (DIM == 1) {
          T dN = l1_dNdx<T>(L.corner[c], l1_inv(g.X(I), g.X(I + 1)));
          cx = cx + (nV * sxx) * dN;
        } else {
          T eta = PROW(DMPM_P_XI + 1)[p];
          T ix, iy, dx, dy;
          q4_invJ(xi, eta, g.X(I), g.X(I + 1), g.Y(J), g.Y(J + 1), ix, iy);
          q4_dNdx(L.corner[c], xi, eta, ix, iy, dx, dy);
          T syy = PROW(DMPM_P_STRESS + 1)[p];
          T sxy = PROW(DMPM_P_STRESS + 3)[p];
          cx = cx + nV * (dx * sxx + dy * sxy);
          cy = cy + nV * (dy * syy + dx * sxy);
        }
      }
      fx = fx + cx;
      fy = fy + cy;
    }
    NROW(DMPM_N_FINT)[nd] = fx;
    if (DIM == 2) NROW(DMPM_N_FINT + 1)[nd] = fy;
  }
}

// ---- grid kernels (A7 constraints, A14 nodal forces, A16 update) -------------
template <typename T>
__global__ void k_constraints(T* __restrict__ nb, int64_t nld, int nc,
                              const int* __restrict__ node, const int* __restrict__ dir,
                              const double* __restrict__ val, int update_mode) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  int i = node[q], d = dir[q];
  T c = T(val[q]);
  T m = NROW(DMPM_N_MASS)[i];
  // constraint.py:38-44; in the nodal update the velocity is snapped after
  // the constraints (element.py:316-324).
  NROW(DMPM_N_V + d)[i] = update_mode ? snap(c) : c;
  NROW(DMPM_N_P + d)[i] = m * c;
  NROW(DMPM_N_A + d)[i] = T(0);
}

template <typename T>
__global__ void k_nodal_forces(T* __restrict__ nb, int64_t nld, int ngroups,
                               const int* __restrict__ gid, const int* __restrict__ gdir,
                               const int* __restrict__ goff, const int* __restrict__ ffn,
                               const double* __restrict__ fval, FnTable fns, T t) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ngroups) return;
  T* row = NROW(DMPM_N_FEXT + gdir[q]);
  int i = gid[q];
  T acc = row[i];
  for (int k = goff[q]; k < goff[q + 1]; ++k)
    acc = acc + fn_value<T>(fns, ffn[k], t) * T(fval[k]);  // element.py:280-284
  row[i] = acc;
}

template <typename T, int DIM>
__global__ void k_grid_update(T* __restrict__ nb, int64_t nld, int nnodes, T dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  T m = NROW(DMPM_N_MASS)[i];
  for (int d = 0; d < DIM; ++d) {
    T F = (NROW(DMPM_N_FINT + d)[i] + NROW(DMPM_N_FEXT + d)[i]) + NROW(DMPM_N_FDAMP + d)[i];
    T a = nan_to_num(F / m);
    T v = NROW(DMPM_N_V + d)[i] + a * dt;
    NROW(DMPM_N_P + d)[i] = m * v;
    NROW(DMPM_N_V + d)[i] = snap(v);
    NROW(DMPM_N_A + d)[i] = snap(a);
  }
}

#undef PROW
#undef NROW

// ----------------------------------------------------------------------------
// plan management
// ----------------------------------------------------------------------------
namespace dmpm {

template <typename T> Geo<T> make_geo(const dmpm_plan* P) {
  Geo<T> g;
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

Mat make_mat(const dmpm_plan* P) {
  Mat m;
  m.kind = P->d.material;
  m.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) m.de[k] = P->d.de[k];
  return m;
}

FnTable fn_table(const dmpm_plan* P) {
  FnTable f;
  f.kind = P->fn_kind;
  f.off = P->fn_off;
  f.cnt = P->fn_cnt;
  f.xs = P->fn_x;
  f.fs = P->fn_f;
  return f;
}

template <typename T> static int dev_upload(T** dst, const std::vector<T>& v) {
  if (*dst) {
    (void)hipFree(*dst);
    *dst = nullptr;
  }
  if (v.empty()) return DMPM_OK;
  DMPM_TRY(hip_check(hipMalloc((void**)dst, v.size() * sizeof(T)), "hipMalloc"));
  return hip_check(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice),
                   "hipMemcpy");
}

static void free_ptr(void* p) {
  if (p) (void)hipFree(p);
}

void plan_free(dmpm_plan* P) {
  if (!P) return;
  void* ptrs[] = {P->key, P->perm, P->counts, P->fill, P->cs, P->bsum, P->flags, P->zflags,
                  P->big,
                  P->perm2, P->cs2, P->mov, P->mrec, P->dL, P->dA, P->headL, P->headA,
                  P->chunkD, P->chunkC, P->subT, P->sims, P->tile_sim, P->clock,
                  P->con_node, P->con_dir, P->con_val, P->nf_id, P->nf_dir, P->nf_off,
                  P->nf_fn, P->nf_val, P->tr_id, P->tr_dir, P->tr_off, P->tr_fn, P->tr_val,
                  P->fn_kind, P->fn_off, P->fn_cnt, P->fn_x, P->fn_f, P->cidx, P->cmask,
                  P->cval, P->nfd};
  for (void* p : ptrs) free_ptr(p);
  P->release_extra();
  delete P;
}

// Groups ordered (id, dir, fn, value) terms by (id, dir), preserving order
// within each group, so that one thread per group reproduces the reference's
// sequential scatter exactly.
struct Grouped {
  std::vector<int> id, dir, off, fn;
  std::vector<double> val;
};
static Grouped group_terms(int32_t n, const int32_t* ids, const int32_t* dirs,
                           const int32_t* fn, const double* vals) {
  std::map<std::pair<int, int>, std::vector<int>> g;
  std::vector<std::pair<int, int>> order;
  for (int k = 0; k < n; ++k) {
    auto key = std::make_pair(ids[k], dirs[k]);
    auto it = g.find(key);
    if (it == g.end()) {
      order.push_back(key);
      g[key] = {k};
    } else {
      it->second.push_back(k);
    }
  }
  Grouped out;
  out.off.push_back(0);
  for (auto& key : order) {
    out.id.push_back(key.first);
    out.dir.push_back(key.second);
    for (int k : g[key]) {
      out.fn.push_back(fn ? fn[k] : -1);
      out.val.push_back(vals[k]);
    }
    out.off.push_back((int)out.fn.size());
  }
  return out;
}

}  // namespace dmpm

extern "C" {

const char* dmpm_last_error(void) { return g_last_error.c_str(); }

int dmpm_abi_version(void) { return 101; }

int dmpm_snapshot_rows(const void* pbuf, int64_t ld, int32_t dtype, int64_t n, void* host_dst,
                       void* stream) {
  if (!pbuf || !host_dst) return fail(DMPM_EINVAL, "null buffer");
  if (dtype != DMPM_F32 && dtype != DMPM_F64) return fail(DMPM_EINVAL, "dtype");
  if (n < 0 || ld < n) return fail(DMPM_EINVAL, "n must lie in [0, ld]");
  if (n == 0) return DMPM_OK;
  const size_t esz = dtype == DMPM_F32 ? 4 : 8;
  const size_t row = (size_t)n * esz, pitch = (size_t)ld * esz;
  const char* src = (const char*)pbuf;
  char* dst = (char*)host_dst;
  hipStream_t s = (hipStream_t)stream;
  // rows 0..3 (loc, velocity), then rows 13..24 (strain, stress)
  DMPM_TRY(hip_check(hipMemcpy2DAsync(dst, row, src, pitch, row, 4, hipMemcpyDeviceToHost, s),
                     "hipMemcpy2DAsync"));
  return hip_check(hipMemcpy2DAsync(dst + 4 * row, row, src + DMPM_P_STRAIN * pitch, pitch, row,
                                    12, hipMemcpyDeviceToHost, s),
                   "hipMemcpy2DAsync");
}

// Allocates and initialises the sort state and the simulation table of a
// plan whose sizes are set; frees it on failure.
static int plan_init(dmpm_plan* P, dmpm_plan** out) {
  if (hipGetDevice(&P->device) != hipSuccess) {
    plan_free(P);
    return fail(DMPM_EHIP, "hipGetDevice failed");
  }
  // the device address of dmpm_step_between_noalloc's scratch array on this
  // device (taken once per device, with it current)
  if (noalloc_scratch_setup(P->device) != DMPM_OK) {
    plan_free(P);
    return DMPM_EHIP;
  }
  int64_t n = std::max<int64_t>(P->n, 1);
  int nb = P->ncells + 1;
  int nbs = (nb + kScanChunk - 1) / kScanChunk;
  int rc = DMPM_OK;
  auto al = [&](int** p, int64_t cnt) {
    if (rc == DMPM_OK) rc = hip_check(hipMalloc((void**)p, cnt * sizeof(int)), "hipMalloc");
  };
  al(&P->key, n);
  al(&P->perm, n);
  al(&P->counts, nb);
  al(&P->fill, nb);
  al(&P->cs, nb + 1);
  al(&P->bsum, nbs);
  al(&P->flags, kFlagWords);
  al(&P->zflags, kFlagWords);
  al(&P->big, P->ncells);
  al(&P->perm2, n);
  al(&P->cs2, nb + 1);
  P->movcap = (int)std::max<int64_t>(1, P->n);
  al(&P->mov, (int64_t)kMovRec * P->movcap);
  al(&P->mrec, n);
  al(&P->dL, nb);
  al(&P->dA, nb);
  al(&P->headL, nb);
  al(&P->headA, nb);
  P->chunk_shift = kMinChunkShift;
  while (((int64_t)nb + (1 << P->chunk_shift) - 1) >> P->chunk_shift > kMaxChunks &&
         P->chunk_shift < 20)
    ++P->chunk_shift;
  P->nchunks = (int)((nb + (1 << P->chunk_shift) - 1) >> P->chunk_shift);
  al(&P->chunkD, P->nchunks);
  al(&P->chunkC, P->nchunks);
  P->nsub = (nb + (1 << kSubShift) - 1) >> kSubShift;
  al(&P->subT, P->nsub);
  auto zero = [&](void* p, int64_t bytes, int v) {
    if (rc == DMPM_OK && p) rc = hip_check(hipMemset(p, v, bytes), "hipMemset");
  };
  zero(P->zflags, kFlagWords * sizeof(int), 0);
  zero(P->counts, nb * sizeof(int), 0);
  zero(P->fill, nb * sizeof(int), 0);
  zero(P->mrec, n * sizeof(int), 0);
  zero(P->dL, nb * sizeof(int), 0);
  zero(P->dA, nb * sizeof(int), 0);
  zero(P->headL, nb * sizeof(int), 0xff);
  zero(P->headA, nb * sizeof(int), 0xff);
  zero(P->chunkD, P->nchunks * sizeof(int), 0);
  zero(P->chunkC, P->nchunks * sizeof(int), 0);
  if (rc == DMPM_OK) rc = hip_check(hipMemset(P->key, 0xff, n * sizeof(int)), "hipMemset");
  if (rc == DMPM_OK) rc = hip_check(hipMemset(P->cs, 0, (nb + 1) * sizeof(int)), "hipMemset");
  if (rc == DMPM_OK) rc = hip_check(hipMemset(P->perm, 0, n * sizeof(int)), "hipMemset");
  P->nsims = (int)P->sims_h.size();
  // tiles of the one-launch step, numbered simulation after simulation (row
  // major within a mesh)
  int64_t ntiles = 0;
  for (SimRec& r : P->sims_h) {
    r.tiles_x = (r.nx + kStepTile - 1) / kStepTile;
    r.tile_off = (int)ntiles;
    ntiles += (int64_t)r.tiles_x * ((r.ny + kStepTile - 1) / kStepTile);
  }
  P->ntiles = (int)ntiles;
  if (rc == DMPM_OK) rc = hip_check(hipMalloc((void**)&P->clock, sizeof(int64_t)), "hipMalloc");
  if (rc == DMPM_OK) rc = hip_check(hipMemset(P->clock, 0, sizeof(int64_t)), "hipMemset");
  if (rc == DMPM_OK && P->nsims > 1) {
    std::vector<int> ts((size_t)ntiles);
    for (int s = 0; s < P->nsims; ++s) {
      const SimRec& r = P->sims_h[s];
      const int64_t end = s + 1 < P->nsims ? P->sims_h[s + 1].tile_off : ntiles;
      for (int64_t k = r.tile_off; k < end; ++k) ts[(size_t)k] = s;
    }
    al(&P->tile_sim, ntiles);
    if (rc == DMPM_OK)
      rc = hip_check(hipMemcpy(P->tile_sim, ts.data(), ts.size() * sizeof(int),
                               hipMemcpyHostToDevice),
                     "hipMemcpy");
  }
  if (rc == DMPM_OK)
    rc = hip_check(hipMalloc((void**)&P->sims, P->sims_h.size() * sizeof(SimRec)), "hipMalloc");
  if (rc == DMPM_OK)
    rc = hip_check(hipMemcpy(P->sims, P->sims_h.data(), P->sims_h.size() * sizeof(SimRec),
                             hipMemcpyHostToDevice),
                   "hipMemcpy");
  int f0[kFlagWords] = {1};
  if (rc == DMPM_OK)
    rc = hip_check(hipMemcpy(P->flags, f0, sizeof(f0), hipMemcpyHostToDevice), "hipMemcpy");
  if (rc != DMPM_OK) {
    plan_free(P);
    return rc;
  }
  *out = P;
  return DMPM_OK;
}

int dmpm_plan_create(const dmpm_plan_desc* desc, dmpm_plan** out) {
  if (!desc || !out) return fail(DMPM_EINVAL, "null argument");
  const dmpm_plan_desc& d = *desc;
  if (d.element != DMPM_LINEAR1D && d.element != DMPM_QUAD4)
    return fail(DMPM_EINVAL, "element must be DMPM_LINEAR1D or DMPM_QUAD4");
  if (d.dtype != DMPM_F32 && d.dtype != DMPM_F64)
    return fail(DMPM_EINVAL, "dtype must be DMPM_F32 or DMPM_F64");
  if (d.nx < 1 || d.ny < 1 || (d.element == DMPM_LINEAR1D && d.ny != 1))
    return fail(DMPM_EINVAL, "invalid mesh size");
  if (!(d.hx > 0) || (d.element == DMPM_QUAD4 && !(d.hy > 0)))
    return fail(DMPM_EINVAL, "element lengths must be positive");
  if ((int64_t)d.nx * d.ny >= (1ll << 30)) return fail(DMPM_EINVAL, "mesh too large");
  if (d.nparticles < 0 || d.nparticles >= (1ll << 31) || d.ld < d.nparticles)
    return fail(DMPM_EINVAL, "invalid particle count or row stride");
  if (d.material != DMPM_LINEAR_ELASTIC && d.material != DMPM_SIMPLE_MATERIAL)
    return fail(DMPM_EINVAL, "unknown material");
  if (d.material == DMPM_SIMPLE_MATERIAL && d.per_particle_E)
    return fail(DMPM_EINVAL,
                "per-particle E takes DMPM_LINEAR_ELASTIC coefficients: de = (1, 0, 1) gives "
                "SimpleMaterial's dsigma = E_p dstrain");
  dmpm_plan* P = new dmpm_plan();
  P->d = d;
  P->n = d.nparticles;
  P->dim = d.element == DMPM_LINEAR1D ? 1 : 2;
  P->ncells = d.nx * d.ny;
  P->nnodes = P->dim == 1 ? d.nx + 1 : (d.nx + 1) * (d.ny + 1);
  if (d.nld < P->nnodes) {
    delete P;
    return fail(DMPM_EINVAL, "node row stride smaller than the node count");
  }
  SimRec r = {};
  r.nx = d.nx;
  r.ny = d.ny;
  r.n = (int)d.nparticles;
  r.hx = d.hx;
  r.hy = d.hy;
  for (int k = 0; k < 3; ++k) r.de[k] = d.de[k];
  P->sims_h.assign(1, r);
  return plan_init(P, out);
}

int dmpm_batch_plan_create(const dmpm_plan_desc* descs, int32_t nsims, int64_t ld, int64_t nld,
                           dmpm_plan** out) {
  if (!descs || !out || nsims < 1) return fail(DMPM_EINVAL, "null argument or empty batch");
  const dmpm_plan_desc& d0 = descs[0];
  std::vector<SimRec> recs(nsims);
  int64_t np = 0, nc = 0, nn = 0;
  for (int s = 0; s < nsims; ++s) {
    const dmpm_plan_desc& d = descs[s];
    if (d.element != DMPM_QUAD4) return fail(DMPM_EINVAL, "batch plans hold DMPM_QUAD4 meshes");
    if (d.dtype != d0.dtype || d.material != d0.material ||
        d.per_particle_E != d0.per_particle_E)
      return fail(DMPM_EINVAL, "the simulations of a batch share dtype and material kind");
    if (d.dtype != DMPM_F32 && d.dtype != DMPM_F64)
      return fail(DMPM_EINVAL, "dtype must be DMPM_F32 or DMPM_F64");
    if (d.material != DMPM_LINEAR_ELASTIC && d.material != DMPM_SIMPLE_MATERIAL)
      return fail(DMPM_EINVAL, "unknown material");
    if (d.nx < 1 || d.ny < 1 || !(d.hx > 0) || !(d.hy > 0) || d.nparticles < 0)
      return fail(DMPM_EINVAL, "invalid simulation in batch");
    SimRec& r = recs[s];
    r.nx = d.nx;
    r.ny = d.ny;
    r.n = (int)d.nparticles;
    r.hx = d.hx;
    r.hy = d.hy;
    for (int k = 0; k < 3; ++k) r.de[k] = d.de[k];
    r.part_off = (int)np;
    r.cell_off = (int)nc;
    r.node_off = (int)nn;
    np += d.nparticles;
    nc += (int64_t)d.nx * d.ny;
    nn += (int64_t)(d.nx + 1) * (d.ny + 1);
  }
  if (np >= (1ll << 31) || nc >= (1ll << 30) || nn >= (1ll << 30))
    return fail(DMPM_EINVAL, "batch too large");
  if (ld < np || nld < nn) return fail(DMPM_EINVAL, "row strides smaller than the batch");
  dmpm_plan* P = new dmpm_plan();
  P->d = d0;
  P->d.nparticles = np;
  P->d.ld = ld;
  P->d.nld = nld;
  P->n = np;
  P->dim = 2;
  P->ncells = (int)nc;
  P->nnodes = (int)nn;
  P->sims_h = recs;
  return plan_init(P, out);
}


static void prof_off(dmpm_plan* P) {
  for (auto& c : P->prof) {
    for (int k = 0; k < kProfRing; ++k) prof_drain_slot(c, k);
    if (c.on)
      for (auto& e : c.ev) (void)hipEventDestroy(e);
    c = dmpm_plan::ProfClass();
  }
  P->prof_single = -1;
}

int dmpm_plan_destroy(dmpm_plan* plan) {
  if (plan) {
    prof_off(plan);
  }
  plan_free(plan);
  return DMPM_OK;
}

int dmpm_timing_enable(dmpm_plan* P, int32_t kernel_class, int32_t max_launches) {
  (void)max_launches;
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (kernel_class >= kNumKC || kernel_class < DMPM_KC_ALL)
    return fail(DMPM_EINVAL, "unknown kernel class");
  prof_off(P);
  if (kernel_class == -1) return DMPM_OK;
  P->prof_single = kernel_class;
  for (int k = 0; k < kNumKC; ++k) {
    if (kernel_class != DMPM_KC_ALL && k != kernel_class) continue;
    dmpm_plan::ProfClass& c = P->prof[k];
    for (auto& e : c.ev) DMPM_TRY(hip_check(hipEventCreate(&e), "hipEventCreate"));
    c.on = true;
  }
  return DMPM_OK;
}

int dmpm_timing_read_class(dmpm_plan* P, int32_t kernel_class, double* total_ms,
                           int64_t* launches) {
  if (!P || !total_ms || !launches) return fail(DMPM_EINVAL, "null argument");
  if (kernel_class < 0 || kernel_class >= kNumKC) return fail(DMPM_EINVAL, "unknown kernel class");
  dmpm_plan::ProfClass& c = P->prof[kernel_class];
  for (int k = 0; k < kProfRing; ++k) prof_drain_slot(c, k);
  *total_ms = c.ms;
  *launches = c.n;
  return DMPM_OK;
}

int dmpm_timing_read(dmpm_plan* P, double* total_ms, int64_t* launches) {
  if (!P || !total_ms || !launches) return fail(DMPM_EINVAL, "null argument");
  *total_ms = 0;
  *launches = 0;
  for (int k = 0; k < kNumKC; ++k) {
    if (P->prof_single != DMPM_KC_ALL && k != P->prof_single) continue;
    double ms = 0;
    int64_t n = 0;
    DMPM_TRY(dmpm_timing_read_class(P, k, &ms, &n));
    *total_ms += ms;
    *launches += n;
  }
  return DMPM_OK;
}

int dmpm_plan_set_functions(dmpm_plan* P, int32_t nfn, const int32_t* kinds,
                            const int32_t* offsets, const int32_t* counts, const double* xs,
                            const double* fs) {
  if (!P || nfn < 0) return fail(DMPM_EINVAL, "invalid function table");
  std::vector<int> k(kinds, kinds + nfn), o(offsets, offsets + nfn), c(counts, counts + nfn);
  int tot = 0;
  for (int i = 0; i < nfn; ++i) {
    if (k[i] != DMPM_FN_UNIT && k[i] != DMPM_FN_LINEAR)
      return fail(DMPM_EINVAL, "unknown function kind");
    if (k[i] == DMPM_FN_LINEAR && c[i] < 1) return fail(DMPM_EINVAL, "empty Linear function");
    tot = std::max(tot, o[i] + c[i]);
  }
  std::vector<double> x(xs, xs + tot), f(fs, fs + tot);
  DMPM_TRY(dev_upload(&P->fn_kind, k));
  DMPM_TRY(dev_upload(&P->fn_off, o));
  DMPM_TRY(dev_upload(&P->fn_cnt, c));
  DMPM_TRY(dev_upload(&P->fn_x, x));
  DMPM_TRY(dev_upload(&P->fn_f, f));
  P->nfn = nfn;
  return DMPM_OK;
}

int dmpm_plan_set_constraints(dmpm_plan* P, int32_t n, const int32_t* node_ids,
                              const int32_t* dirs, const double* velocities) {
  if (!P || n < 0) return fail(DMPM_EINVAL, "invalid constraint table");
  // Later entries overwrite earlier ones (sequential .at[].set), so only the
  // last value per (node, dir) matters.
  std::map<std::pair<int, int>, double> last;
  std::vector<std::pair<int, int>> order;
  for (int k = 0; k < n; ++k) {
    int id = node_ids[k];
    if (id < 0) id += P->nnodes;  // JAX normalises negative indices
    if (id < 0 || id >= P->nnodes) continue;  // and drops out-of-range ones
    if (dirs[k] < 0 || dirs[k] >= P->dim) return fail(DMPM_EINVAL, "constraint dir out of range");
    auto key = std::make_pair(id, (int)dirs[k]);
    if (!last.count(key)) order.push_back(key);
    last[key] = velocities[k];
  }
  std::vector<int> nd, dr;
  std::vector<double> vl;
  for (auto& key : order) {
    nd.push_back(key.first);
    dr.push_back(key.second);
    vl.push_back(last[key]);
  }
  DMPM_TRY(dev_upload(&P->con_node, nd));
  DMPM_TRY(dev_upload(&P->con_dir, dr));
  DMPM_TRY(dev_upload(&P->con_val, vl));
  P->ncon = (int)nd.size();
  // dense per-node records for the tile kernels
  std::vector<int> cidx(P->nnodes, -1), cmask;
  std::vector<double> cval;
  for (size_t k = 0; k < nd.size(); ++k) {
    int& r = cidx[nd[k]];
    if (r < 0) {
      r = (int)cmask.size();
      cmask.push_back(0);
      cval.push_back(0.0);
      cval.push_back(0.0);
    }
    cmask[r] |= 1 << dr[k];
    cval[2 * r + dr[k]] = vl[k];
  }
  DMPM_TRY(dev_upload(&P->cidx, cidx));
  DMPM_TRY(dev_upload(&P->cmask, cmask));
  DMPM_TRY(dev_upload(&P->cval, cval));
  return DMPM_OK;
}

int dmpm_plan_set_nodal_forces(dmpm_plan* P, int32_t n, const int32_t* node_ids,
                               const int32_t* dirs, const int32_t* fn, const double* forces) {
  if (!P || n < 0) return fail(DMPM_EINVAL, "invalid nodal force table");
  std::vector<int32_t> ids, ds, fs;
  std::vector<double> vs;
  for (int k = 0; k < n; ++k) {
    int id = node_ids[k];
    if (id < 0) id += P->nnodes;
    if (id < 0 || id >= P->nnodes) continue;
    if (dirs[k] < 0 || dirs[k] >= P->dim) return fail(DMPM_EINVAL, "force dir out of range");
    if (fn && fn[k] >= P->nfn) return fail(DMPM_EINVAL, "force function index out of range");
    ids.push_back(id);
    ds.push_back(dirs[k]);
    fs.push_back(fn ? fn[k] : -1);
    vs.push_back(forces[k]);
  }
  Grouped g = group_terms((int)ids.size(), ids.data(), ds.data(), fs.data(), vs.data());
  DMPM_TRY(dev_upload(&P->nf_id, g.id));
  DMPM_TRY(dev_upload(&P->nf_dir, g.dir));
  DMPM_TRY(dev_upload(&P->nf_off, g.off));
  DMPM_TRY(dev_upload(&P->nf_fn, g.fn));
  DMPM_TRY(dev_upload(&P->nf_val, g.val));
  P->nnf = (int)g.id.size();
  std::vector<int> nfd(2 * (size_t)P->nnodes, -1);
  for (size_t q = 0; q < g.id.size(); ++q) nfd[(size_t)g.dir[q] * P->nnodes + g.id[q]] = (int)q;
  DMPM_TRY(dev_upload(&P->nfd, nfd));
  return DMPM_OK;
}

int dmpm_plan_set_tractions(dmpm_plan* P, int32_t n, const int32_t* pids, const int32_t* dirs,
                            const int32_t* fn, const double* tractions) {
  if (!P || n < 0) return fail(DMPM_EINVAL, "invalid traction table");
  std::vector<int32_t> ids, ds, fs;
  std::vector<double> vs;
  for (int k = 0; k < n; ++k) {
    int64_t id = pids[k];
    if (id < 0) id += P->n;
    if (id < 0 || id >= P->n) continue;
    if (dirs[k] < 0 || dirs[k] >= P->dim) return fail(DMPM_EINVAL, "traction dir out of range");
    if (fn && fn[k] >= P->nfn) return fail(DMPM_EINVAL, "traction function index out of range");
    ids.push_back((int32_t)id);
    ds.push_back(dirs[k]);
    fs.push_back(fn ? fn[k] : -1);
    vs.push_back(tractions[k]);
  }
  Grouped g = group_terms((int)ids.size(), ids.data(), ds.data(), fs.data(), vs.data());
  DMPM_TRY(dev_upload(&P->tr_id, g.id));
  DMPM_TRY(dev_upload(&P->tr_dir, g.dir));
  DMPM_TRY(dev_upload(&P->tr_off, g.off));
  DMPM_TRY(dev_upload(&P->tr_fn, g.fn));
  DMPM_TRY(dev_upload(&P->tr_val, g.val));
  P->ntr = (int)g.id.size();
  P->has_tractions = n > 0;
  return DMPM_OK;
}

int dmpm_plan_set_block(dmpm_plan* P, int32_t cells_per_block) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (cells_per_block != 0 && cells_per_block != 64 && cells_per_block != 128 &&
      cells_per_block != 256)
    return fail(DMPM_EINVAL, "cells per block must be 64, 128 or 256 (0: default)");
  P->cell_block = cells_per_block ? cells_per_block : 128;
  return DMPM_OK;
}

int dmpm_plan_set_option(dmpm_plan* P, int32_t option, int32_t value) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  switch (option) {
    case DMPM_OPT_USF_SPLIT:
      P->usf_split = value ? 1 : 0;
      return DMPM_OK;
    case DMPM_OPT_STEP_KERNEL:
      if (value < 0 || value > 2) return fail(DMPM_EINVAL, "step kernel must be 0, 1 or 2");
      P->step_kernel = value;
      return DMPM_OK;
    default:
      return fail(DMPM_EINVAL, "unknown option");
  }
}

int dmpm_mark_unsorted(dmpm_plan* P, void* stream) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  static const int one = 1;
  return hip_check(hipMemcpyAsync(P->flags, &one, sizeof(int), hipMemcpyHostToDevice,
                                  (hipStream_t)stream),
                   "hipMemcpyAsync");
}

}  // extern "C"

// ----------------------------------------------------------------------------
// typed drivers
// ----------------------------------------------------------------------------
namespace dmpm {

template <typename T, int DIM> struct Ops {
  static int locate(dmpm_plan* P, void* pbuf, int32_t* eid, int write_xi, hipStream_t s) {
    ProfScope ps_(P, KC_LOCATE, s);
    if (P->n > 0)
      k_locate<T, DIM><<<nblocks(P->n), kBlock, 0, s>>>(make_geo<T>(P), (T*)pbuf, P->d.ld,
                                                         (int)P->n, eid, P->key, P->flags,
                                                         write_xi);
    return launch_check("locate");
  }
  static int natural(dmpm_plan* P, void* pbuf, const int32_t* eid, hipStream_t s) {
    if (P->n > 0)
      k_natural_coords<T, DIM><<<nblocks(P->n), kBlock, 0, s>>>(make_geo<T>(P), (T*)pbuf,
                                                                P->d.ld, (int)P->n, eid);
    return launch_check("natural_coords");
  }
  static int pvolume(dmpm_plan* P, void* pbuf, const int32_t* eid, double ve, hipStream_t s) {
    DMPM_TRY(ensure_sorted(P, eid, s));
    if (P->n > 0)
      k_particle_volume<T, DIM><<<nblocks(P->n), kBlock, 0, s>>>(
          make_geo<T>(P), (T*)pbuf, P->d.ld, (int)P->n, eid, sort_ref(P), T(ve));
    return launch_check("particle_volume");
  }
  static int kin(dmpm_plan* P, const void* pbuf, void* nbuf, int mode, hipStream_t s) {
    ProfScope ps_(P, KC_P2G_KIN, s);
    k_p2g_kin<T, DIM><<<nblocks(P->nnodes), kBlock, 0, s>>>(
        make_geo<T>(P), (const T*)pbuf, P->d.ld, sort_ref(P), (T*)nbuf, P->d.nld, mode);
    return launch_check("p2g_kin");
  }
  static int force(dmpm_plan* P, const void* pbuf, void* nbuf, int mode, const double* grav,
                   hipStream_t s) {
    ProfScope ps_(P, KC_P2G_FORCE, s);
    T gx = grav ? T(grav[0]) : T(0);
    T gy = (grav && DIM == 2) ? T(grav[1]) : T(0);
    k_p2g_force<T, DIM><<<nblocks(P->nnodes), kBlock, 0, s>>>(
        make_geo<T>(P), (const T*)pbuf, P->d.ld, sort_ref(P), (T*)nbuf, P->d.nld, mode, gx,
        gy);
    return launch_check("p2g_force");
  }
  static int constraints(dmpm_plan* P, void* nbuf, int update_mode, hipStream_t s) {
    if (P->ncon > 0)
      k_constraints<T><<<nblocks(P->ncon), kBlock, 0, s>>>(
          (T*)nbuf, P->d.nld, P->ncon, P->con_node, P->con_dir, P->con_val, update_mode);
    return launch_check("constraints");
  }
  static int nodal_forces(dmpm_plan* P, void* nbuf, T t, hipStream_t s) {
    if (P->nnf > 0)
      k_nodal_forces<T><<<nblocks(P->nnf), kBlock, 0, s>>>((T*)nbuf, P->d.nld, P->nnf,
                                                           P->nf_id, P->nf_dir, P->nf_off,
                                                           P->nf_fn, P->nf_val, fn_table(P), t);
    return launch_check("nodal_forces");
  }
  static int tractions(dmpm_plan* P, void* pbuf, T t, hipStream_t s) {
    if (P->n > 0)
      k_traction_zero<T><<<nblocks(P->n), kBlock, 0, s>>>((T*)pbuf, P->d.ld, (int)P->n);
    if (P->ntr > 0)
      k_traction_assign<T><<<nblocks(P->ntr), kBlock, 0, s>>>(
          (T*)pbuf, P->d.ld, P->ntr, P->tr_id, P->tr_dir, P->tr_off, P->tr_fn, P->tr_val,
          fn_table(P), t);
    return launch_check("tractions");
  }
  static int grid_update(dmpm_plan* P, void* nbuf, double dt, hipStream_t s) {
    ProfScope ps_(P, KC_GRID, s);
    k_grid_update<T, DIM><<<nblocks(P->nnodes), kBlock, 0, s>>>((T*)nbuf, P->d.nld, P->nnodes,
                                                                T(dt));
    DMPM_TRY(launch_check("grid_update"));
    return constraints(P, nbuf, 1, s);
  }
  static int stress(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf, double dt,
                    int mode, hipStream_t s) {
    ProfScope ps_(P, KC_STRESS, s);
    if (P->n > 0)
      k_stress<T, DIM><<<nblocks(P->n), kBlock, 0, s>>>(make_geo<T>(P), make_mat(P), (T*)pbuf,
                                                        P->d.ld, (int)P->n, eid,
                                                        (const T*)nbuf, P->d.nld, T(dt), mode);
    return launch_check("stress");
  }
  static int g2p(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf, double dt,
                 int vu, int write_mom, hipStream_t s) {
    ProfScope ps_(P, KC_G2P, s);
    if (P->n > 0)
      k_g2p<T, DIM><<<nblocks(P->n), kBlock, 0, s>>>(make_geo<T>(P), (T*)pbuf, P->d.ld,
                                                     (int)P->n, eid, (const T*)nbuf, P->d.nld,
                                                     T(dt), vu, write_mom);
    return launch_check("g2p");
  }
  static T step_time(int64_t step, double dt) { return T(step) * T(dt); }
};

// Entry points whose body declares kBatch = true also accept batch plans.
template <class F, class = void> struct BatchOK : std::false_type {};
template <class F>
struct BatchOK<F, std::void_t<decltype(F::kBatch)>> : std::integral_constant<bool, F::kBatch> {};

template <template <typename, int> class F, typename... A>
int dispatch(dmpm_plan* P, A... a) {
  if (P->nsims > 1 && !BatchOK<F<float, 2>>::value)
    return fail(DMPM_EINVAL, "batch plans run the fused 2D step (dmpm_run) and its adjoint only");
  if (P->d.dtype == DMPM_F32)
    return P->dim == 1 ? F<float, 1>::run(P, a...) : F<float, 2>::run(P, a...);
  return P->dim == 1 ? F<double, 1>::run(P, a...) : F<double, 2>::run(P, a...);
}

// Phase bodies (scheme.py:22-67).
template <typename T, int DIM> struct PhaseKin {
  static int run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, hipStream_t s) {
    using O = Ops<T, DIM>;
    DMPM_TRY(O::locate(P, pbuf, eid, 1, s));
    DMPM_TRY(enqueue_sort(P, s));
    DMPM_TRY(O::kin(P, pbuf, nbuf, K_MASS | K_MOM | K_VEL, s));
    return O::constraints(P, nbuf, 0, s);
  }
};
template <typename T, int DIM> struct PhaseStress {
  static int run(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf, double dt,
                 int temps, hipStream_t s) {
    return Ops<T, DIM>::stress(P, pbuf, eid, nbuf, dt,
                               S_STRAIN | S_VOLUME | S_STRESS | (temps ? S_TEMPS : 0), s);
  }
};
// sorted: the caller sorted by these element ids in this step's kinematics
// phase (no element id changes in between), so no re-check is enqueued.
template <typename T, int DIM> struct PhaseForces {
  static int run(dmpm_plan* P, void* pbuf, const int32_t* eid, void* nbuf, const double* grav,
                 int64_t step, double dt, int pfext, hipStream_t s, bool sorted = false) {
    using O = Ops<T, DIM>;
    T t = O::step_time(step, dt);
    if (!sorted) DMPM_TRY(ensure_sorted(P, eid, s));
    if (P->has_tractions) DMPM_TRY(O::tractions(P, pbuf, t, s));
    int mode = F_EXT_ZERO | F_BODY | F_INT | (pfext ? F_PFEXT : 0) |
               (P->has_tractions ? F_TRAC : 0);
    DMPM_TRY(O::force(P, pbuf, nbuf, mode, grav, s));
    return O::nodal_forces(P, nbuf, t, s);
  }
};
template <typename T, int DIM> struct PhasePkin {
  static int run(dmpm_plan* P, void* pbuf, const int32_t* eid, void* nbuf, double dt, int vu,
                 int temps, hipStream_t s) {
    using O = Ops<T, DIM>;
    DMPM_TRY(O::grid_update(P, nbuf, dt, s));
    return O::g2p(P, pbuf, eid, nbuf, dt, vu, temps, s);
  }
};

}  // namespace dmpm

// ----------------------------------------------------------------------------
// C ABI: per-method entry points
// ----------------------------------------------------------------------------
namespace {
template <typename T, int DIM> struct R_locate {
  static int run(dmpm_plan* P, void* pb, int32_t* eid, hipStream_t s) {
    return Ops<T, DIM>::locate(P, pb, eid, 0, s);
  }
};
template <typename T, int DIM> struct R_natural {
  static int run(dmpm_plan* P, void* pb, const int32_t* eid, hipStream_t s) {
    return Ops<T, DIM>::natural(P, pb, eid, s);
  }
};
template <typename T, int DIM> struct R_pvol {
  static int run(dmpm_plan* P, void* pb, const int32_t* eid, double ve, hipStream_t s) {
    return Ops<T, DIM>::pvolume(P, pb, eid, ve, s);
  }
};
template <typename T, int DIM> struct R_kin {
  static int run(dmpm_plan* P, const void* pb, const int32_t* eid, void* nb, int mode,
                 hipStream_t s) {
    if (mode & (K_MASS | K_MOM)) DMPM_TRY(ensure_sorted(P, eid, s));
    return Ops<T, DIM>::kin(P, pb, nb, mode, s);
  }
};
template <typename T, int DIM> struct R_force {
  static int run(dmpm_plan* P, const void* pb, const int32_t* eid, void* nb, int mode,
                 const double* grav, hipStream_t s) {
    DMPM_TRY(ensure_sorted(P, eid, s));
    return Ops<T, DIM>::force(P, pb, nb, mode, grav, s);
  }
};
template <typename T, int DIM> struct R_cons {
  static int run(dmpm_plan* P, void* nb, hipStream_t s) {
    return Ops<T, DIM>::constraints(P, nb, 0, s);
  }
};
template <typename T, int DIM> struct R_nodal {
  static int run(dmpm_plan* P, void* nb, double t, hipStream_t s) {
    return Ops<T, DIM>::nodal_forces(P, nb, T(t), s);
  }
};
template <typename T, int DIM> struct R_trac {
  static int run(dmpm_plan* P, void* pb, double t, hipStream_t s) {
    return Ops<T, DIM>::tractions(P, pb, T(t), s);
  }
};
template <typename T, int DIM> struct R_grid {
  static int run(dmpm_plan* P, void* nb, double dt, hipStream_t s) {
    return Ops<T, DIM>::grid_update(P, nb, dt, s);
  }
};
template <typename T, int DIM> struct R_g2p {
  static int run(dmpm_plan* P, void* pb, const int32_t* eid, const void* nb, double dt, int vu,
                 hipStream_t s) {
    return Ops<T, DIM>::g2p(P, pb, eid, nb, dt, vu, 1, s);
  }
};
template <typename T, int DIM> struct R_stress {
  static int run(dmpm_plan* P, void* pb, const int32_t* eid, const void* nb, double dt, int mode,
                 hipStream_t s) {
    return Ops<T, DIM>::stress(P, pb, eid, nb, dt, mode, s);
  }
};

int check_plan(dmpm_plan* P) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  return DMPM_OK;
}
}  // namespace

#define ST(s) ((hipStream_t)(s))
#define PCHK(P)                              \
  do {                                       \
    int rc__ = check_plan(P);                \
    if (rc__ != DMPM_OK) return rc__;        \
  } while (0)

extern "C" {

int dmpm_set_particle_element_ids(dmpm_plan* P, void* pbuf, int32_t* eid, void* stream) {
  PCHK(P);
  return dispatch<R_locate>(P, pbuf, eid, ST(stream));
}

int dmpm_update_natural_coords(dmpm_plan* P, void* pbuf, const int32_t* eid, void* stream) {
  PCHK(P);
  return dispatch<R_natural>(P, pbuf, eid, ST(stream));
}

int dmpm_element_volume(int32_t element, int32_t dtype, double hx, double hy, double* out) {
  if (!out) return fail(DMPM_EINVAL, "null output");
  if (element != DMPM_LINEAR1D && element != DMPM_QUAD4)
    return fail(DMPM_EINVAL, "element must be DMPM_LINEAR1D or DMPM_QUAD4");
  if (dtype != DMPM_F32 && dtype != DMPM_F64) return fail(DMPM_EINVAL, "bad dtype");
  if (element == DMPM_LINEAR1D) {
    // Intended value X_1 - X_0 = fl(1*h) - 0 (element.py:545-546 crashes for
    // more than one element).
    *out = dtype == DMPM_F32 ? (double)(float)hx : hx;
    return DMPM_OK;
  }
  // Bretschneider's formula of element.py:903-909 in the working precision.
  if (dtype == DMPM_F32) {
    volatile float a = (float)hy, b = (float)hx;
    float c = a, dd = b;
    float p = sqrtf(a * a + b * b), q = p;
    float t = a * a + c * c - b * b - dd * dd;
    *out = (double)(0.25f * sqrtf(4.0f * p * p * q * q - t * t));
  } else {
    volatile double a = hy, b = hx;
    double c = a, dd = b;
    double p = sqrt(a * a + b * b), q = p;
    double t = a * a + c * c - b * b - dd * dd;
    *out = 0.25 * sqrt(4.0 * p * p * q * q - t * t);
  }
  return DMPM_OK;
}

int dmpm_compute_particle_volume(dmpm_plan* P, void* pbuf, const int32_t* eid,
                                 double element_volume, void* stream) {
  PCHK(P);
  return dispatch<R_pvol>(P, pbuf, eid, element_volume, ST(stream));
}

int dmpm_compute_nodal_mass(dmpm_plan* P, const void* pbuf, const int32_t* eid, void* nbuf,
                            void* stream) {
  PCHK(P);
  return dispatch<R_kin>(P, pbuf, eid, nbuf, (int)K_MASS, ST(stream));
}

int dmpm_compute_nodal_momentum(dmpm_plan* P, const void* pbuf, const int32_t* eid, void* nbuf,
                                void* stream) {
  PCHK(P);
  return dispatch<R_kin>(P, pbuf, eid, nbuf, (int)K_MOM, ST(stream));
}

int dmpm_compute_velocity(dmpm_plan* P, void* nbuf, void* stream) {
  PCHK(P);
  return dispatch<R_kin>(P, (const void*)nullptr, (const int32_t*)nullptr, nbuf, (int)K_VEL,
                         ST(stream));
}

int dmpm_apply_boundary_constraints(dmpm_plan* P, void* nbuf, void* stream) {
  PCHK(P);
  return dispatch<R_cons>(P, nbuf, ST(stream));
}

int dmpm_compute_external_force(dmpm_plan* P, const void* pbuf, const int32_t* eid, void* nbuf,
                                void* stream) {
  PCHK(P);
  return dispatch<R_force>(P, pbuf, eid, nbuf, (int)(F_EXT_ZERO | F_PFEXT),
                           (const double*)nullptr, ST(stream));
}

int dmpm_compute_body_force(dmpm_plan* P, const void* pbuf, const int32_t* eid, void* nbuf,
                            const double* gravity, void* stream) {
  PCHK(P);
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  return dispatch<R_force>(P, pbuf, eid, nbuf, (int)F_BODY, gravity, ST(stream));
}

int dmpm_assign_tractions(dmpm_plan* P, void* pbuf, double t, void* stream) {
  PCHK(P);
  return dispatch<R_trac>(P, pbuf, t, ST(stream));
}

int dmpm_apply_particle_traction_forces(dmpm_plan* P, const void* pbuf, const int32_t* eid,
                                        void* nbuf, void* stream) {
  PCHK(P);
  return dispatch<R_force>(P, pbuf, eid, nbuf, (int)F_TRAC, (const double*)nullptr, ST(stream));
}

int dmpm_apply_concentrated_nodal_forces(dmpm_plan* P, void* nbuf, double t, void* stream) {
  PCHK(P);
  return dispatch<R_nodal>(P, nbuf, t, ST(stream));
}

int dmpm_compute_internal_force(dmpm_plan* P, const void* pbuf, const int32_t* eid, void* nbuf,
                                void* stream) {
  PCHK(P);
  return dispatch<R_force>(P, pbuf, eid, nbuf, (int)F_INT, (const double*)nullptr, ST(stream));
}

int dmpm_update_nodal_acceleration_velocity(dmpm_plan* P, void* nbuf, double dt, void* stream) {
  PCHK(P);
  return dispatch<R_grid>(P, nbuf, dt, ST(stream));
}

int dmpm_update_position_velocity(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf,
                                  double dt, int32_t velocity_update, void* stream) {
  PCHK(P);
  return dispatch<R_g2p>(P, pbuf, eid, nbuf, dt, (int)velocity_update, ST(stream));
}

int dmpm_compute_strain(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf,
                        double dt, void* stream) {
  PCHK(P);
  return dispatch<R_stress>(P, pbuf, eid, nbuf, dt, (int)(S_STRAIN | S_TEMPS), ST(stream));
}

int dmpm_update_volume(dmpm_plan* P, void* pbuf, void* stream) {
  PCHK(P);
  return dispatch<R_stress>(P, pbuf, (const int32_t*)nullptr, (const void*)nullptr, 0.0,
                            (int)S_VOLUME, ST(stream));
}

int dmpm_compute_stress(dmpm_plan* P, void* pbuf, void* stream) {
  PCHK(P);
  return dispatch<R_stress>(P, pbuf, (const int32_t*)nullptr, (const void*)nullptr, 0.0,
                            (int)S_STRESS, ST(stream));
}

int dmpm_phase_nodal_kinematics(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf,
                                void* stream) {
  PCHK(P);
  return dispatch<PhaseKin>(P, pbuf, eid, nbuf, ST(stream));
}

int dmpm_phase_stress_strain(dmpm_plan* P, void* pbuf, const int32_t* eid, const void* nbuf,
                             double dt, void* stream) {
  PCHK(P);
  return dispatch<PhaseStress>(P, pbuf, eid, nbuf, dt, 1, ST(stream));
}

int dmpm_phase_forces(dmpm_plan* P, void* pbuf, const int32_t* eid, void* nbuf,
                      const double* gravity, int64_t step, double dt, void* stream) {
  PCHK(P);
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  return dispatch<PhaseForces>(P, pbuf, eid, nbuf, gravity, step, dt, 1, ST(stream));
}

int dmpm_phase_particle_kinematics(dmpm_plan* P, void* pbuf, const int32_t* eid, void* nbuf,
                                   double dt, int32_t velocity_update, void* stream) {
  PCHK(P);
  return dispatch<PhasePkin>(P, pbuf, eid, nbuf, dt, (int)velocity_update, 1, ST(stream));
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Multi-step driver (solver.py:179-213)
// ----------------------------------------------------------------------------
namespace dmpm {

void* plan_workspace(dmpm_plan* P, int slot, size_t bytes) {
  if (slot < 0 || slot >= 16) return nullptr;
  if (P->extra_bytes[slot] >= bytes && P->extra[slot]) return P->extra[slot];
  if (P->extra[slot]) (void)hipFree(P->extra[slot]);
  P->extra[slot] = nullptr;
  P->extra_bytes[slot] = 0;
  if (hipMalloc(&P->extra[slot], bytes) != hipSuccess) return nullptr;
  P->extra_bytes[slot] = bytes;
  return P->extra[slot];
}

template <typename T, int DIM> struct RunSteps {
  static constexpr bool kBatch = true;
  static int run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
                 const double* grav, double dt, int64_t step0, int64_t nsteps, int flags,
                 hipStream_t s) {
    const int temps = flags & 1;
    const int pfext = (flags >> 1) & 1;
    if (!(flags & 4) && tile_eligible(P, pfext))
      return tile_run(P, pbuf, eid, nbuf, scheme, vu, grav, dt, step0, nsteps, temps, s,
                      (flags & DMPM_RUN_CLOCK) ? P->clock : nullptr);
    if (flags & DMPM_RUN_CLOCK)
      return fail(DMPM_EINVAL,
                  "DMPM_RUN_CLOCK covers the fused 2D step (Quadrilateral4Node, no particle "
                  "tractions or particle external forces, flag bit 2 clear)");
    if (P->nsims > 1)
      return fail(DMPM_EINVAL,
                  "batch plans run the fused 2D step only (no per-phase path, particle "
                  "tractions or particle external forces)");
    for (int64_t i = step0; i < step0 + nsteps; ++i) {
      DMPM_TRY((PhaseKin<T, DIM>::run(P, pbuf, eid, nbuf, s)));
      if (scheme == DMPM_USF) DMPM_TRY((PhaseStress<T, DIM>::run(P, pbuf, eid, nbuf, dt, temps, s)));
      DMPM_TRY((PhaseForces<T, DIM>::run(P, pbuf, eid, nbuf, grav, i, dt, pfext, s, true)));
      DMPM_TRY((PhasePkin<T, DIM>::run(P, pbuf, eid, nbuf, dt, vu, temps, s)));
      if (scheme == DMPM_USL) DMPM_TRY((PhaseStress<T, DIM>::run(P, pbuf, eid, nbuf, dt, temps, s)));
    }
    return DMPM_OK;
  }
};

// One step of the per-phase path (scheme.py:22-67, as dmpm_run takes it for
// meshes the fused step does not cover), copying the kinematic nodal velocity
// (after the constraints, before the forces) to vkin, DIM rows of nnodes.
template <typename T, int DIM> struct StepKeepVkin {
  static int run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
                 const double* grav, double dt, int64_t step, void* vkin, hipStream_t s) {
    DMPM_TRY((PhaseKin<T, DIM>::run(P, pbuf, eid, nbuf, s)));
    DMPM_TRY(hip_check(hipMemcpy2DAsync(vkin, (size_t)P->nnodes * sizeof(T),
                                        (T*)nbuf + (int64_t)DMPM_N_V * P->d.nld,
                                        (size_t)P->d.nld * sizeof(T),
                                        (size_t)P->nnodes * sizeof(T), DIM,
                                        hipMemcpyDeviceToDevice, s),
                       "hipMemcpy2DAsync"));
    if (scheme == DMPM_USF) DMPM_TRY((PhaseStress<T, DIM>::run(P, pbuf, eid, nbuf, dt, 0, s)));
    DMPM_TRY((PhaseForces<T, DIM>::run(P, pbuf, eid, nbuf, grav, step, dt, 0, s, true)));
    DMPM_TRY((PhasePkin<T, DIM>::run(P, pbuf, eid, nbuf, dt, vu, 0, s)));
    if (scheme == DMPM_USL) DMPM_TRY((PhaseStress<T, DIM>::run(P, pbuf, eid, nbuf, dt, 0, s)));
    return DMPM_OK;
  }
};

int phase_step_keep_vkin(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
                         const double* grav, double dt, int64_t step, void* vkin,
                         hipStream_t s) {
  return dispatch<StepKeepVkin>(P, pbuf, eid, nbuf, scheme, vu, grav, dt, step, vkin, s);
}

}  // namespace dmpm

extern "C" int dmpm_step_between(dmpm_plan* P, const void* pbuf_in, void* pbuf_out,
                                 int32_t* eid, void* nbuf, int32_t scheme,
                                 int32_t velocity_update, const double* gravity, double dt,
                                 int64_t step, int32_t flags, const void* state_in,
                                 void* state_out, int64_t state_ld, void* stream) {
  PCHK(P);
  if (!pbuf_in || !pbuf_out || !eid || !nbuf) return fail(DMPM_EINVAL, "null buffer");
  if (state_in && !state_out) return fail(DMPM_EINVAL, "state_in needs state_out");
  if (state_in && state_in == state_out) return fail(DMPM_EINVAL, "state_in and state_out must differ");
  if (state_out && state_ld < P->n) return fail(DMPM_EINVAL, "state_ld below the particle count");
  if (pbuf_in == pbuf_out) return fail(DMPM_EINVAL, "input and output blocks must differ");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "dmpm_step_between covers the fused 2D step (Quadrilateral4Node, no particle "
                "tractions or particle external forces)");
  return tile_between(P, pbuf_in, pbuf_out, eid, nbuf, (int)scheme, (int)velocity_update,
                      gravity, dt, step, (int)(flags & 1), (int)((flags >> 1) & 1), state_in,
                      state_out, state_ld, ST(stream));
}

extern "C" int dmpm_step_between_saved(dmpm_plan* P, const void* pbuf_in, void* pbuf_out,
                                       int32_t* eid, const void* nprev, void* nout,
                                       int32_t scheme, int32_t velocity_update,
                                       const double* gravity, double dt, int64_t step,
                                       int32_t flags, const void* state_in, void* state_out,
                                       int64_t state_ld, void* vkin_out, int32_t* perm_out,
                                       int32_t* cs_out, void* stream) {
  PCHK(P);
  if (!pbuf_in || !pbuf_out || !eid || !nprev || !nout) return fail(DMPM_EINVAL, "null buffer");
  if (!vkin_out || !perm_out || !cs_out) return fail(DMPM_EINVAL, "null export buffer");
  if (state_in && !state_out) return fail(DMPM_EINVAL, "state_in needs state_out");
  if (state_in && state_in == state_out) return fail(DMPM_EINVAL, "state_in and state_out must differ");
  if (state_out && state_ld < P->n) return fail(DMPM_EINVAL, "state_ld below the particle count");
  if (pbuf_in == pbuf_out) return fail(DMPM_EINVAL, "input and output blocks must differ");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  if (flags & 1) return fail(DMPM_EINVAL, "a saved step keeps no temporaries");
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "dmpm_step_between_saved covers the fused 2D step (Quadrilateral4Node, no "
                "particle tractions or particle external forces)");
  return tile_between_saved(P, pbuf_in, pbuf_out, eid, nprev, nout, (int)scheme,
                            (int)velocity_update, gravity, dt, step, (int)((flags >> 1) & 1),
                            state_in, state_out, state_ld, vkin_out, perm_out, cs_out,
                            ST(stream));
}

extern "C" int dmpm_step_between_noalloc(dmpm_plan* P, const void* pbuf_in, void* pbuf_out,
                                         int32_t* eid, void* nbuf, int32_t scheme,
                                         int32_t velocity_update, const double* gravity,
                                         double dt, int64_t step, int32_t flags,
                                         const void* state_in, void* state_out,
                                         int64_t state_ld, void* stream) {
  PCHK(P);
  if (!pbuf_in || !pbuf_out || !eid || !nbuf) return fail(DMPM_EINVAL, "null buffer");
  if (state_in && !state_out) return fail(DMPM_EINVAL, "state_in needs state_out");
  if (state_in && state_in == state_out) return fail(DMPM_EINVAL, "state_in and state_out must differ");
  if (state_out && state_ld < P->n) return fail(DMPM_EINVAL, "state_ld below the particle count");
  if (pbuf_in == pbuf_out) return fail(DMPM_EINVAL, "input and output blocks must differ");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "dmpm_step_between_noalloc covers the fused 2D step (Quadrilateral4Node, no "
                "particle tractions or particle external forces)");
  return tile_between_noalloc(P, pbuf_in, pbuf_out, eid, nbuf, (int)scheme, (int)velocity_update,
                              gravity, dt, step, (int)(flags & 1), (int)((flags >> 1) & 1),
                              state_in, state_out, state_ld, ST(stream));
}

namespace dmpm {
__global__ void k_clock_set(int64_t* clock, int64_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clock = v;
}
}  // namespace dmpm

extern "C" int dmpm_plan_set_clock(dmpm_plan* P, int64_t step, void* stream) {
  PCHK(P);
  dmpm::k_clock_set<<<1, 64, 0, ST(stream)>>>(P->clock, step);
  return launch_check("clock_set");
}

extern "C" int dmpm_plan_get_clock(dmpm_plan* P, int64_t* step, void* stream) {
  PCHK(P);
  if (!step) return fail(DMPM_EINVAL, "null output");
  DMPM_TRY(hip_check(hipMemcpyAsync(step, P->clock, sizeof(int64_t), hipMemcpyDeviceToHost,
                                    ST(stream)),
                     "hipMemcpyAsync"));
  return hip_check(hipStreamSynchronize(ST(stream)), "hipStreamSynchronize");
}

extern "C" int dmpm_prepare_run(dmpm_plan* P, int32_t scheme, int32_t flags, void* stream) {
  PCHK(P);
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  // the per-phase kernels use only the plan's own buffers
  if ((flags & 4) || !tile_eligible(P, (flags >> 1) & 1)) return DMPM_OK;
  DMPM_TRY(tile_prepare(P, (int)scheme, ST(stream)));
  return hip_check(hipStreamSynchronize(ST(stream)), "hipStreamSynchronize");
}

extern "C" int dmpm_run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int32_t scheme,
                        int32_t velocity_update, const double* gravity, double dt, int64_t step0,
                        int64_t nsteps, int32_t flags, void* stream) {
  PCHK(P);
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!gravity) return fail(DMPM_EINVAL, "null gravity");
  if (nsteps < 0) return fail(DMPM_EINVAL, "negative step count");
  return dispatch<RunSteps>(P, pbuf, eid, nbuf, (int)scheme, (int)velocity_update, gravity, dt,
                            step0, nsteps, (int)flags, ST(stream));
}
