// Reverse-mode (adjoint) of one Linear1D explicit step (gfx950).
//
// The reference differentiates its 1D bar (examples/optim_1d.py: Linear1D,
// SimpleMaterial, USL) with jax.grad through solve_jit; this is the adjoint of
// the 1D step as this library computes it (the per-phase kernels of
// mpm_forward.hip, scheme.py:22-67), for the particle state
//   S = (x, y, vx, vy, sigma_xx, sigma_yy, sigma_xy, V)
// (y and vy are not 1D quantities and pass their cotangents through) and the
// per-particle log Young's modulus, with the particle mass, the mesh, the
// constraints and the nodal forces held fixed.
//
// In 1D the strain rate is one number per particle, L = dN0 v0 + dN1 v1
// (dN = -+1/h, independent of xi), and the reference writes it into the rate
// vector with its out-of-range indices clamped: rate = [L, L, 0, 2L, 0, 0]
// (particle.py:281-313), each entry snapped; the centroid rate is the same L,
// so dvol = dt snap(L).  The stress increment is E rate dt (SimpleMaterial,
// material.py:130-131) or De(E) rate dt (LinearElastic): sigma_yy and
// sigma_xy change too, and their cotangents reach L.
//
// dmpm_step_adjoint on a Linear1D plan recomputes the forward step from its
// input state with the per-phase kernels (keeping the kinematic nodal
// velocity), then runs, for cotangents Sbar' of the step output:
//   k_adj1_out     G2P adjoint (per-particle weights of the nodal velocity
//                  and acceleration cotangents, xi cotangent); USL: stress
//                  and volume adjoint (weight of the strain-rate cotangent)
//   k_adj1_node<1> gather of those weights at each node over the particles of
//                  its two cells in sorted order (deterministic), nodal
//                  update adjoint (a = nan_to_num(F/m), v' = v + a dt, snaps,
//                  constraints); USL: kinematics adjoint (v = snap(P)/m)
//   USF: k_adj1_usf_stress  internal/body force adjoint on the updated
//                  V, sigma, then stress and volume adjoint on the kinematic
//                  velocity; k_adj1_node<2> its strain-rate scatter plus the
//                  kinematics adjoint
//   k_adj1_p2g     USL: internal/body force adjoint; P2G of mass and momentum
//                  adjoint (gather of the nodal cotangents); chain to x
//                  through xi = (x - (X0 + X1)/2) (2/(X1 - X0))
// Element ids are piecewise constant in x: no cotangent flows through them.
// Particles outside the mesh get zero cotangents.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mpm_common.h"
#include "mpm_plan.h"

using namespace dmpm;

namespace {

constexpr int kB = 256;

// per-particle weight rows: nodal-velocity and acceleration cotangents of
// G2P, strain-rate cotangent, natural-coordinate cotangent
enum { W_NV = 0, W_NA, W_L, W_XI, W_ROWS };
// per-node cotangent rows: total force, kinematic velocity (USF), snapped
// momentum, mass
enum { A_F = 0, A_VK, A_P, A_M, A_ROWS };
// rows of the cotangent block
enum { C_X = 0, C_Y, C_VX, C_VY, C_SXX, C_SYY, C_SXY, C_V, C_ROWS };

struct Mat1 {
  int kind;
  int per_E;
  double de[3];  // (a1, a2, G) or (E, -, -); per unit E with a per-particle E
};

template <typename T> struct Adj1 {
  int nx, n, nnodes;
  T hx;
  const T* pb;  // step input block (x, v, mass, volume, stress, E)
  const T* pw;  // recomputed block: xi; after a USF step its V and stress
  int64_t ld;
  const int* eid;
  const T* nb;  // node block after the step
  int64_t nld;
  const T* vkin;  // kinematic nodal velocity
  T* cot;
  T* lebar;
  T* wk;
  T* na;
  const int* cidx;
  const int* cmask;
  Mat1 mat;
  T dt, gx;
  int vu;
  __device__ __forceinline__ T X(int i) const { return T(i) * hx; }
  __device__ __forceinline__ const T* P(int r) const { return pb + (int64_t)r * ld; }
  __device__ __forceinline__ const T* W(int r) const { return pw + (int64_t)r * ld; }
  __device__ __forceinline__ const T* N(int r) const { return nb + (int64_t)r * nld; }
  __device__ __forceinline__ T* C(int r) const { return cot + (int64_t)r * ld; }
  __device__ __forceinline__ T* K(int r) const { return wk + (int64_t)r * ld; }
  __device__ __forceinline__ T* A(int r) const { return na + (int64_t)r * nnodes; }
};

// Stress increment of the 1D rate vector r = [r0, r1, 0, r3, 0, 0] (the
// entries the 2D state holds), as mpm_forward.hip's stress_increment.
template <typename T>
__device__ __forceinline__ void inc1(const Mat1& m, T E_p, const T d[4], T ds[4]) {
  if (m.kind == DMPM_SIMPLE_MATERIAL) {  // (a per-particle E comes as LinearElastic)
    const T E = T(m.de[0]);
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

// D^T applied to the stress cotangent [sxx, syy, -, sxy] (D is symmetric).
template <typename T>
__device__ __forceinline__ void dT1(const Mat1& m, T E_p, const T sb[4], T out[4]) {
  if (m.kind == DMPM_SIMPLE_MATERIAL) {  // (a per-particle E comes as LinearElastic)
    const T E = T(m.de[0]);
    for (int k = 0; k < 4; ++k) out[k] = sb[k] * E;
    return;
  }
  T a1 = T(m.de[0]), a2 = T(m.de[1]), G = T(m.de[2]);
  if (m.per_E) {
    a1 = a1 * E_p;
    a2 = a2 * E_p;
    G = G * E_p;
  }
  out[0] = a1 * sb[0] + a2 * sb[1];
  out[1] = a2 * sb[0] + a1 * sb[1];
  out[2] = T(0);
  out[3] = G * sb[3];
}

// Adjoint of strain/volume/stress (particle.py:249-327) at one particle of
// cell e with nodal velocities v0, v1: given the cotangents sb (sigma' xx, yy,
// xy) and Vb (V'), adds sb . dsigma to *lebar (dsigma is linear in E, so
// that is dL/dlog E), returns the strain-rate cotangent and the cotangent of
// the step-start volume (sigma' = sigma + dsigma passes sb unchanged).
template <typename T>
__device__ __forceinline__ T stress_adj1(const Mat1& mat, T E_p, T dt, T inv, T v0, T v1, T V_in,
                                         const T sb[3], T Vb, T* lebar_p, T& Vb_in) {
  const T L = l1_dNdx<T>(0, inv) * v0 + l1_dNdx<T>(1, inv) * v1;
  const T r0 = snap(L), r3 = snap(L + L);
  const T d[4] = {r0 * dt, r0 * dt, T(0), r3 * dt};
  const T dvol = dt * r0;
  T ds[4];
  inc1(mat, E_p, d, ds);
  if (lebar_p) *lebar_p = *lebar_p + ((sb[0] * ds[0] + sb[1] * ds[1]) + sb[2] * ds[3]);
  const T sg[4] = {sb[0], sb[1], T(0), sb[2]};
  T de[4];
  dT1(mat, E_p, sg, de);
  T Lb = T(0);
  if (r0 != T(0)) Lb = Lb + dt * ((de[0] + de[1]) + Vb * V_in);
  if (r3 != T(0)) Lb = Lb + T(2) * (dt * de[3]);
  Vb_in = Vb * (T(1) + dvol);
  return Lb;
}

// G2P adjoint (particle.py:208-247); USL: stress/volume adjoint on the
// nodal velocity after the update.
template <typename T, bool USL>
__global__ __launch_bounds__(kB) void k_adj1_out(Adj1<T> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const int e = a.eid[p];
  if (e < 0 || e >= a.nx) {
    for (int r = 0; r < C_ROWS; ++r) a.C(r)[p] = T(0);
    for (int r = 0; r < W_ROWS; ++r) a.K(r)[p] = T(0);
    return;
  }
  const T xi = a.W(DMPM_P_XI)[p];
  const T v0 = a.N(DMPM_N_V)[e], v1 = a.N(DMPM_N_V)[e + 1];
  const T a0 = a.N(DMPM_N_A)[e], a1 = a.N(DMPM_N_A)[e + 1];
  const T xb = a.C(C_X)[p], vxb = a.C(C_VX)[p];
  // x' = x + nv dt; v' = nv (velocity update) or v + na dt
  T nvb = xb * a.dt, nab = T(0);
  if (a.vu) {
    nvb = nvb + vxb;
    a.C(C_VX)[p] = T(0);
  } else {
    nab = vxb * a.dt;
  }
  // d/dxi of sum N_k f_k = (f1 - f0)/2
  T xib = nvb * (T(0.5) * (v1 - v0)) + nab * (T(0.5) * (a1 - a0));
  T Lb = T(0);
  if (USL) {
    const T inv = l1_inv(a.X(e), a.X(e + 1));
    const T sb[3] = {a.C(C_SXX)[p], a.C(C_SYY)[p], a.C(C_SXY)[p]};
    const T E_p = a.mat.per_E ? a.P(DMPM_P_YOUNGS)[p] : T(0);
    T Vb_in;
    Lb = stress_adj1(a.mat, E_p, a.dt, inv, v0, v1, a.P(DMPM_P_VOLUME)[p], sb, a.C(C_V)[p],
                     a.lebar ? a.lebar + p : (T*)nullptr, Vb_in);
    a.C(C_V)[p] = Vb_in;
  }
  (void)xi;
  a.K(W_NV)[p] = nvb;
  a.K(W_NA)[p] = nab;
  a.K(W_L)[p] = Lb;
  a.K(W_XI)[p] = xib;
}

// Kinematics adjoint at one node (element.py:191-202, constraint.py:29-44):
// v = snap(snap(P)/m), or the constraint value.
template <typename T>
__device__ __forceinline__ void kin_adj1(T m, T v, bool fixed, T vb, T& Pb, T& mb) {
  Pb = T(0);
  if (fixed || m == T(0) || v == T(0)) return;
  Pb = vb / m;
  mb = mb - vb * v / m;
}

// Sum over the particles of node nd's two cells (cell nd-1 at its right
// corner, then cell nd at its left one; slots in sorted order) of
// N * w_N + dN * w_D (pass 1: w_N = W_NV and W_NA, w_D = W_L; pass 2: the
// dN * W_L term only).
template <typename T, int PASS>
__device__ __forceinline__ void gather1(const Adj1<T>& a, SortRef sr, int nd, T& sv, T& sa) {
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  sv = T(0);
  sa = T(0);
  for (int c = 0; c < 2; ++c) {
    const int cell = nd - 1 + c, corner = 1 - c;
    if (cell < 0 || cell >= a.nx) continue;
    const T dN = l1_dNdx<T>(corner, l1_inv(a.X(cell), a.X(cell + 1)));
    T cv = T(0), ca = T(0);
    for (int s = cs[cell]; s < cs[cell + 1]; ++s) {
      const int p = perm[s];
      if (PASS == 1) {
        const T N = l1_N<T>(corner, a.W(DMPM_P_XI)[p]);
        cv = cv + (N * a.K(W_NV)[p] + dN * a.K(W_L)[p]);
        ca = ca + N * a.K(W_NA)[p];
      } else {
        cv = cv + dN * a.K(W_L)[p];
      }
    }
    sv = sv + cv;
    sa = sa + ca;
  }
}

// Pass 1: adjoint of the nodal update (element.py:305-329): F = f_int +
// f_ext + f_damp, a0 = nan_to_num(F/m), v' = snap(v_kin + a0 dt), a' =
// snap(a0), constraints; USL: kinematics adjoint.  Pass 2 (USF): the
// kinematic velocity's cotangent from pass 1 plus the strain-rate scatter,
// kinematics adjoint.
template <typename T, int PASS, bool USL>
__global__ __launch_bounds__(kB) void k_adj1_node(Adj1<T> a, SortRef sr) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= a.nnodes) return;
  T sv, sa;
  gather1<T, PASS>(a, sr, nd, sv, sa);
  const int cr = a.cidx ? a.cidx[nd] : -1;
  const bool fixed = cr >= 0 && (a.cmask[cr] & 1);
  const T m = a.N(DMPM_N_MASS)[nd];
  const T vk = a.vkin[nd];
  if (PASS == 2) {
    T Pb, mb = a.A(A_M)[nd];
    kin_adj1(m, vk, fixed, a.A(A_VK)[nd] + sv, Pb, mb);
    a.A(A_P)[nd] = Pb;
    a.A(A_M)[nd] = mb;
    return;
  }
  T Fb = T(0), vkb = T(0), mb = T(0);
  if (!fixed) {
    const T F = (a.N(DMPM_N_FINT)[nd] + a.N(DMPM_N_FEXT)[nd]) + a.N(DMPM_N_FDAMP)[nd];
    const T r = F / m;
    const T v1b = a.N(DMPM_N_V)[nd] != T(0) ? sv : T(0);
    const T a0b = (a.N(DMPM_N_A)[nd] != T(0) ? sa : T(0)) + a.dt * v1b;
    if (isfinite(r)) {
      Fb = a0b / m;
      mb = mb - a0b * r / m;
    }
    vkb = v1b;
  }
  a.A(A_F)[nd] = Fb;
  if (USL) {
    T Pb;
    kin_adj1(m, vk, fixed, vkb, Pb, mb);
    a.A(A_P)[nd] = Pb;
  } else {
    a.A(A_VK)[nd] = vkb;
  }
  a.A(A_M)[nd] = mb;
}

// USF: internal and body force adjoint on the updated volume and stress
// (element.py:548-600, 236-268), then the stress/volume adjoint on the
// kinematic velocity.
template <typename T>
__global__ __launch_bounds__(kB) void k_adj1_usf_stress(Adj1<T> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const int e = a.eid[p];
  if (e < 0 || e >= a.nx) return;
  const T inv = l1_inv(a.X(e), a.X(e + 1));
  const T Fb0 = a.A(A_F)[e], Fb1 = a.A(A_F)[e + 1];
  const T g = l1_dNdx<T>(0, inv) * Fb0 + l1_dNdx<T>(1, inv) * Fb1;
  // f_int_k += (-V' sigma'_xx) dN_k
  const T Vb = a.C(C_V)[p] - a.W(DMPM_P_STRESS)[p] * g;
  const T sb[3] = {a.C(C_SXX)[p] - a.W(DMPM_P_VOLUME)[p] * g, a.C(C_SYY)[p], a.C(C_SXY)[p]};
  const T m = a.P(DMPM_P_MASS)[p];
  a.K(W_XI)[p] = a.K(W_XI)[p] + (T(0.5) * (Fb1 - Fb0)) * (m * a.gx);
  const T E_p = a.mat.per_E ? a.P(DMPM_P_YOUNGS)[p] : T(0);
  T Vb_in;
  a.K(W_L)[p] = stress_adj1(a.mat, E_p, a.dt, inv, a.vkin[e], a.vkin[e + 1],
                            a.P(DMPM_P_VOLUME)[p], sb, Vb, a.lebar ? a.lebar + p : (T*)nullptr,
                            Vb_in);
  a.C(C_V)[p] = Vb_in;
  a.C(C_SXX)[p] = sb[0];
}

// USL: internal and body force adjoint on the step-start volume and stress;
// P2G of mass and momentum adjoint (element.py:122-189); x cotangent through
// the natural coordinate (particle.py:180-206).
template <typename T, bool USL>
__global__ __launch_bounds__(kB) void k_adj1_p2g(Adj1<T> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const int e = a.eid[p];
  if (e < 0 || e >= a.nx) return;
  const T inv = l1_inv(a.X(e), a.X(e + 1));
  const T xi = a.W(DMPM_P_XI)[p];
  const T m = a.P(DMPM_P_MASS)[p], v = a.P(DMPM_P_V)[p];
  T xib = a.K(W_XI)[p];
  if (USL) {
    const T Fb0 = a.A(A_F)[e], Fb1 = a.A(A_F)[e + 1];
    const T g = l1_dNdx<T>(0, inv) * Fb0 + l1_dNdx<T>(1, inv) * Fb1;
    a.C(C_V)[p] = a.C(C_V)[p] - a.P(DMPM_P_STRESS)[p] * g;
    a.C(C_SXX)[p] = a.C(C_SXX)[p] - a.P(DMPM_P_VOLUME)[p] * g;
    xib = xib + (T(0.5) * (Fb1 - Fb0)) * (m * a.gx);
  }
  const T Pb0 = a.A(A_P)[e], Pb1 = a.A(A_P)[e + 1];
  const T mb0 = a.A(A_M)[e], mb1 = a.A(A_M)[e + 1];
  // M_k += N_k m, P_k += N_k (m v)
  xib = xib + m * (T(0.5) * (mb1 - mb0)) + (m * v) * (T(0.5) * (Pb1 - Pb0));
  a.C(C_VX)[p] = a.C(C_VX)[p] + m * (l1_N<T>(0, xi) * Pb0 + l1_N<T>(1, xi) * Pb1);
  a.C(C_X)[p] = a.C(C_X)[p] + xib * inv;
}

unsigned nblk1(int n) { return (unsigned)std::max(1, (n + kB - 1) / kB); }

template <typename T>
int adjoint_1d_t(dmpm_plan* P, const T* pb, T* pw, int32_t* eid, T* nb, T* cot, T* lebar,
                 int scheme, int vu, const double* grav, double dt, int64_t step, hipStream_t s) {
  const int64_t ld = P->d.ld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  // the rows the forward step reads: x, v, mass, volume, density, stress, E
  const int spans[3][2] = {{DMPM_P_X, DMPM_P_DENSITY + 1},
                           {DMPM_P_STRESS, DMPM_P_STRESS + 6},
                           {DMPM_P_YOUNGS, DMPM_P_YOUNGS + 1}};
  for (const auto& sp : spans)
    DMPM_TRY(hip_check(hipMemcpyAsync(pw + (int64_t)sp[0] * ld, pb + (int64_t)sp[0] * ld,
                                      (size_t)(sp[1] - sp[0]) * ld * sizeof(T),
                                      hipMemcpyDeviceToDevice, s),
                       "hipMemcpyAsync"));
  T* vkin = (T*)plan_workspace(P, 1, (size_t)P->nnodes * sizeof(T));
  T* wk = (T*)plan_workspace(P, 3, (size_t)W_ROWS * ld * sizeof(T));
  T* na = (T*)plan_workspace(P, 4, (size_t)A_ROWS * P->nnodes * sizeof(T));
  if (!vkin || !wk || !na) return fail(DMPM_EHIP, "workspace allocation failed");
  DMPM_TRY(phase_step_keep_vkin(P, pw, eid, nb, scheme, vu, grav, dt, step, vkin, s));
  ProfScopeT ps_(P, KC_ADJ, s);
  Adj1<T> a;
  a.nx = P->d.nx;
  a.n = n;
  a.nnodes = P->nnodes;
  a.hx = T(P->d.hx);
  a.pb = pb;
  a.pw = pw;
  a.ld = ld;
  a.eid = eid;
  a.nb = nb;
  a.nld = P->d.nld;
  a.vkin = vkin;
  a.cot = cot;
  a.lebar = lebar;
  a.wk = wk;
  a.na = na;
  a.cidx = P->cidx;
  a.cmask = P->cmask;
  a.mat.kind = P->d.material;
  a.mat.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) a.mat.de[k] = P->d.de[k];
  a.dt = T(dt);
  a.gx = T(grav[0]);
  a.vu = vu;
  const SortRef sr = sort_ref(P);
  const unsigned np = nblk1(n), nn = nblk1(P->nnodes);
  if (usf) {
    k_adj1_out<T, false><<<np, kB, 0, s>>>(a);
    k_adj1_node<T, 1, false><<<nn, kB, 0, s>>>(a, sr);
    k_adj1_usf_stress<T><<<np, kB, 0, s>>>(a);
    k_adj1_node<T, 2, false><<<nn, kB, 0, s>>>(a, sr);
    k_adj1_p2g<T, false><<<np, kB, 0, s>>>(a);
  } else {
    k_adj1_out<T, true><<<np, kB, 0, s>>>(a);
    k_adj1_node<T, 1, true><<<nn, kB, 0, s>>>(a, sr);
    k_adj1_p2g<T, true><<<np, kB, 0, s>>>(a);
  }
  return launch_check("step_adjoint_1d");
}

}  // namespace

namespace dmpm {
int adjoint_1d(dmpm_plan* P, const void* pin, void* pwork, int32_t* eid, void* nwork, void* cot,
               void* lebar, int scheme, int vu, const double* grav, double dt, int64_t step,
               hipStream_t s) {
  if (P->dim != 1) return fail(DMPM_EINVAL, "not a Linear1D plan");
  if (P->nsims > 1) return fail(DMPM_EINVAL, "batch plans are 2D");
  if (P->has_tractions)
    return fail(DMPM_EINVAL, "the 1D adjoint step covers steps without particle tractions");
  if (P->d.dtype == DMPM_F32)
    return adjoint_1d_t<float>(P, (const float*)pin, (float*)pwork, eid, (float*)nwork,
                               (float*)cot, (float*)lebar, scheme, vu, grav, dt, step, s);
  return adjoint_1d_t<double>(P, (const double*)pin, (double*)pwork, eid, (double*)nwork,
                              (double*)cot, (double*)lebar, scheme, vu, grav, dt, step, s);
}
}  // namespace dmpm
