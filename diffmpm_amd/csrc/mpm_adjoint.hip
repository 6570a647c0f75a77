// Reverse-mode (adjoint) of one fused 2D explicit step (gfx950).
//
// The reference differentiates its rollout with jax.grad; there is no
// hand-written adjoint in it.  This is the adjoint of the step as this
// library computes it (mpm_tile.hip), for the dynamic particle state
//   S = (x, y, vx, vy, sigma_xx, sigma_yy, sigma_xy, V)
// and the per-particle log Young's modulus, with the particle mass, the mesh,
// the constraints and the nodal forces held fixed.  Strain, density,
// volumetric strain and sigma_zz do not feed back into the dynamics of a 2D
// LinearElastic step; their cotangents are not propagated.
//
// dmpm_step_adjoint recomputes the forward step from its input state (on a
// scratch block, keeping the nodal arrays and the kinematic nodal velocity),
// then runs, for cotangents Sbar' of the step output:
//   k_adj_g2p     G2P adjoint, particle-local part: cotangents of the nodal
//                 velocity/acceleration as per-particle weights, the
//                 natural-coordinate cotangent of G2P; USL: stress adjoint
//   cell + node 1 scatter of those weights to the nodes (deterministic cell
//                 pass), grid-update adjoint (a = F/m, v' = v + a dt, snaps,
//                 constraints); USL: kinematics adjoint (v = P/m)
//   k_adj_force   internal/body force adjoint (gather of Fbar); USF: stress
//                 adjoint (sigma' = sigma + D deps, V' = V (1 + dvol))
//   cell + node 2 (USF) scatter of the strain-rate cotangents to the kinematic
//                 nodal velocities, kinematics adjoint
//   k_adj_kin     P2G of mass/momentum adjoint (gather of Pbar, mbar), chain
//                 to x through the natural coordinates
// Element ids are piecewise constant in x: no cotangent flows through them.
// snap(x) = where(|x| < 1e-12, 0, x) passes a cotangent where it kept x;
// nan_to_num passes it where the value was finite.  Particles outside the
// mesh get zero cotangents.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mpm_common.h"
#include "mpm_plan.h"
#include "mpm_step.h"

using namespace dmpm;

namespace {

#define PR(r) (pb + (int64_t)(r) * ld)
#define PW(r) (pw + (int64_t)(r) * ld)
#define NR(r) (nb + (int64_t)(r) * nld)
#define CR(r) (cot + (int64_t)(r) * ld)
#define WR(r) (wk + (int64_t)(r) * ld)
#define AR(r) (na + (int64_t)(r) * M.nnodes)

// rows of the per-particle weight workspace: N-weighted nodal-velocity and
// acceleration cotangents, strain-rate cotangents (L00, L11, L01, L10 at the
// particle, L00, L11 at the centroid), natural-coordinate cotangent
enum { W_NVX = 0, W_NVY, W_NAX, W_NAY, W_L00, W_L11, W_L01, W_L10, W_C00, W_C11, W_XI, W_ETA,
       W_ROWS };
// rows of the nodal adjoint workspace
enum { A_FX = 0, A_FY, A_M, A_VKX, A_VKY, A_PX, A_PY, A_ROWS };

// The particle rows the adjoint reads (element p at [p]): the step's input
// (from a particle block, row r at pb + r*ld, or from a state tensor plus a
// block of the constant rows) and the stress and volume after the step.
template <typename T> struct PRows {
  const T* x;
  const T* y;
  const T* vx;
  const T* vy;
  const T* m;
  const T* V;
  const T* sxx;
  const T* syy;
  const T* sxy;
  const T* E;
};
template <typename T> struct POut {
  const T* V;
  const T* sxx;
  const T* syy;
  const T* sxy;
};
template <typename T> PRows<T> rows_of_block(const T* pb, int64_t ld) {
  PRows<T> r;
  r.x = pb + (int64_t)DMPM_P_X * ld;
  r.y = pb + (int64_t)(DMPM_P_X + 1) * ld;
  r.vx = pb + (int64_t)DMPM_P_V * ld;
  r.vy = pb + (int64_t)(DMPM_P_V + 1) * ld;
  r.m = pb + (int64_t)DMPM_P_MASS * ld;
  r.V = pb + (int64_t)DMPM_P_VOLUME * ld;
  r.sxx = pb + (int64_t)DMPM_P_STRESS * ld;
  r.syy = pb + (int64_t)(DMPM_P_STRESS + 1) * ld;
  r.sxy = pb + (int64_t)(DMPM_P_STRESS + 3) * ld;
  r.E = pb + (int64_t)DMPM_P_YOUNGS * ld;
  return r;
}
template <typename T> POut<T> out_of_block(const T* pw, int64_t ld) {
  POut<T> o;
  o.V = pw + (int64_t)DMPM_P_VOLUME * ld;
  o.sxx = pw + (int64_t)DMPM_P_STRESS * ld;
  o.syy = pw + (int64_t)(DMPM_P_STRESS + 1) * ld;
  o.sxy = pw + (int64_t)(DMPM_P_STRESS + 3) * ld;
  return o;
}
// state tensor rows x, y, vx, vy, sxx, syy, sxy, V (row q at st + q*sld)
template <typename T> PRows<T> rows_of_state(const T* st, int64_t sld, const T* pconst,
                                             int64_t ld) {
  PRows<T> r = rows_of_block(pconst, ld);
  r.x = st;
  r.y = st + sld;
  r.vx = st + 2 * sld;
  r.vy = st + 3 * sld;
  r.sxx = st + 4 * sld;
  r.syy = st + 5 * sld;
  r.sxy = st + 6 * sld;
  r.V = st + 7 * sld;
  return r;
}
template <typename T> POut<T> out_of_state(const T* st, int64_t sld) {
  POut<T> o;
  o.sxx = st + 4 * sld;
  o.syy = st + 5 * sld;
  o.sxy = st + 6 * sld;
  o.V = st + 7 * sld;
  return o;
}

// Per-particle geometry at the step-start position: simulation, cell of its
// own mesh, natural coordinates, inverse-Jacobian scales, corner nodes.
template <typename T> struct PGeo {
  bool in;
  int s, I, J;
  T xi, eta, ix, iy;
  int ids[4];
};
template <typename T>
__device__ __forceinline__ PGeo<T> pgeo(const Mesh2<T>& M, int p, T x, T y) {
  PGeo<T> q;
  q.s = M.sim_p(p);
  const Grid2<T> g = M.grid(q.s);
  const int e = element_id<T, 2>(g.geo(), x, y);
  q.in = e >= 0 && e < g.ncells;
  q.I = q.in ? e % g.nx : 0;
  q.J = q.in ? e / g.nx : 0;
  const T X0 = g.X(q.I), X1 = g.X(q.I + 1), Y0 = g.Y(q.J), Y1 = g.Y(q.J + 1);
  q.xi = natural(x, X0, X1);
  q.eta = natural(y, Y0, Y1);
  q4_invJ(q.xi, q.eta, X0, X1, Y0, Y1, q.ix, q.iy);
  cell_nodes(g.nx, q.I, q.J, q.ids, M.noff(q.s));
  return q;
}

// d^2 N_k / (dxi deta) of the bilinear shape functions (element.py:748-755).
__device__ __forceinline__ float mixed2(int k) { return (k & 1) ? -0.25f : 0.25f; }

// Material tangent D (symmetric) applied to a 2D [xx, yy, -, xy] vector.
template <typename T>
__device__ __forceinline__ void d_apply(const Mat2& m, T E_p, const T in[4], T out[4]) {
  if (m.kind == DMPM_SIMPLE_MATERIAL) {
    const T E = T(m.de[0]);
    for (int k = 0; k < 4; ++k) out[k] = in[k] * E;
    return;
  }
  T a1 = T(m.de[0]), a2 = T(m.de[1]), G = T(m.de[2]);
  if (m.per_E) {
    a1 = a1 * E_p;
    a2 = a2 * E_p;
    G = G * E_p;
  }
  out[0] = a1 * in[0] + a2 * in[1];
  out[1] = a2 * in[0] + a1 * in[1];
  out[2] = T(0);
  out[3] = G * in[3];
}

// Adjoint of the strain/volume/stress update (particle.py:249-327) at one
// particle with nodal velocities vx, vy: given the cotangents sb (sigma'
// xx, yy, xy) and Vb (V'), adds sigma' . dsigma to lebar (dL/dlogE: D is
// linear in E), returns the strain-rate cotangents Lb[6] (L00, L11, L01, L10
// at the particle, L00, L11 at the centroid), the natural-coordinate
// cotangent xb[2] and the cotangent of the step-start volume.
template <typename T>
__device__ __forceinline__ void stress_adjoint(const Mat2& mat, T E_p, T dt, T V_in,
                                               const PGeo<T>& q, const T vx[4], const T vy[4],
                                               const T sb[3], T Vb, T* lebar, T Lb[6], T xb[2],
                                               T& Vb_in) {
  T rr[6], rc[6], d[6];
  rate2s(q.ix, q.iy, q.xi, q.eta, vx, vy, rr);
  rate2s(q.ix, q.iy, T(0), T(0), vx, vy, rc);
  for (int k = 0; k < 6; ++k) d[k] = rr[k] * dt;
  const T dvol = dt * (rc[0] + rc[1]);
  T ds[4];
  stress_inc(mat, E_p, d, ds);
  *lebar = *lebar + (sb[0] * ds[0] + sb[1] * ds[1] + sb[2] * ds[3]);
  const T sg[4] = {sb[0], sb[1], T(0), sb[2]};
  T de[4];
  d_apply(mat, E_p, sg, de);
  Lb[0] = rr[0] != T(0) ? dt * de[0] : T(0);
  Lb[1] = rr[1] != T(0) ? dt * de[1] : T(0);
  Lb[2] = rr[3] != T(0) ? dt * de[3] : T(0);
  Lb[3] = Lb[2];
  const T dvb = Vb * V_in;
  Lb[4] = rc[0] != T(0) ? dt * dvb : T(0);
  Lb[5] = rc[1] != T(0) ? dt * dvb : T(0);
  Vb_in = Vb * (T(1) + dvol);
  // L00 = sum dx_k vx_k, L01 = sum dx_k vy_k depend on eta; L11, L10 on xi
  T sxi = T(0), seta = T(0);
  for (int k = 0; k < 4; ++k) {
    const T s = T(mixed2(k));
    sxi = sxi + s * (Lb[1] * vy[k] + Lb[3] * vx[k]);
    seta = seta + s * (Lb[0] * vx[k] + Lb[2] * vy[k]);
  }
  xb[0] = sxi * q.iy;
  xb[1] = seta * q.ix;
}

// G2P adjoint (particle.py:208-247), particle-local part; USL also the
// adjoint of the stress update made with the updated nodal velocities
// (scheme.py:127-133).
// G2P adjoint (particle.py:208-247) of particle p (inside the mesh, geometry
// q), particle-local part: cotangent updates, the natural-coordinate
// cotangent (rows W_XI, W_ETA of wk) and the weights w[W_NVX .. W_C11] the
// scatter to the nodes needs; USL also the adjoint of the stress update made
// with the updated nodal velocities (scheme.py:127-133).
template <typename T, bool USL>
__device__ __forceinline__ void adj_g2p_particle(const Mesh2<T>& M, const PRows<T>& R,
                                                 int64_t ld, const T* __restrict__ nb,
                                                 int64_t nld, T* __restrict__ cot,
                                                 T* __restrict__ wk, T* __restrict__ lebar,
                                                 const Mat2& mat, T dt, int vu, int p,
                                                 const PGeo<T>& q, T w[10]) {
  const int* ids = q.ids;
  const T xbx = CR(DMPM_C_X)[p], xby = CR(DMPM_C_X + 1)[p];
  const T vbx = CR(DMPM_C_V)[p], vby = CR(DMPM_C_V + 1)[p];
  T vx[4], vy[4];
  T sxi = T(0), seta = T(0);
  for (int k = 0; k < 4; ++k) {
    vx[k] = NR(DMPM_N_V)[ids[k]];
    vy[k] = NR(DMPM_N_V + 1)[ids[k]];
    T s = dt * (vx[k] * xbx + vy[k] * xby);
    if (vu) s = s + (vx[k] * vbx + vy[k] * vby);
    else s = s + dt * (NR(DMPM_N_A)[ids[k]] * vbx + NR(DMPM_N_A + 1)[ids[k]] * vby);
    T gx, gy;
    q4_dNdxi(k, q.xi, q.eta, gx, gy);
    sxi = sxi + gx * s;
    seta = seta + gy * s;
  }
  w[W_NVX] = vu ? dt * xbx + vbx : dt * xbx;
  w[W_NVY] = vu ? dt * xby + vby : dt * xby;
  w[W_NAX] = vu ? T(0) : dt * vbx;
  w[W_NAY] = vu ? T(0) : dt * vby;
  if (vu) {
    CR(DMPM_C_V)[p] = T(0);
    CR(DMPM_C_V + 1)[p] = T(0);
  }
  T Lb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (USL) {
    const T sb[3] = {CR(DMPM_C_SXX)[p], CR(DMPM_C_SYY)[p], CR(DMPM_C_SXY)[p]};
    T xb[2], Vb_in, unused = T(0);
    stress_adjoint(M.mat(mat, q.s), mat.per_E ? R.E[p] : T(0), dt,
                   R.V[p], q, vx, vy, sb, CR(DMPM_C_VOL)[p],
                   lebar ? lebar + p : &unused, Lb, xb, Vb_in);
    CR(DMPM_C_VOL)[p] = Vb_in;
    sxi = sxi + xb[0];
    seta = seta + xb[1];
  }
  for (int k = 0; k < 6; ++k) w[W_L00 + k] = Lb[k];
  WR(W_XI)[p] = sxi;
  WR(W_ETA)[p] = seta;
}

// Force P2G adjoint (element.py:236-268, 842-901): gathers the nodal force
// cotangent; USF then the adjoint of the stress update made with the
// kinematic nodal velocities.
// Force P2G adjoint (element.py:236-268, 842-901) of particle p (inside the
// mesh, geometry q): gathers the nodal force cotangent; USF then the adjoint
// of the stress update made with the kinematic nodal velocities, whose
// strain-rate cotangents (L00, L11, L01, L10, C00, C11) it returns in w.
template <typename T, bool USF>
__device__ __forceinline__ void adj_force_particle(
    const Mesh2<T>& M, const PRows<T>& R, const POut<T>& O, int64_t ld,
    const T* __restrict__ na, const T* __restrict__ vkin, T* __restrict__ cot,
    T* __restrict__ wk, T* __restrict__ lebar, const Mat2& mat, T Gx, T Gy, T dt, int p,
    const PGeo<T>& q, T w[6]) {
  const int* ids = q.ids;
  // the stress and volume the force P2G used: after the USF update, or the
  // step-start ones (USL)
  const T V = USF ? O.V[p] : R.V[p];
  const T sxx = USF ? O.sxx[p] : R.sxx[p];
  const T syy = USF ? O.syy[p] : R.syy[p];
  const T sxy = USF ? O.sxy[p] : R.sxy[p];
  const T m = R.m[p];
  T Vb = T(0), sbxx = T(0), sbyy = T(0), sbxy = T(0), sxi = T(0), seta = T(0);
  for (int k = 0; k < 4; ++k) {
    const T fx = AR(A_FX)[ids[k]], fy = AR(A_FY)[ids[k]];
    T gx, gy, dx, dy;
    q4_dNdxi(k, q.xi, q.eta, gx, gy);
    dx = gx * q.ix;
    dy = gy * q.iy;
    const T body = m * (Gx * fx + Gy * fy);
    sxi = sxi + gx * body;
    seta = seta + gy * body;
    Vb = Vb - (fx * (dx * sxx + dy * sxy) + fy * (dy * syy + dx * sxy));
    sbxx = sbxx + fx * dx;
    sbyy = sbyy + fy * dy;
    sbxy = sbxy + (fx * dy + fy * dx);
    const T s = T(mixed2(k));
    sxi = sxi - V * (s * q.iy) * (fx * sxy + fy * syy);
    seta = seta - V * (s * q.ix) * (fx * sxx + fy * sxy);
  }
  const T sb[3] = {CR(DMPM_C_SXX)[p] - V * sbxx, CR(DMPM_C_SYY)[p] - V * sbyy,
                   CR(DMPM_C_SXY)[p] - V * sbxy};
  const T Vbt = CR(DMPM_C_VOL)[p] + Vb;
  CR(DMPM_C_SXX)[p] = sb[0];
  CR(DMPM_C_SYY)[p] = sb[1];
  CR(DMPM_C_SXY)[p] = sb[2];
  if (USF) {
    T vx[4], vy[4];
    for (int k = 0; k < 4; ++k) {
      vx[k] = vkin[ids[k]];
      vy[k] = vkin[M.nnodes + ids[k]];
    }
    T Lb[6], xb[2], Vb_in;
    T dummy = T(0);
    stress_adjoint(M.mat(mat, q.s), mat.per_E ? R.E[p] : T(0), dt,
                   R.V[p], q, vx, vy, sb, Vbt, lebar ? lebar + p : &dummy, Lb, xb,
                   Vb_in);
    CR(DMPM_C_VOL)[p] = Vb_in;
    for (int k = 0; k < 6; ++k) w[k] = Lb[k];
    sxi = sxi + xb[0];
    seta = seta + xb[1];
  } else {
    CR(DMPM_C_VOL)[p] = Vbt;
  }
  WR(W_XI)[p] = WR(W_XI)[p] + sxi;
  WR(W_ETA)[p] = WR(W_ETA)[p] + seta;
}

// USL: the force adjoint of every particle in the mesh.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_force(
    Mesh2<T> M, PRows<T> R, int64_t ld, int n, const T* __restrict__ na,
    const T* __restrict__ vkin, T* __restrict__ cot, T* __restrict__ wk, T* __restrict__ lebar,
    Mat2 mat, T Gx, T Gy, T dt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const PGeo<T> q = pgeo(M, p, R.x[p], R.y[p]);
  if (!q.in) return;
  T w[6];
  POut<T> O;
  O.V = R.V;
  O.sxx = R.sxx;
  O.syy = R.syy;
  O.sxy = R.sxy;
  adj_force_particle<T, false>(M, R, O, ld, na, vkin, cot, wk, lebar, mat, Gx, Gy, dt, p, q, w);
}

// P2G of mass and momentum adjoint (element.py:122-189): gathers the nodal
// momentum and mass cotangents; the natural-coordinate cotangent goes to x.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_kin(Mesh2<T> M, PRows<T> R,
                                                      int64_t ld, int n,
                                                      const T* __restrict__ na,
                                                      T* __restrict__ cot, T* __restrict__ wk) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const PGeo<T> q = pgeo(M, p, R.x[p], R.y[p]);
  if (!q.in) return;
  const int* ids = q.ids;
  const T m = R.m[p];
  const T vx = R.vx[p], vy = R.vy[p];
  T N[4];
  q4_N4(q.xi, q.eta, N);
  T sxi = WR(W_XI)[p], seta = WR(W_ETA)[p];
  T vbx = T(0), vby = T(0);
  for (int k = 0; k < 4; ++k) {
    const T px = AR(A_PX)[ids[k]], py = AR(A_PY)[ids[k]], mb = AR(A_M)[ids[k]];
    vbx = vbx + N[k] * px;
    vby = vby + N[k] * py;
    T gx, gy;
    q4_dNdxi(k, q.xi, q.eta, gx, gy);
    const T s = m * (vx * px + vy * py) + m * mb;
    sxi = sxi + gx * s;
    seta = seta + gy * s;
  }
  CR(DMPM_C_V)[p] = CR(DMPM_C_V)[p] + m * vbx;
  CR(DMPM_C_V + 1)[p] = CR(DMPM_C_V + 1)[p] + m * vby;
  CR(DMPM_C_X)[p] = CR(DMPM_C_X)[p] + sxi * q.ix;
  CR(DMPM_C_X + 1)[p] = CR(DMPM_C_X + 1)[p] + seta * q.iy;
}

// Particle pass + cell pass of a cotangent scatter to the nodes.  The slot
// threads run a particle-local adjoint for the particles the sort put in the
// workgroup's cells (MODE 0, 1: G2P, adj_g2p_particle; MODE 2: force,
// adj_force_particle) and stage their natural coordinates and weights in
// LDS; then per corner k of a particle's cell the quantities
//   q0 = N_k w_nvx + dNk/dx L00b + dNk/dy L10b + dNk/dx(0) C00b
//   q1 = N_k w_nvy + dNk/dy L11b + dNk/dx L01b + dNk/dy(0) C11b
//   q2 = N_k w_nax,  q3 = N_k w_nay
// are summed per cell in slot order into partials[NQ * 4][ncells]
// (node_gather then adds the four cells of a node in the canonical order).
// MODE 0 (USL) forms all four; the USF step scatters in two passes whose
// other weights are zero: MODE 1 the N-weighted terms, MODE 2 the
// strain-rate terms (q0, q1 only; NQ = 2).  Workgroups from ncb on (MODE 0,
// 1) clear the cotangents of the particles outside the mesh.
constexpr int kSlotsA = 512;
template <typename T> struct AdjArgs {
  PRows<T> R;
  POut<T> O;
  int64_t ld;
  const T* nb;
  int64_t nld;
  const T* na;
  const T* vkin;
  T* cot;
  T* wk;
  T* lebar;
  Mat2 mat;
  T Gx, Gy, dt;
  int vu, n, ncb;
};
template <typename T, int CB, int MODE>
__global__ __launch_bounds__(kThreads) void k_adj_scatter(Mesh2<T> M, AdjArgs<T> a, SortRef sr,
                                                          T* __restrict__ part) {
  constexpr int NW = MODE == 0 ? 10 : (MODE == 1 ? 4 : 6);  // staged weights
  constexpr int W0 = MODE == 2 ? W_L00 : W_NVX;              // first staged weight
  constexpr int NQ = MODE == 2 ? 2 : 4;
  __shared__ T st[2 + NW][kSlotsA];
  const int* __restrict__ perm = sr.perm();
  const int* __restrict__ cs = sr.cs();
  const int ident = sr.flags[1];
  const PRows<T>& R = a.R;
  const int64_t ld = a.ld;
  T* __restrict__ cot = a.cot;
  if ((int)blockIdx.x >= a.ncb) {
    for (int s = cs[M.ncells] + ((int)blockIdx.x - a.ncb) * kThreads + threadIdx.x; s < a.n;
         s += (gridDim.x - a.ncb) * kThreads) {
      const int p = ident ? s : perm[s];
      for (int r = 0; r < DMPM_C_ROWS; ++r) CR(r)[p] = T(0);
    }
    return;
  }
  const int c0 = blockIdx.x * CB, t = threadIdx.x;
  const int cend = min(c0 + CB, M.ncells);
  const int mycell = c0 + t;
  const bool celler = t < CB && mycell < cend;
  const int ms0 = celler ? cs[mycell] : 0, ms1 = celler ? cs[mycell + 1] : 0;
  const int s0 = cs[c0], s1 = cs[cend];
  // the cell's inverse-Jacobian scales and shape-function gradients at the
  // centroid
  T ix = T(0), iy = T(0), cx[4], cy[4];
  if (celler) {
    const int cs_ = M.sim_c(mycell);
    const Grid2<T> g = M.grid(cs_);
    const int lc = mycell - M.coff(cs_);
    const int I = lc % g.nx, J = lc / g.nx;
    q4_invJ(T(0), T(0), g.X(I), g.X(I + 1), g.Y(J), g.Y(J + 1), ix, iy);
    for (int k = 0; k < 4; ++k) q4_dNdx(k, T(0), T(0), ix, iy, cx[k], cy[k]);
  }
  T acc[4 * NQ];
  for (int v = 0; v < 4 * NQ; ++v) acc[v] = T(0);
  for (int k0 = s0; k0 < s1; k0 += kSlotsA) {
    const int kend = min(k0 + kSlotsA, s1);
    for (int s = k0 + t; s < kend; s += kThreads) {
      const int p = ident ? s : perm[s];
      const int qq = s - k0;
      const PGeo<T> q = pgeo(M, p, R.x[p], R.y[p]);
      st[0][qq] = q.xi;
      st[1][qq] = q.eta;
      if (MODE == 2) {
        T w[6];
        adj_force_particle<T, true>(M, R, a.O, ld, a.na, a.vkin, cot, a.wk, a.lebar, a.mat,
                                    a.Gx, a.Gy, a.dt, p, q, w);
        for (int r = 0; r < 6; ++r) st[2 + r][qq] = w[r];
      } else {
        T w[10];
        adj_g2p_particle<T, MODE == 0>(M, R, ld, a.nb, a.nld, cot, a.wk, a.lebar, a.mat, a.dt,
                                       a.vu, p, q, w);
        for (int r = 0; r < NW; ++r) st[2 + r][qq] = w[W0 + r];
      }
    }
    __syncthreads();
    if (celler) {
      const int lo = max(ms0, k0) - k0, hi = min(ms1, kend) - k0;
      for (int qq = lo; qq < hi; ++qq) {
        const T xi = st[0][qq], eta = st[1][qq];
        T N[4];
        q4_N4(xi, eta, N);
#define WQ(r) st[2 + (r) - W0][qq]
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (MODE == 1) {
            acc[4 * k + 0] = acc[4 * k + 0] + N[k] * WQ(W_NVX);
            acc[4 * k + 1] = acc[4 * k + 1] + N[k] * WQ(W_NVY);
            acc[4 * k + 2] = acc[4 * k + 2] + N[k] * WQ(W_NAX);
            acc[4 * k + 3] = acc[4 * k + 3] + N[k] * WQ(W_NAY);
            continue;
          }
          T dx, dy;
          q4_dNdx(k, xi, eta, ix, iy, dx, dy);
          if (MODE == 2) {
            acc[2 * k + 0] =
                acc[2 * k + 0] + ((dx * WQ(W_L00) + dy * WQ(W_L10)) + cx[k] * WQ(W_C00));
            acc[2 * k + 1] =
                acc[2 * k + 1] + ((dy * WQ(W_L11) + dx * WQ(W_L01)) + cy[k] * WQ(W_C11));
            continue;
          }
          const T q0 = ((N[k] * WQ(W_NVX) + dx * WQ(W_L00)) + dy * WQ(W_L10)) + cx[k] * WQ(W_C00);
          const T q1 = ((N[k] * WQ(W_NVY) + dy * WQ(W_L11)) + dx * WQ(W_L01)) + cy[k] * WQ(W_C11);
          acc[4 * k + 0] = acc[4 * k + 0] + q0;
          acc[4 * k + 1] = acc[4 * k + 1] + q1;
          acc[4 * k + 2] = acc[4 * k + 2] + N[k] * WQ(W_NAX);
          acc[4 * k + 3] = acc[4 * k + 3] + N[k] * WQ(W_NAY);
        }
#undef WQ
      }
    }
    __syncthreads();
  }
  if (celler)
    for (int v = 0; v < 4 * NQ; ++v) part[(int64_t)v * M.ncells + mycell] = acc[v];
}

// Kinematics adjoint at one node and direction (element.py:191-202 with
// constraint.py:29-44): v = snap(snap(P)/m) (or the constraint value).
template <typename T>
__device__ __forceinline__ void kin_adjoint(T m, T v, bool constrained, T vb, T& Pb, T& mb) {
  Pb = T(0);
  if (constrained || m == T(0) || v == T(0)) return;
  Pb = vb / m;
  mb = mb - vb * v / m;
}

// Node pass 1: nodal-velocity/acceleration cotangents of G2P, adjoint of the
// nodal update (element.py:305-329): F = f_int + f_ext + f_damp,
// a0 = nan_to_num(F/m), v = snap(v_kin + a0 dt), a = snap(a0), constraints.
// USL: the kinematic velocity's cotangent is then complete and the
// kinematics adjoint follows.
template <typename T, bool USL>
__global__ __launch_bounds__(kThreads) void k_adj_node1(
    Mesh2<T> M, const T* __restrict__ part, const T* __restrict__ nb, int64_t nld,
    const T* __restrict__ vkin, T* __restrict__ na, const int* __restrict__ cidx,
    const int* __restrict__ cmask, T dt) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= M.nnodes) return;
  const int sm = M.sim_n(nd);
  const Grid2<T> g = M.grid(sm);
  const int nl = nd - M.noff(sm);
  const int i = nl % (g.nx + 1), j = nl / (g.nx + 1);
  T acc[4];
  node_gather<T, 4>(g, M.coff(sm), M.ncells, part, i, j, acc);
  const int cr = cidx ? cidx[nd] : -1;
  const int mask = cr >= 0 ? cmask[cr] : 0;
  const T m = NR(DMPM_N_MASS)[nd];
  T mb = T(0);
  for (int d = 0; d < 2; ++d) {
    T Fb = T(0), vkb = T(0);
    if (!(mask & (1 << d))) {
      const T F = (NR(DMPM_N_FINT + d)[nd] + NR(DMPM_N_FEXT + d)[nd]) + NR(DMPM_N_FDAMP + d)[nd];
      const T r = F / m;
      const T v1b = NR(DMPM_N_V + d)[nd] != T(0) ? acc[d] : T(0);
      const T a0b = (NR(DMPM_N_A + d)[nd] != T(0) ? acc[2 + d] : T(0)) + dt * v1b;
      if (isfinite(r)) {
        Fb = a0b / m;
        mb = mb - a0b * r / m;
      }
      vkb = v1b;
    }
    AR(A_FX + d)[nd] = Fb;
    if (USL) {
      T Pb;
      kin_adjoint(m, vkin[(int64_t)d * M.nnodes + nd], (mask >> d) & 1, vkb, Pb, mb);
      AR(A_PX + d)[nd] = Pb;
    } else {
      AR(A_VKX + d)[nd] = vkb;
    }
  }
  AR(A_M)[nd] = mb;
}

// Node pass 2 (USF): the kinematic velocity's cotangent from the nodal update
// plus the strain-rate scatter; kinematics adjoint.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_adj_node2(Mesh2<T> M, const T* __restrict__ part,
                                                        const T* __restrict__ nb, int64_t nld,
                                                        const T* __restrict__ vkin,
                                                        T* __restrict__ na,
                                                        const int* __restrict__ cidx,
                                                        const int* __restrict__ cmask) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= M.nnodes) return;
  const int sm = M.sim_n(nd);
  const Grid2<T> g = M.grid(sm);
  const int nl = nd - M.noff(sm);
  const int i = nl % (g.nx + 1), j = nl / (g.nx + 1);
  T acc[2];
  node_gather<T, 2>(g, M.coff(sm), M.ncells, part, i, j, acc);
  const int cr = cidx ? cidx[nd] : -1;
  const int mask = cr >= 0 ? cmask[cr] : 0;
  const T m = NR(DMPM_N_MASS)[nd];
  T mb = AR(A_M)[nd];
  for (int d = 0; d < 2; ++d) {
    const T vb = AR(A_VKX + d)[nd] + acc[d];
    T Pb;
    kin_adjoint(m, vkin[(int64_t)d * M.nnodes + nd], (mask >> d) & 1, vb, Pb, mb);
    AR(A_PX + d)[nd] = Pb;
  }
  AR(A_M)[nd] = mb;
}

template <typename T, int CB>
int adjoint_kernels(dmpm_plan* P, const Mesh2<T>& M, const PRows<T>& R, const POut<T>& O,
                    const T* nb, const T* vkin, SortRef sr, T* cot, T* lebar, int scheme, int vu,
                    const double* grav, double dt, hipStream_t s);

template <typename T, int CB>
int adjoint_t(dmpm_plan* P, const T* pb, T* pw, int32_t* eid, T* nb, T* cot, T* lebar,
              int scheme, int vu, const double* grav, double dt, int64_t step, hipStream_t s) {
  const Mesh2<T> M = mesh_of<T>(P);
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  // the rows the forward step reads: x, v, mass, volume, density, stress, E
  const int spans[3][2] = {{DMPM_P_X, DMPM_P_DENSITY + 1},
                           {DMPM_P_STRESS, DMPM_P_STRESS + 4},
                           {DMPM_P_YOUNGS, DMPM_P_YOUNGS + 1}};
  for (const auto& sp : spans)
    DMPM_TRY(hip_check(hipMemcpyAsync(pw + (int64_t)sp[0] * ld, pb + (int64_t)sp[0] * ld,
                                      (size_t)(sp[1] - sp[0]) * ld * sizeof(T),
                                      hipMemcpyDeviceToDevice, s),
                       "hipMemcpyAsync"));
  DMPM_TRY(tile_run(P, pw, eid, nb, scheme, vu, grav, dt, step, 1, 0, s, nullptr));
  const T* vkin = (const T*)plan_workspace(P, 1, (size_t)2 * P->nnodes * sizeof(T));
  if (!vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  return adjoint_kernels<T, CB>(P, M, rows_of_block(pb, ld), out_of_block<T>(pw, ld), nb, vkin,
                                sort_ref(P), cot, lebar, scheme, vu, grav, dt, s);
}

// The adjoint kernels of one step from its forward quantities: pb the step's
// input block, pw a block whose stress and volume rows hold the values after
// the (USF) stress update, nb the node block after the step, vkin the
// kinematic nodal velocity, sr the step's sort by cell.
template <typename T, int CB>
int adjoint_kernels(dmpm_plan* P, const Mesh2<T>& M, const PRows<T>& R, const POut<T>& O,
                    const T* nb, const T* vkin, SortRef sr, T* cot, T* lebar, int scheme, int vu,
                    const double* grav, double dt, hipStream_t s) {
  const int64_t ld = P->d.ld, nld = P->d.nld;
  const int n = (int)P->n;
  const bool usf = scheme == DMPM_USF;
  ProfScopeT ps_(P, KC_ADJ, s);
  T* part = (T*)plan_workspace(P, 0, (size_t)16 * M.ncells * sizeof(T));
  T* wk = (T*)plan_workspace(P, 3, (size_t)W_ROWS * ld * sizeof(T));
  T* na = (T*)plan_workspace(P, 4, (size_t)A_ROWS * P->nnodes * sizeof(T));
  if (!part || !vkin || !wk || !na) return fail(DMPM_EHIP, "workspace allocation failed");
  Mat2 mat;
  mat.kind = P->d.material;
  mat.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) mat.de[k] = P->d.de[k];
  const T tdt = T(dt), Gx = T(grav[0]), Gy = T(grav[1]);
  const unsigned np = nblk(n), nn = nblk(P->nnodes);
  const unsigned ncb = (unsigned)((M.ncells + CB - 1) / CB);
  AdjArgs<T> aa;
  aa.R = R;
  aa.O = O;
  aa.ld = ld;
  aa.nb = nb;
  aa.nld = nld;
  aa.na = na;
  aa.vkin = vkin;
  aa.cot = cot;
  aa.wk = wk;
  aa.lebar = lebar;
  aa.mat = mat;
  aa.Gx = Gx;
  aa.Gy = Gy;
  aa.dt = tdt;
  aa.vu = vu;
  aa.n = n;
  aa.ncb = (int)ncb;
  constexpr int kOutsideBlocks = 8;
  if (usf)
    k_adj_scatter<T, CB, 1><<<ncb + kOutsideBlocks, kThreads, 0, s>>>(M, aa, sr, part);
  else
    k_adj_scatter<T, CB, 0><<<ncb + kOutsideBlocks, kThreads, 0, s>>>(M, aa, sr, part);
  if (usf)
    k_adj_node1<T, false><<<nn, kThreads, 0, s>>>(M, part, nb, nld, vkin, na, P->cidx, P->cmask,
                                                   tdt);
  else
    k_adj_node1<T, true><<<nn, kThreads, 0, s>>>(M, part, nb, nld, vkin, na, P->cidx, P->cmask,
                                                  tdt);
  if (usf) {
    k_adj_scatter<T, CB, 2><<<ncb, kThreads, 0, s>>>(M, aa, sr, part);
    k_adj_node2<T><<<nn, kThreads, 0, s>>>(M, part, nb, nld, vkin, na, P->cidx, P->cmask);
  } else {
    k_adj_force<T><<<np, kThreads, 0, s>>>(M, R, ld, n, na, vkin, cot, wk, lebar, mat, Gx, Gy,
                                          tdt);
  }
  k_adj_kin<T><<<np, kThreads, 0, s>>>(M, R, ld, n, na, cot, wk);
  return launch_check("step_adjoint");
}

// Copies the sort order of the current step (buffer selected by flags[4]).
__global__ void k_export_sort(const int* __restrict__ perm0, const int* __restrict__ perm1,
                              const int* __restrict__ cs0, const int* __restrict__ cs1,
                              const int* __restrict__ flags, int n, int ncs,
                              int* __restrict__ perm_out, int* __restrict__ cs_out) {
  const int cur = flags[4];
  const int* __restrict__ perm = cur ? perm1 : perm0;
  const int* __restrict__ cs = cur ? cs1 : cs0;
  const int gs = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gs) perm_out[i] = perm[i];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncs; i += gs) cs_out[i] = cs[i];
}

#undef PR
#undef PW
#undef NR
#undef CR
#undef WR
#undef AR

}  // namespace

extern "C" int dmpm_plan_export_step(dmpm_plan* P, void* vkin_out, int32_t* perm_out,
                                     int32_t* cs_out, void* stream) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (!vkin_out || !perm_out || !cs_out) return fail(DMPM_EINVAL, "null buffer");
  if (!tile_eligible(P, 0)) return fail(DMPM_EINVAL, "no fused 2D step on this plan");
  hipStream_t s = (hipStream_t)stream;
  const size_t esz = P->d.dtype == DMPM_F32 ? sizeof(float) : sizeof(double);
  void* vkin = plan_workspace(P, 1, 2 * (size_t)P->nnodes * esz);
  if (!vkin) return fail(DMPM_EHIP, "workspace allocation failed");
  DMPM_TRY(hip_check(hipMemcpyAsync(vkin_out, vkin, 2 * (size_t)P->nnodes * esz,
                                    hipMemcpyDeviceToDevice, s),
                     "hipMemcpyAsync"));
  k_export_sort<<<1024, kThreads, 0, s>>>(P->perm, P->perm2, P->cs, P->cs2, P->flags, (int)P->n,
                                          P->ncells + 2, perm_out, cs_out);
  return launch_check("export_step");
}

extern "C" int dmpm_step_adjoint_saved(dmpm_plan* P, const void* pin, const void* pout,
                                       const void* nsaved, const void* vkin_saved,
                                       const int32_t* perm_saved, const int32_t* cs_saved,
                                       void* cot, void* lebar, int32_t scheme,
                                       int32_t velocity_update, const double* gravity,
                                       double dt, void* stream) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (!pin || !pout || !nsaved || !vkin_saved || !perm_saved || !cs_saved || !cot || !gravity)
    return fail(DMPM_EINVAL, "null buffer");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "the adjoint step covers 2D Quadrilateral4Node meshes without particle "
                "tractions or particle external forces");
  SortRef sr;
  sr.perm0 = sr.perm1 = perm_saved;
  sr.cs0 = sr.cs1 = cs_saved;
  sr.flags = P->zflags;  // not the identity, buffer 0
  hipStream_t s = (hipStream_t)stream;
  const int64_t ld = P->d.ld;
  if (P->d.dtype == DMPM_F32)
    return adjoint_kernels<float, 128>(
        P, mesh_of<float>(P), rows_of_block((const float*)pin, ld),
        out_of_block((const float*)pout, ld), (const float*)nsaved, (const float*)vkin_saved, sr,
        (float*)cot, (float*)lebar, scheme, velocity_update, gravity, dt, s);
  return adjoint_kernels<double, 128>(
      P, mesh_of<double>(P), rows_of_block((const double*)pin, ld),
      out_of_block((const double*)pout, ld), (const double*)nsaved, (const double*)vkin_saved,
      sr, (double*)cot, (double*)lebar, scheme, velocity_update, gravity, dt, s);
}

extern "C" int dmpm_step_adjoint_saved_state(
    dmpm_plan* P, const void* pconst, const void* state_in, const void* state_out,
    int64_t state_ld, const void* nsaved, const void* vkin_saved, const int32_t* perm_saved,
    const int32_t* cs_saved, void* cot, void* lebar, int32_t scheme, int32_t velocity_update,
    const double* gravity, double dt, void* stream) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (!pconst || !state_in || !state_out || !nsaved || !vkin_saved || !perm_saved ||
      !cs_saved || !cot || !gravity)
    return fail(DMPM_EINVAL, "null buffer");
  if (state_ld < P->n) return fail(DMPM_EINVAL, "state_ld below the particle count");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "the adjoint step covers 2D Quadrilateral4Node meshes without particle "
                "tractions or particle external forces");
  SortRef sr;
  sr.perm0 = sr.perm1 = perm_saved;
  sr.cs0 = sr.cs1 = cs_saved;
  sr.flags = P->zflags;  // not the identity, buffer 0
  hipStream_t s = (hipStream_t)stream;
  const int64_t ld = P->d.ld;
  if (P->d.dtype == DMPM_F32)
    return adjoint_kernels<float, 128>(
        P, mesh_of<float>(P),
        rows_of_state((const float*)state_in, state_ld, (const float*)pconst, ld),
        out_of_state((const float*)state_out, state_ld), (const float*)nsaved,
        (const float*)vkin_saved, sr, (float*)cot, (float*)lebar, scheme, velocity_update,
        gravity, dt, s);
  return adjoint_kernels<double, 128>(
      P, mesh_of<double>(P),
      rows_of_state((const double*)state_in, state_ld, (const double*)pconst, ld),
      out_of_state((const double*)state_out, state_ld), (const double*)nsaved,
      (const double*)vkin_saved, sr, (double*)cot, (double*)lebar, scheme, velocity_update,
      gravity, dt, s);
}

extern "C" int dmpm_step_adjoint(dmpm_plan* P, const void* pin, void* pwork, int32_t* eid_work,
                                 void* nwork, void* cot, void* lebar, int32_t scheme,
                                 int32_t velocity_update, const double* gravity, double dt,
                                 int64_t step, void* stream) {
  if (!P) return fail(DMPM_EINVAL, "null plan");
  if (!pin || !pwork || !eid_work || !nwork || !cot || !gravity)
    return fail(DMPM_EINVAL, "null buffer");
  if (scheme != DMPM_USF && scheme != DMPM_USL)
    return fail(DMPM_EINVAL, "scheme must be DMPM_USF or DMPM_USL");
  if (P->d.material != DMPM_LINEAR_ELASTIC && P->d.material != DMPM_SIMPLE_MATERIAL)
    return fail(DMPM_EINVAL, "unknown material");
  hipStream_t s = (hipStream_t)stream;
  if (P->dim == 1)
    return adjoint_1d(P, pin, pwork, eid_work, nwork, cot, lebar, scheme, velocity_update,
                      gravity, dt, step, s);
  if (!tile_eligible(P, 0))
    return fail(DMPM_EINVAL,
                "the adjoint step covers Linear1D meshes without particle tractions and 2D "
                "Quadrilateral4Node meshes without particle tractions or particle external "
                "forces");
  if (P->d.dtype == DMPM_F32)
    return adjoint_t<float, 128>(P, (const float*)pin, (float*)pwork, eid_work, (float*)nwork,
                                 (float*)cot, (float*)lebar, scheme, velocity_update, gravity, dt,
                                 step, s);
  return adjoint_t<double, 128>(P, (const double*)pin, (double*)pwork, eid_work, (double*)nwork,
                                (double*)cot, (double*)lebar, scheme, velocity_update, gravity,
                                dt, step, s);
}
