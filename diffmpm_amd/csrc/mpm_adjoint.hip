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
//   k_adj_tile1   per node tile (mpm_tiles.h): G2P adjoint of the particles of
//                 the tile's cells and of the column and row of cells left of
//                 and below it (cotangents of the nodal velocity/acceleration
//                 as per-particle weights, the natural-coordinate cotangent of
//                 G2P; USL: stress adjoint), their scatter to the nodes the
//                 tile owns (cell sums, node sums) and the grid-update adjoint
//                 there (a = F/m, v' = v + a dt, snaps, constraints; USL:
//                 kinematics adjoint, v = P/m)
//   k_adj_tile2   (USF) per node tile: internal/body force adjoint (gather of
//                 Fbar), stress adjoint (sigma' = sigma + D deps, V' = V (1 +
//                 dvol)), scatter of the strain-rate cotangents to the owned
//                 nodes, kinematics adjoint there
//   k_adj_force   (USL) the force adjoint, per particle
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
#include "mpm_tiles.h"

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
// rows of the weight workspace that carry the updated stress and volume
// cotangents of the particles in boundary cells between two launches (the
// weights themselves stay in LDS)
enum { W_BSXX = 0, W_BSYY, W_BSXY, W_BVOL };
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

// Whether cell (I, J) of a mesh is in the last column or row of its tile
// (a cell whose particles another tile reads).
__device__ __forceinline__ bool boundary_cell(int I, int J) {
  return (I % kStepTile) == kStepTile - 1 || (J % kStepTile) == kStepTile - 1;
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

// G2P adjoint (particle.py:208-247) of one particle (geometry q) from its
// position and velocity cotangents (xbx, xby, vbx, vby) and the velocities
// (vx, vy) and accelerations (ax, ay) of its cell's nodes: the weights
// w[W_NVX .. W_C11] that the scatter to the nodes needs and the
// natural-coordinate cotangent (sxi, seta).  USL also the adjoint of the
// stress update made with the updated nodal velocities (scheme.py:127-133):
// sb and Vb are the stress and volume cotangents of the step output, Vb_in
// receives the step-start volume's, *lebar accumulates dL/dlogE.
template <typename T, bool USL>
__device__ __forceinline__ void adj_g2p_core(const Mat2& mat, T E_p, T V_in, T dt, int vu,
                                             const PGeo<T>& q, T xbx, T xby, T vbx, T vby,
                                             const T vx[4], const T vy[4], const T ax[4],
                                             const T ay[4], const T sb[3], T Vb, T* lebar,
                                             T& Vb_in, T& sxi_out, T& seta_out, T w[10]) {
  T sxi = T(0), seta = T(0);
  for (int k = 0; k < 4; ++k) {
    T s = dt * (vx[k] * xbx + vy[k] * xby);
    if (vu) s = s + (vx[k] * vbx + vy[k] * vby);
    else s = s + dt * (ax[k] * vbx + ay[k] * vby);
    T gx, gy;
    q4_dNdxi(k, q.xi, q.eta, gx, gy);
    sxi = sxi + gx * s;
    seta = seta + gy * s;
  }
  w[W_NVX] = vu ? dt * xbx + vbx : dt * xbx;
  w[W_NVY] = vu ? dt * xby + vby : dt * xby;
  w[W_NAX] = vu ? T(0) : dt * vbx;
  w[W_NAY] = vu ? T(0) : dt * vby;
  T Lb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (USL) {
    T xb[2];
    stress_adjoint(mat, E_p, dt, V_in, q, vx, vy, sb, Vb, lebar, Lb, xb, Vb_in);
    sxi = sxi + xb[0];
    seta = seta + xb[1];
  }
  for (int k = 0; k < 6; ++k) w[W_L00 + k] = Lb[k];
  sxi_out = sxi;
  seta_out = seta;
}

// Force P2G adjoint (element.py:236-268, 842-901) of one particle (geometry
// q, mass m, the stress and volume the force P2G used) from the force
// cotangents (fx, fy) of its cell's nodes and the stress and volume
// cotangents cs[3], cvol of the step output: the new stress cotangents sb,
// the volume cotangent vol (USF: of the step-start volume V_in, after the
// adjoint of the stress update made with the kinematic nodal velocities vkx,
// vky, whose strain-rate cotangents go to w), the natural-coordinate
// cotangent (sxi, seta); *lebar accumulates dL/dlogE.
template <typename T, bool USF>
__device__ __forceinline__ void adj_force_core(const Mat2& mat, T E_p, T dt, T V_in, T V, T sxx,
                                               T syy, T sxy, T m, const PGeo<T>& q, T Gx, T Gy,
                                               const T fx[4], const T fy[4], const T vkx[4],
                                               const T vky[4], const T cs[3], T cvol, T* lebar,
                                               T sb[3], T& vol, T w[6], T& sxi_out,
                                               T& seta_out) {
  T Vb = T(0), sbxx = T(0), sbyy = T(0), sbxy = T(0), sxi = T(0), seta = T(0);
  for (int k = 0; k < 4; ++k) {
    T gx, gy, dx, dy;
    q4_dNdxi(k, q.xi, q.eta, gx, gy);
    dx = gx * q.ix;
    dy = gy * q.iy;
    const T body = m * (Gx * fx[k] + Gy * fy[k]);
    sxi = sxi + gx * body;
    seta = seta + gy * body;
    Vb = Vb - (fx[k] * (dx * sxx + dy * sxy) + fy[k] * (dy * syy + dx * sxy));
    sbxx = sbxx + fx[k] * dx;
    sbyy = sbyy + fy[k] * dy;
    sbxy = sbxy + (fx[k] * dy + fy[k] * dx);
    const T s = T(mixed2(k));
    sxi = sxi - V * (s * q.iy) * (fx[k] * sxy + fy[k] * syy);
    seta = seta - V * (s * q.ix) * (fx[k] * sxx + fy[k] * sxy);
  }
  sb[0] = cs[0] - V * sbxx;
  sb[1] = cs[1] - V * sbyy;
  sb[2] = cs[2] - V * sbxy;
  const T Vbt = cvol + Vb;
  if (USF) {
    T Lb[6], xb[2], Vb_in;
    stress_adjoint(mat, E_p, dt, V_in, q, vkx, vky, sb, Vbt, lebar, Lb, xb, Vb_in);
    vol = Vb_in;
    for (int k = 0; k < 6; ++k) w[k] = Lb[k];
    sxi = sxi + xb[0];
    seta = seta + xb[1];
  } else {
    vol = Vbt;
  }
  sxi_out = sxi;
  seta_out = seta;
}

// The same for particle p (inside the mesh, geometry q) with the nodal values
// read from the nodal adjoint rows and vkin: cotangent updates, the
// natural-coordinate cotangent added to rows W_XI, W_ETA of wk, the
// strain-rate cotangents to w (USF).
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
  T fx[4], fy[4], vkx[4], vky[4];
  for (int k = 0; k < 4; ++k) {
    fx[k] = AR(A_FX)[ids[k]];
    fy[k] = AR(A_FY)[ids[k]];
    vkx[k] = USF ? vkin[ids[k]] : T(0);
    vky[k] = USF ? vkin[M.nnodes + ids[k]] : T(0);
  }
  const T cs[3] = {CR(DMPM_C_SXX)[p], CR(DMPM_C_SYY)[p], CR(DMPM_C_SXY)[p]};
  // the volume cotangent after the G2P adjoint (a boundary cell's particle
  // left it in the workspace)
  const T cvol = boundary_cell(q.I, q.J) ? WR(W_BVOL)[p] : CR(DMPM_C_VOL)[p];
  T sb[3], vol, sxi, seta, dummy = T(0);
  adj_force_core<T, USF>(M.mat(mat, q.s), mat.per_E ? R.E[p] : T(0), dt, R.V[p], V, sxx, syy,
                         sxy, R.m[p], q, Gx, Gy, fx, fy, vkx, vky, cs, cvol,
                         lebar ? lebar + p : &dummy, sb, vol, w, sxi, seta);
  CR(DMPM_C_SXX)[p] = sb[0];
  CR(DMPM_C_SYY)[p] = sb[1];
  CR(DMPM_C_SXY)[p] = sb[2];
  CR(DMPM_C_VOL)[p] = vol;
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
                                                      T* __restrict__ cot, T* __restrict__ wk,
                                                      int vu, int usf) {
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
  // (with the velocity update the G2P adjoint leaves a zero velocity
  // cotangent)
  CR(DMPM_C_V)[p] = (vu ? T(0) : CR(DMPM_C_V)[p]) + m * vbx;
  CR(DMPM_C_V + 1)[p] = (vu ? T(0) : CR(DMPM_C_V + 1)[p]) + m * vby;
  CR(DMPM_C_X)[p] = CR(DMPM_C_X)[p] + sxi * q.ix;
  CR(DMPM_C_X + 1)[p] = CR(DMPM_C_X + 1)[p] + seta * q.iy;
  if (usf && boundary_cell(q.I, q.J)) {
    CR(DMPM_C_SXX)[p] = WR(W_BSXX)[p];
    CR(DMPM_C_SYY)[p] = WR(W_BSYY)[p];
    CR(DMPM_C_SXY)[p] = WR(W_BSXY)[p];
    CR(DMPM_C_VOL)[p] = WR(W_BVOL)[p];
  }
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

// ---- node-tile adjoint ---------------------------------------------------------
//
// Two launches over the tiles of the forward step, each workgroup a node tile
// as in the split forward step (mpm_tiles.h): tile (ti, tj) owns the cells R0
// and the nodes NR; the cells of those nodes are RC = R0 grown by one cell to
// the left and below, whose particles run the particle-local adjoint (only
// the tile's own particles write their cotangents).  The nodal values the
// particles read are staged in LDS for the nodes of RC's cells, the weights
// are summed per cell in slot order and per node in node_gather's order, and
// the node pass runs for the owned nodes in the same launch, so no cell
// partials go through HBM:
//   k_adj_tile1<0>  USL: G2P and stress adjoint, scatter, nodal-update and
//                   kinematics adjoint
//   k_adj_tile1<1>  USF: G2P adjoint, scatter, nodal-update adjoint
//   k_adj_tile2     USF: force and stress adjoint, strain-rate scatter,
//                   kinematics adjoint
// Workgroups from the tile count on (k_adj_tile1) clear the cotangents of the
// particles outside the mesh.
//
// A tile reads the cotangents of the particles in the column of cells left
// of it and the row below it while the tiles that own them update them.  So
// a particle in the last column or row of a tile (a boundary cell, the only
// cells another tile reads) writes its updated stress and volume cotangents
// to rows of the weight workspace (W_BSXX..W_BVOL), and the next particle
// launch (k_adj_kin; USL: k_adj_force for the volume) moves them into place;
// the G2P adjoint's zero velocity cotangent (velocity update) is applied by
// k_adj_kin.
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

template <typename T> struct AdjTile {
  AdjArgs<T> a;
  TileTab<T> tt;
  SortRef sr;
  const int* cidx;
  const int* cmask;
  int nn_all, nc_all;
};

// threads, slots per thread and round, and waves per SIMD to fit (register
// budget) of each tile kernel
template <int MODE> struct AdjTileShape;
template <> struct AdjTileShape<0> { static constexpr int NT = 512, RPT = 3, WPE = 4; };
template <> struct AdjTileShape<1> { static constexpr int NT = 512, RPT = 3, WPE = 1; };
template <> struct AdjTileShape<2> { static constexpr int NT = 512, RPT = 3, WPE = 1; };
constexpr int kACells = (kStepTile + 1) * (kStepTile + 1);  // cells of RC (at most)
constexpr int kANodes = (kStepTile + 1) * (kStepTile + 1);  // owned nodes (at most)
constexpr int kAVN = (kStepTile + 2) * (kStepTile + 2);     // nodes of RC's cells
constexpr int kAdjOutsideBlocks = 8;
constexpr int kARow32 = 64;  // entries of a tile's row table (virtual indices up to 2048)

// Particle geometry of a slot of RC: the slot's cell (local column ci, row r
// of RC) from the slot table, natural coordinates and scales from the
// tables, corner indices in the node grid of RC's cells (row length wV).
template <typename T>
__device__ __forceinline__ PGeo<T> tile_geo(T x, T y, int ci, int r, int wV, const T* colC,
                                            const T* colS, const T* rowC, const T* rowS) {
  PGeo<T> q;
  q.in = true;
  q.s = 0;
  q.I = ci;
  q.J = r;
  q.xi = (x - colC[ci]) * colS[ci];
  q.eta = (y - rowC[r]) * rowS[r];
  q.ix = colS[ci];
  q.iy = rowS[r];
  const int ll = ci + r * wV;
  q.ids[0] = ll;
  q.ids[1] = ll + 1;
  q.ids[2] = ll + wV + 1;
  q.ids[3] = ll + wV;
  return q;
}

template <typename T, bool B, int MODE, int NT = AdjTileShape<MODE>::NT,
          int RPT = AdjTileShape<MODE>::RPT, int WPE = AdjTileShape<MODE>::WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_adj_tile1(
    AdjTile<T> A) {
  constexpr int CAP = NT * RPT;
  constexpr int NW = MODE == 0 ? 10 : 4;  // staged weights, from W_NVX
  constexpr int NQ = 4;
  constexpr int CC = (kACells + NT - 1) / NT;
  constexpr int NPT = (kANodes + NT - 1) / NT;
  constexpr bool USL = MODE == 0;
  static_assert(4 * NQ * kACells <= (2 + NW) * CAP, "node-sum slices");
  __shared__ T st[(2 + NW) * CAP];
  // USL: slot q's table entry lies in slot q's element of staged row NW (the
  // slot's thread reads it before it writes its weights there), which keeps
  // the kernel under half the CU's LDS
  __shared__ uint16_t tab_[USL ? 1 : CAP];
  static_assert(sizeof(T) % sizeof(uint16_t) == 0, "slot table in a weight element");
  constexpr int kTs = USL ? (int)(sizeof(T) / sizeof(uint16_t)) : 1;
  uint16_t* const tab = USL ? reinterpret_cast<uint16_t*>(st + NW * CAP) : tab_;
  __shared__ T nvv[4 * kAVN];  // nodal vx, vy, ax, ay at the nodes of RC's cells
  __shared__ T colC[kStepTile + 1], colS[kStepTile + 1], rowC[kStepTile + 1],
      rowS[kStepTile + 1];
  __shared__ int segA[kStepTile + 1], segP[kStepTile + 2], rowOf[kARow32];
  const AdjArgs<T>& a = A.a;
  const PRows<T>& R = a.R;
  const int64_t ld = a.ld, nld = a.nld;
  const T* __restrict__ nb = a.nb;
  T* __restrict__ cot = a.cot;
  T* __restrict__ wk = a.wk;
  T* __restrict__ na = const_cast<T*>(a.na);
  const int* __restrict__ cs = A.sr.cs();
  const int* __restrict__ perm = A.sr.perm();
  const int ident = A.sr.flags[1];
  const int t = threadIdx.x;
  const int ntiles = A.tt.ntiles;
  if ((int)blockIdx.x >= ntiles) {
    for (int s = cs[A.nc_all] + ((int)blockIdx.x - ntiles) * NT + t; s < a.n;
         s += kAdjOutsideBlocks * NT) {
      const int p = ident ? s : perm[s];
      for (int r = 0; r < DMPM_C_ROWS; ++r) CR(r)[p] = T(0);
    }
    return;
  }
  const TileCtx<T> c = tile_ctx_of<T, B>(A.tt, xcd_pos(blockIdx.x, ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx, ny = g.ny;
  const Rgn R0{c.ti * kStepTile, min(c.ti * kStepTile + kStepTile, nx), c.tj * kStepTile,
               min(c.tj * kStepTile + kStepTile, ny)};
  const Rgn RC{max(R0.i0 - 1, 0), R0.i1, max(R0.j0 - 1, 0), R0.j1};
  const Rgn NR = owned_nodes(R0, nx, ny);
  const int hC = RC.j1 - RC.j0, wC = RC.i1 - RC.i0;
  const int wV = wC + 1, nV = wV * (hC + 1);
  const int oi = R0.i0 - RC.i0, oj = R0.j0 - RC.j0;
  const int nnodes = A.nn_all;
  int cr_[CC], ccol[CC], cs0[CC], cs1[CC];
  cell_slots_rc<CC, NT>(RC, nx, c.coff, cs, cr_, ccol, cs0, cs1);
  if (t < 64) row_segments(RC, nx, c.coff, cs, segA, segP, rowOf, kARow32);
  else if (t < 128) cell_tables(g, RC, colC, colS, rowC, rowS);
  {
    const SmallDiv dwv = small_div(wV);
    for (int ln = t; ln < nV; ln += NT) {
      const int jr = sdiv(dwv, ln);
      const int nd = c.noff + (RC.j0 + jr) * (nx + 1) + RC.i0 + (ln - jr * wV);
      nvv[ln] = NR(DMPM_N_V)[nd];
      nvv[kAVN + ln] = NR(DMPM_N_V + 1)[nd];
      nvv[2 * kAVN + ln] = NR(DMPM_N_A)[nd];
      nvv[3 * kAVN + ln] = NR(DMPM_N_A + 1)[nd];
    }
  }
  // the owned nodes' rows the node pass reads, loaded ahead
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const SmallDiv dwn = small_div(wn);
  int ndx[NPT], mask[NPT];
  T nm[NPT], nF[NPT][2], nv1[NPT][2], na1[NPT][2], vk1[NPT][2];
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = min(t + w * NT, nn - 1);
    const int jr = sdiv(dwn, ln);
    const int nd = c.noff + (NR.j0 + jr) * (nx + 1) + NR.i0 + (ln - jr * wn);
    ndx[w] = nd;
    const int cr = A.cidx ? A.cidx[nd] : -1;
    mask[w] = cr >= 0 ? A.cmask[cr] : 0;
    nm[w] = NR(DMPM_N_MASS)[nd];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      nF[w][d] = (NR(DMPM_N_FINT + d)[nd] + NR(DMPM_N_FEXT + d)[nd]) + NR(DMPM_N_FDAMP + d)[nd];
      nv1[w][d] = NR(DMPM_N_V + d)[nd];
      na1[w][d] = NR(DMPM_N_A + d)[nd];
      vk1[w][d] = USL ? a.vkin[(int64_t)d * nnodes + nd] : T(0);
    }
  }
  __syncthreads();
  const int S = segP[hC];
  int cv0[CC], cv1[CC];
  cell_ranges_of<CC>(cr_, cs0, cs1, segA, segP, cv0, cv1);
  // per cell: inverse-Jacobian scales and the shape-function gradients at
  // the centroid (USL's strain-rate terms)
  T cix[CC], ciy[CC], cx[CC][4], cy[CC][4];
#pragma unroll
  for (int u = 0; u < CC; ++u) {
    cix[u] = ciy[u] = T(0);
    for (int k = 0; k < 4; ++k) cx[u][k] = cy[u][k] = T(0);
    if (USL && cr_[u] >= 0) {
      cix[u] = colS[ccol[u]];
      ciy[u] = rowS[cr_[u]];
      for (int k = 0; k < 4; ++k) q4_dNdx(k, T(0), T(0), cix[u], ciy[u], cx[u][k], cy[u][k]);
    }
  }
  T acc[CC][4 * NQ];
#pragma unroll
  for (int u = 0; u < CC; ++u)
#pragma unroll
    for (int v = 0; v < 4 * NQ; ++v) acc[u][v] = T(0);
  for (int v0 = 0; v0 < S; v0 += CAP) {
    const int vend = min(v0 + CAP, S);
    if (v0 > 0) __syncthreads();
    // the particle of each of this thread's slots of the round (from the row
    // table: the permutation loads are in flight while the slot table is
    // built)
    int pp[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int v = v0 + t + j * NT;
      const int s = v < vend ? slot_of_virtual(v, rowOf, kARow32, segA, segP) : 0;
      pp[j] = ident ? s : perm[s];
    }
    fill_slot_cells<CC>(tab, cr_, ccol, cv0, cv1, v0, vend, kTs);
    if constexpr (USL) {
      // (slot q's particle shares LDS with slot q's element of the last
      // weight row, which its thread writes after reading the particle)
      static_assert(sizeof(T) % sizeof(int) == 0, "slot particle in a weight element");
      constexpr int kSp = sizeof(T) / sizeof(int);
      int* const spp = reinterpret_cast<int*>(st + (1 + NW) * CAP);
#pragma unroll
      for (int j = 0; j < RPT; ++j) spp[(t + j * NT) * kSp] = pp[j];
      __syncthreads();
      // one slot at a time (the G2P and stress adjoint of several slots
      // held together spill registers)
#pragma unroll 1
      for (int j = 0; j < RPT; ++j) {
        const int q = t + j * NT;
        if (v0 + q >= vend) break;
        const int code = tab[q * kTs];
        const int r = code >> 8, ci = code & 255;
        const int p = spp[q * kSp];
        // (both entries are read before this thread's weight stores reuse them)
        asm volatile("" ::: "memory");
        const bool mine = r >= oj && ci >= oi;
        const PGeo<T> gq = tile_geo(R.x[p], R.y[p], ci, r, wV, colC, colS, rowC, rowS);
        T vx[4], vy[4], ax[4], ay[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          vx[k] = nvv[gq.ids[k]];
          vy[k] = nvv[kAVN + gq.ids[k]];
          ax[k] = nvv[2 * kAVN + gq.ids[k]];
          ay[k] = nvv[3 * kAVN + gq.ids[k]];
        }
        const T sb[3] = {CR(DMPM_C_SXX)[p], CR(DMPM_C_SYY)[p], CR(DMPM_C_SXY)[p]};
        T w[10], Vb_in = T(0), sxi, seta, dummy = T(0);
        adj_g2p_core<T, true>(c.mat, a.mat.per_E ? R.E[p] : T(0), R.V[p], a.dt, a.vu, gq,
                              CR(DMPM_C_X)[p], CR(DMPM_C_X + 1)[p], CR(DMPM_C_V)[p],
                              CR(DMPM_C_V + 1)[p], vx, vy, ax, ay, sb, CR(DMPM_C_VOL)[p],
                              (mine && a.lebar) ? a.lebar + p : &dummy, Vb_in, sxi, seta, w);
        if (mine) {
          if (boundary_cell(RC.i0 + ci, RC.j0 + r)) WR(W_BVOL)[p] = Vb_in;
          else CR(DMPM_C_VOL)[p] = Vb_in;
          WR(W_XI)[p] = sxi;
          WR(W_ETA)[p] = seta;
        }
        st[q] = gq.xi;
        st[CAP + q] = gq.eta;
#pragma unroll
        for (int r2 = 0; r2 < NW; ++r2) st[(2 + r2) * CAP + q] = w[W_NVX + r2];
      }
    } else {
      // the slots' fields, issued before the barrier
      T x[RPT], y[RPT], xbx[RPT], xby[RPT], vbx[RPT], vby[RPT];
      T sbx[RPT], sby[RPT], sbq[RPT], vb[RPT], Ep[RPT], Vin[RPT];
  #pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int p = pp[j];
        x[j] = R.x[p];
        y[j] = R.y[p];
        xbx[j] = CR(DMPM_C_X)[p];
        xby[j] = CR(DMPM_C_X + 1)[p];
        vbx[j] = CR(DMPM_C_V)[p];
        vby[j] = CR(DMPM_C_V + 1)[p];
        if (USL) {
          sbx[j] = CR(DMPM_C_SXX)[p];
          sby[j] = CR(DMPM_C_SYY)[p];
          sbq[j] = CR(DMPM_C_SXY)[p];
          vb[j] = CR(DMPM_C_VOL)[p];
          Vin[j] = R.V[p];
          Ep[j] = a.mat.per_E ? R.E[p] : T(0);
        }
      }
      __syncthreads();
      int code[RPT];
      unsigned own = 0;
  #pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int q = t + j * NT;
        const bool live = v0 + q < vend;
        code[j] = live ? (int)tab[q * kTs] : 0;
        const int r = code[j] >> 8, ci = code[j] & 255;
        if (live && r >= oj && ci >= oi) own |= 1u << j;
      }
  #pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int q = t + j * NT;
        if (v0 + q >= vend) continue;
        const int p = pp[j];
        const PGeo<T> gq = tile_geo(x[j], y[j], code[j] & 255, code[j] >> 8, wV, colC, colS,
                                    rowC, rowS);
        T vx[4], vy[4], ax[4], ay[4];
  #pragma unroll
        for (int k = 0; k < 4; ++k) {
          vx[k] = nvv[gq.ids[k]];
          vy[k] = nvv[kAVN + gq.ids[k]];
          ax[k] = nvv[2 * kAVN + gq.ids[k]];
          ay[k] = nvv[3 * kAVN + gq.ids[k]];
        }
        const bool mine = own & (1u << j);
        T sb[3] = {T(0), T(0), T(0)}, Vb = T(0);
        if (USL) {
          sb[0] = sbx[j];
          sb[1] = sby[j];
          sb[2] = sbq[j];
          Vb = vb[j];
        }
        T w[10], Vb_in = T(0), sxi, seta, dummy = T(0);
        adj_g2p_core<T, USL>(c.mat, USL ? Ep[j] : T(0), USL ? Vin[j] : T(0), a.dt, a.vu, gq,
                             xbx[j], xby[j], vbx[j], vby[j], vx, vy, ax, ay, sb, Vb,
                             (mine && a.lebar) ? a.lebar + p : &dummy, Vb_in, sxi, seta, w);
        if (mine) {
          if (USL) {
            if (boundary_cell(RC.i0 + gq.I, RC.j0 + gq.J)) WR(W_BVOL)[p] = Vb_in;
            else CR(DMPM_C_VOL)[p] = Vb_in;
          }
          WR(W_XI)[p] = sxi;
          WR(W_ETA)[p] = seta;
        }
        st[q] = gq.xi;
        st[CAP + q] = gq.eta;
  #pragma unroll
        for (int r = 0; r < NW; ++r) st[(2 + r) * CAP + q] = w[W_NVX + r];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CC; ++u) {
      if (cr_[u] < 0) continue;
      const int lo = max(cv0[u], v0) - v0, hi = min(cv1[u], vend) - v0;
      for (int qq = lo; qq < hi; ++qq) {
        const T xi = st[qq], eta = st[CAP + qq];
        T N[4];
        q4_N4(xi, eta, N);
#define WQ(r) st[(2 + (r)) * CAP + qq]
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (MODE == 1) {
            acc[u][4 * k + 0] = acc[u][4 * k + 0] + N[k] * WQ(W_NVX);
            acc[u][4 * k + 1] = acc[u][4 * k + 1] + N[k] * WQ(W_NVY);
            acc[u][4 * k + 2] = acc[u][4 * k + 2] + N[k] * WQ(W_NAX);
            acc[u][4 * k + 3] = acc[u][4 * k + 3] + N[k] * WQ(W_NAY);
            continue;
          }
          T dx, dy;
          q4_dNdx(k, xi, eta, cix[u], ciy[u], dx, dy);
          const T q0 =
              ((N[k] * WQ(W_NVX) + dx * WQ(W_L00)) + dy * WQ(W_L10)) + cx[u][k] * WQ(W_C00);
          const T q1 =
              ((N[k] * WQ(W_NVY) + dy * WQ(W_L11)) + dx * WQ(W_L01)) + cy[u][k] * WQ(W_C11);
          acc[u][4 * k + 0] = acc[u][4 * k + 0] + q0;
          acc[u][4 * k + 1] = acc[u][4 * k + 1] + q1;
          acc[u][4 * k + 2] = acc[u][4 * k + 2] + N[k] * WQ(W_NAX);
          acc[u][4 * k + 3] = acc[u][4 * k + 3] + N[k] * WQ(W_NAY);
        }
#undef WQ
      }
    }
  }
  __syncthreads();
  T nacc[NPT][NQ];
  node_sums_nt<T, NQ, CC, NPT, NT>(st, acc, RC, NR, nx, ny, nacc);
  // node pass 1 of the owned nodes: adjoint of the nodal update
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    if (t + w * NT >= nn) continue;
    const int nd = ndx[w];
    const T m = nm[w];
    T mb = T(0);
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      T Fb = T(0), vkb = T(0);
      if (!(mask[w] & (1 << d))) {
        const T r = nF[w][d] / m;
        const T v1b = nv1[w][d] != T(0) ? nacc[w][d] : T(0);
        const T a0b = (na1[w][d] != T(0) ? nacc[w][2 + d] : T(0)) + a.dt * v1b;
        if (isfinite(r)) {
          Fb = a0b / m;
          mb = mb - a0b * r / m;
        }
        vkb = v1b;
      }
      na[(int64_t)(A_FX + d) * nnodes + nd] = Fb;
      if (USL) {
        T Pb;
        kin_adjoint(m, vk1[w][d], (mask[w] >> d) & 1, vkb, Pb, mb);
        na[(int64_t)(A_PX + d) * nnodes + nd] = Pb;
      } else {
        na[(int64_t)(A_VKX + d) * nnodes + nd] = vkb;
      }
    }
    na[(int64_t)A_M * nnodes + nd] = mb;
  }
}

template <typename T, bool B, int NT = AdjTileShape<2>::NT, int RPT = AdjTileShape<2>::RPT,
          int WPE = AdjTileShape<2>::WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_adj_tile2(
    AdjTile<T> A) {
  constexpr int CAP = NT * RPT;
  constexpr int NW = 5;  // strain-rate cotangents L00, L11, L01 (= L10), C00, C11
  constexpr int NQ = 2;
  constexpr int CC = (kACells + NT - 1) / NT;
  constexpr int NPT = (kANodes + NT - 1) / NT;
  static_assert(4 * NQ * kACells <= (2 + NW) * CAP, "node-sum slices");
  __shared__ T st[(2 + NW) * CAP];
  __shared__ uint16_t tab[CAP];
  __shared__ T nvv[4 * kAVN];  // force cotangents fx, fy and kinematic velocity vx, vy
  __shared__ T colC[kStepTile + 1], colS[kStepTile + 1], rowC[kStepTile + 1],
      rowS[kStepTile + 1];
  __shared__ int segA[kStepTile + 1], segP[kStepTile + 2], rowOf[kARow32];
  // slot q's particle shares LDS with slot q's element of the last staged
  // row, which its thread writes after reading the particle
  static_assert(sizeof(T) % sizeof(int) == 0, "slot particle in a staged element");
  constexpr int kSp = sizeof(T) / sizeof(int);
  int* const spp = reinterpret_cast<int*>(st + (1 + NW) * CAP);
  const AdjArgs<T>& a = A.a;
  const PRows<T>& R = a.R;
  const POut<T>& O = a.O;
  const int64_t ld = a.ld, nld = a.nld;
  const T* __restrict__ nb = a.nb;
  T* __restrict__ cot = a.cot;
  T* __restrict__ wk = a.wk;
  T* __restrict__ na = const_cast<T*>(a.na);
  const int* __restrict__ cs = A.sr.cs();
  const int* __restrict__ perm = A.sr.perm();
  const int ident = A.sr.flags[1];
  const int t = threadIdx.x;
  const int ntiles = A.tt.ntiles;
  const TileCtx<T> c = tile_ctx_of<T, B>(A.tt, xcd_pos(blockIdx.x, ntiles), true);
  const Grid2<T> g = c.g;
  const int nx = g.nx, ny = g.ny;
  const Rgn R0{c.ti * kStepTile, min(c.ti * kStepTile + kStepTile, nx), c.tj * kStepTile,
               min(c.tj * kStepTile + kStepTile, ny)};
  const Rgn RC{max(R0.i0 - 1, 0), R0.i1, max(R0.j0 - 1, 0), R0.j1};
  const Rgn NR = owned_nodes(R0, nx, ny);
  const int hC = RC.j1 - RC.j0, wC = RC.i1 - RC.i0;
  const int wV = wC + 1, nV = wV * (hC + 1);
  const int oi = R0.i0 - RC.i0, oj = R0.j0 - RC.j0;
  const int nnodes = A.nn_all;
  int cr_[CC], ccol[CC], cs0[CC], cs1[CC];
  cell_slots_rc<CC, NT>(RC, nx, c.coff, cs, cr_, ccol, cs0, cs1);
  if (t < 64) row_segments(RC, nx, c.coff, cs, segA, segP, rowOf, kARow32);
  else if (t < 128) cell_tables(g, RC, colC, colS, rowC, rowS);
  {
    const SmallDiv dwv = small_div(wV);
    for (int ln = t; ln < nV; ln += NT) {
      const int jr = sdiv(dwv, ln);
      const int nd = c.noff + (RC.j0 + jr) * (nx + 1) + RC.i0 + (ln - jr * wV);
      nvv[ln] = na[(int64_t)A_FX * nnodes + nd];
      nvv[kAVN + ln] = na[(int64_t)A_FY * nnodes + nd];
      nvv[2 * kAVN + ln] = a.vkin[nd];
      nvv[3 * kAVN + ln] = a.vkin[(int64_t)nnodes + nd];
    }
  }
  const int wn = NR.i1 - NR.i0, nn = wn * (NR.j1 - NR.j0);
  const SmallDiv dwn = small_div(wn);
  int ndx[NPT], mask[NPT];
  T nm[NPT], nmb[NPT], nvb[NPT][2], vk1[NPT][2];
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    const int ln = min(t + w * NT, nn - 1);
    const int jr = sdiv(dwn, ln);
    const int nd = c.noff + (NR.j0 + jr) * (nx + 1) + NR.i0 + (ln - jr * wn);
    ndx[w] = nd;
    const int cr = A.cidx ? A.cidx[nd] : -1;
    mask[w] = cr >= 0 ? A.cmask[cr] : 0;
    nm[w] = NR(DMPM_N_MASS)[nd];
    nmb[w] = na[(int64_t)A_M * nnodes + nd];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      nvb[w][d] = na[(int64_t)(A_VKX + d) * nnodes + nd];
      vk1[w][d] = a.vkin[(int64_t)d * nnodes + nd];
    }
  }
  __syncthreads();
  const int S = segP[hC];
  int cv0[CC], cv1[CC];
  cell_ranges_of<CC>(cr_, cs0, cs1, segA, segP, cv0, cv1);
  T cix[CC], ciy[CC], cx[CC][4], cy[CC][4];
#pragma unroll
  for (int u = 0; u < CC; ++u) {
    cix[u] = ciy[u] = T(0);
    for (int k = 0; k < 4; ++k) cx[u][k] = cy[u][k] = T(0);
    if (cr_[u] >= 0) {
      cix[u] = colS[ccol[u]];
      ciy[u] = rowS[cr_[u]];
      for (int k = 0; k < 4; ++k) q4_dNdx(k, T(0), T(0), cix[u], ciy[u], cx[u][k], cy[u][k]);
    }
  }
  T acc[CC][4 * NQ];
#pragma unroll
  for (int u = 0; u < CC; ++u)
#pragma unroll
    for (int v = 0; v < 4 * NQ; ++v) acc[u][v] = T(0);
  for (int v0 = 0; v0 < S; v0 += CAP) {
    const int vend = min(v0 + CAP, S);
    if (v0 > 0) __syncthreads();
    // the particle of each of this thread's slots of the round (from the row
    // table: the permutation loads are in flight while the slot table is
    // built), kept in LDS for the loop below
    {
      int pp[RPT];
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int v = v0 + t + j * NT;
        const int s = v < vend ? slot_of_virtual(v, rowOf, kARow32, segA, segP) : 0;
        pp[j] = ident ? s : perm[s];
      }
      fill_slot_cells<CC>(tab, cr_, ccol, cv0, cv1, v0, vend);
#pragma unroll
      for (int j = 0; j < RPT; ++j) spp[(t + j * NT) * kSp] = pp[j];
    }
    __syncthreads();
    // one slot at a time (the force and stress adjoint of several slots held
    // together spill registers)
#pragma unroll 1
    for (int j = 0; j < RPT; ++j) {
      const int q = t + j * NT;
      if (v0 + q >= vend) break;
      const int code = tab[q];
      const int r = code >> 8, ci = code & 255;
      const int p = spp[q * kSp];
      // (the particle is read before this thread's staged stores reuse it)
      asm volatile("" ::: "memory");
      const bool mine = r >= oj && ci >= oi;
      const PGeo<T> gq = tile_geo(R.x[p], R.y[p], ci, r, wV, colC, colS, rowC, rowS);
      T fx[4], fy[4], vkx[4], vky[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        fx[k] = nvv[gq.ids[k]];
        fy[k] = nvv[kAVN + gq.ids[k]];
        vkx[k] = nvv[2 * kAVN + gq.ids[k]];
        vky[k] = nvv[3 * kAVN + gq.ids[k]];
      }
      const T csv[3] = {CR(DMPM_C_SXX)[p], CR(DMPM_C_SYY)[p], CR(DMPM_C_SXY)[p]};
      T sb[3], vol, w[6], sxi, seta, dummy = T(0);
      adj_force_core<T, true>(c.mat, a.mat.per_E ? R.E[p] : T(0), a.dt, R.V[p], O.V[p],
                              O.sxx[p], O.syy[p], O.sxy[p], R.m[p], gq, a.Gx, a.Gy, fx, fy,
                              vkx, vky, csv, CR(DMPM_C_VOL)[p],
                              (mine && a.lebar) ? a.lebar + p : &dummy, sb, vol, w, sxi, seta);
      if (mine) {
        if (boundary_cell(RC.i0 + ci, RC.j0 + r)) {
          WR(W_BSXX)[p] = sb[0];
          WR(W_BSYY)[p] = sb[1];
          WR(W_BSXY)[p] = sb[2];
          WR(W_BVOL)[p] = vol;
        } else {
          CR(DMPM_C_SXX)[p] = sb[0];
          CR(DMPM_C_SYY)[p] = sb[1];
          CR(DMPM_C_SXY)[p] = sb[2];
          CR(DMPM_C_VOL)[p] = vol;
        }
        WR(W_XI)[p] = WR(W_XI)[p] + sxi;
        WR(W_ETA)[p] = WR(W_ETA)[p] + seta;
      }
      // the strain-rate cotangents (L01 and L10 are the same number)
      st[q] = gq.xi;
      st[CAP + q] = gq.eta;
      st[2 * CAP + q] = w[W_L00 - W_L00];
      st[3 * CAP + q] = w[W_L11 - W_L00];
      st[4 * CAP + q] = w[W_L01 - W_L00];
      st[5 * CAP + q] = w[W_C00 - W_L00];
      st[6 * CAP + q] = w[W_C11 - W_L00];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CC; ++u) {
      if (cr_[u] < 0) continue;
      const int lo = max(cv0[u], v0) - v0, hi = min(cv1[u], vend) - v0;
      for (int qq = lo; qq < hi; ++qq) {
        const T xi = st[qq], eta = st[CAP + qq];
        const T l00 = st[2 * CAP + qq], l11 = st[3 * CAP + qq], l01 = st[4 * CAP + qq];
        const T c00 = st[5 * CAP + qq], c11 = st[6 * CAP + qq];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          T dx, dy;
          q4_dNdx(k, xi, eta, cix[u], ciy[u], dx, dy);
          acc[u][2 * k + 0] = acc[u][2 * k + 0] + ((dx * l00 + dy * l01) + cx[u][k] * c00);
          acc[u][2 * k + 1] = acc[u][2 * k + 1] + ((dy * l11 + dx * l01) + cy[u][k] * c11);
        }
      }
    }
  }
  __syncthreads();
  T nacc[NPT][NQ];
  node_sums_nt<T, NQ, CC, NPT, NT>(st, acc, RC, NR, nx, ny, nacc);
  // node pass 2 of the owned nodes: kinematics adjoint
#pragma unroll
  for (int w = 0; w < NPT; ++w) {
    if (t + w * NT >= nn) continue;
    const int nd = ndx[w];
    T mb = nmb[w];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const T vb = nvb[w][d] + nacc[w][d];
      T Pb;
      kin_adjoint(nm[w], vk1[w][d], (mask[w] >> d) & 1, vb, Pb, mb);
      na[(int64_t)(A_PX + d) * nnodes + nd] = Pb;
    }
    na[(int64_t)A_M * nnodes + nd] = mb;
  }
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
  T* wk = (T*)plan_workspace(P, 3, (size_t)W_ROWS * ld * sizeof(T));
  T* na = (T*)plan_workspace(P, 4, (size_t)A_ROWS * P->nnodes * sizeof(T));
  if (!vkin || !wk || !na) return fail(DMPM_EHIP, "workspace allocation failed");
  Mat2 mat;
  mat.kind = P->d.material;
  mat.per_E = P->d.per_particle_E;
  for (int k = 0; k < 3; ++k) mat.de[k] = P->d.de[k];
  const T tdt = T(dt), Gx = T(grav[0]), Gy = T(grav[1]);
  const unsigned np = nblk(n);
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
  aa.ncb = 0;
  AdjTile<T> at;
  at.a = aa;
  at.tt.sims = P->sims;
  at.tt.tile_sim = P->tile_sim;
  at.tt.tiles_x = P->sims_h[0].tiles_x;
  at.tt.ntiles = P->ntiles;
  at.tt.g = M.g0;
  at.tt.mat = mat;
  at.sr = sr;
  at.cidx = P->cidx;
  at.cmask = P->cmask;
  at.nn_all = P->nnodes;
  at.nc_all = P->ncells;
  const unsigned nt = (unsigned)P->ntiles;
  const bool batch = P->nsims > 1;
  if (usf) {
    constexpr int NT1 = AdjTileShape<1>::NT, NT2 = AdjTileShape<2>::NT;
    if (batch) {
      k_adj_tile1<T, true, 1><<<nt + kAdjOutsideBlocks, NT1, 0, s>>>(at);
      k_adj_tile2<T, true><<<nt, NT2, 0, s>>>(at);
    } else {
      k_adj_tile1<T, false, 1><<<nt + kAdjOutsideBlocks, NT1, 0, s>>>(at);
      k_adj_tile2<T, false><<<nt, NT2, 0, s>>>(at);
    }
  } else {
    constexpr int NT0 = AdjTileShape<0>::NT;
    if (batch) k_adj_tile1<T, true, 0><<<nt + kAdjOutsideBlocks, NT0, 0, s>>>(at);
    else k_adj_tile1<T, false, 0><<<nt + kAdjOutsideBlocks, NT0, 0, s>>>(at);
    k_adj_force<T><<<np, kThreads, 0, s>>>(M, R, ld, n, na, vkin, cot, wk, lebar, mat, Gx, Gy,
                                          tdt);
  }
  k_adj_kin<T><<<np, kThreads, 0, s>>>(M, R, ld, n, na, cot, wk, vu, usf ? 1 : 0);
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

// The same copy in 16-byte words (every pointer 16-byte aligned): the
// permutation and the cell offsets of the current sort buffer.
__global__ void k_export_sort4(const int* __restrict__ perm0, const int* __restrict__ perm1,
                               const int* __restrict__ cs0, const int* __restrict__ cs1,
                               const int* __restrict__ flags, int n, int ncs,
                               int* __restrict__ perm_out, int* __restrict__ cs_out) {
  const int cur = flags[4];
  const int* __restrict__ perm = cur ? perm1 : perm0;
  const int* __restrict__ cs = cur ? cs1 : cs0;
  const int gs = gridDim.x * blockDim.x, gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int n4 = n >> 2, c4 = ncs >> 2;
  for (int i = gt; i < n4; i += gs)
    reinterpret_cast<int4*>(perm_out)[i] = reinterpret_cast<const int4*>(perm)[i];
  for (int i = gt; i < c4; i += gs)
    reinterpret_cast<int4*>(cs_out)[i] = reinterpret_cast<const int4*>(cs)[i];
  for (int i = (n4 << 2) + gt; i < n; i += gs) perm_out[i] = perm[i];
  for (int i = (c4 << 2) + gt; i < ncs; i += gs) cs_out[i] = cs[i];
}

#undef PR
#undef PW
#undef NR
#undef CR
#undef WR
#undef AR

}  // namespace

namespace dmpm {

// Copies the sort by cell of the plan's last step (permutation, cell offsets
// of the current sort buffer) to perm_out [n] and cs_out [ncells + 2].
int export_sort(dmpm_plan* P, int32_t* perm_out, int32_t* cs_out, hipStream_t s) {
  const int n = (int)P->n, ncs = P->ncells + 2;
  const bool a16 = ((uintptr_t)perm_out % 16 == 0) && ((uintptr_t)cs_out % 16 == 0) &&
                   ((uintptr_t)P->perm % 16 == 0) && ((uintptr_t)P->perm2 % 16 == 0) &&
                   ((uintptr_t)P->cs % 16 == 0) && ((uintptr_t)P->cs2 % 16 == 0);
  if (a16) {
    const int words = std::max(n, ncs) / 4 + 1;
    const unsigned blocks = (unsigned)std::min((words + kThreads - 1) / kThreads, 4096);
    k_export_sort4<<<blocks, kThreads, 0, s>>>(P->perm, P->perm2, P->cs, P->cs2, P->flags, n,
                                               ncs, perm_out, cs_out);
  } else {
    k_export_sort<<<1024, kThreads, 0, s>>>(P->perm, P->perm2, P->cs, P->cs2, P->flags, n, ncs,
                                            perm_out, cs_out);
  }
  return launch_check("export_sort");
}

}  // namespace dmpm

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
  return export_sort(P, perm_out, cs_out, s);
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
