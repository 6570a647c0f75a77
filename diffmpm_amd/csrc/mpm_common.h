// Device helpers shared by the diffmpm_amd HIP sources (gfx950).
//
// All sources are compiled with -ffp-contract=on: a product added within the
// same source expression (a * b + c) is one fused multiply-add, and nothing is
// fused across statements.  Rounding is therefore fixed by the source text, so
// the per-method kernels, the fused step kernels and the adjoint recompute,
// which write each expression the same way, produce bit-identical values for
// the same inputs, independent of inlining; the operation order follows the
// reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <math.h>

#include "../../include/diffmpm_amd.h"

namespace dmpm {

constexpr int kBlock = 256;

template <typename T> struct Num;
template <> struct Num<float> {
  __host__ __device__ static constexpr float max_finite() { return FLT_MAX; }
};
template <> struct Num<double> {
  __host__ __device__ static constexpr double max_finite() { return DBL_MAX; }
};

// where(|x| < 1e-12, 0, x)  (element.py:185-189; the threshold rounds to T).
template <typename T> __device__ __forceinline__ T snap(T x) {
  return fabs(x) < T(1e-12) ? T(0) : x;
}

// jnp.nan_to_num (element.py:311).
template <typename T> __device__ __forceinline__ T nan_to_num(T x) {
  if (isnan(x)) return T(0);
  if (isinf(x)) return x > T(0) ? Num<T>::max_finite() : -Num<T>::max_finite();
  return x;
}

template <typename T> __device__ __forceinline__ T qnan() {
  return T(NAN);
}

// Uniform structured mesh: node coordinates fl(i*h) in T (element.py:667-680).
template <typename T> struct Geo {
  int nx, ny;        // cells per axis (ny = 1 in 1D)
  T hx, hy;
  T ihx, ihy;        // 1/h, only used to guess the element index
  int ncells;        // nx*ny
  int nnodes;        // (nx+1)*(ny+1) in 2D, nx+1 in 1D
  __device__ __forceinline__ T X(int i) const { return T(i) * hx; }
  __device__ __forceinline__ T Y(int j) const { return T(j) * hy; }
};

// max{i in [0,n] : fl(i*h) <= x}, or -1 when no node coordinate is <= x
// (also for NaN).  Node coordinates are nondecreasing in i.
template <typename T> __device__ __forceinline__ int last_le(T x, int n, T h, T ih) {
  if (!(x >= T(0))) return -1;
  T q = x * ih;  // a guess; the loops below make the result exact
  int g = (q >= T(n)) ? n : (int)q;
  while (g < n && T(g + 1) * h <= x) ++g;
  while (g >= 0 && T(g) * h > x) --g;
  return g;
}

// Element search of the reference (A1).
//  Q4 (element.py:820-840): lower_left = max{n : X_n <= x and Y_n <= y} or -1,
//  eid = lower_left - lower_left // (nx+1)  ==  J*nx + I, or 0 when empty.
//  1D (element.py:517-541): eid = I when X_I <= x < X_{I+1}, else -1.
template <typename T, int DIM>
__device__ __forceinline__ int element_id(const Geo<T>& g, T x, T y) {
  int I = last_le(x, g.nx, g.hx, g.ihx);
  if (DIM == 1) return (I >= 0 && I < g.nx) ? I : -1;
  int J = last_le(y, g.ny, g.hy, g.ihy);
  return (I >= 0 && J >= 0) ? J * g.nx + I : 0;
}

// particle.py:180-206, per axis: xi = (x - (X0 + X1)/2) * (2/(X1 - X0)).
template <typename T> __device__ __forceinline__ T natural(T x, T X0, T X1) {
  return (x - (X0 + X1) / T(2)) * (T(2) / (X1 - X0));
}

// Q4 shape functions (element.py:748-755), corner order ll, lr, ur, ul.
template <typename T> __device__ __forceinline__ T q4_N(int k, T xi, T eta) {
  switch (k) {
    case 0: return T(0.25) * (T(1) - xi) * (T(1) - eta);
    case 1: return T(0.25) * (T(1) + xi) * (T(1) - eta);
    case 2: return T(0.25) * (T(1) + xi) * (T(1) + eta);
    default: return T(0.25) * (T(1) - xi) * (T(1) + eta);
  }
}
template <typename T> __device__ __forceinline__ void q4_N4(T xi, T eta, T N[4]) {
  N[0] = T(0.25) * (T(1) - xi) * (T(1) - eta);
  N[1] = T(0.25) * (T(1) + xi) * (T(1) - eta);
  N[2] = T(0.25) * (T(1) + xi) * (T(1) + eta);
  N[3] = T(0.25) * (T(1) - xi) * (T(1) + eta);
}
// 1D (element.py:453): N = [0.5(1-xi), 0.5(1+xi)].
template <typename T> __device__ __forceinline__ T l1_N(int k, T xi) {
  return k == 0 ? T(0.5) * (T(1) - xi) : T(0.5) * (T(1) + xi);
}

// Natural gradients (element.py:781-788) of corner k.
template <typename T> __device__ __forceinline__ void q4_dNdxi(int k, T xi, T eta, T& gx, T& gy) {
  switch (k) {
    case 0: gx = T(-0.25) * (T(1) - eta); gy = T(-0.25) * (T(1) - xi); break;
    case 1: gx = T(0.25) * (T(1) - eta);  gy = T(-0.25) * (T(1) + xi); break;
    case 2: gx = T(0.25) * (T(1) + eta);  gy = T(0.25) * (T(1) + xi); break;
    default: gx = T(-0.25) * (T(1) + eta); gy = T(0.25) * (T(1) - xi); break;
  }
}

// Inverse Jacobian of an axis-aligned Q4 cell (element.py:814-817).  In
// exact arithmetic J = dN/dxi^T X_e = diag((X1-X0)/2, (Y1-Y0)/2) for any
// (xi, eta), so inv(J) = diag(2/(X1-X0), 2/(Y1-Y0)): the same scale as the
// natural coordinates (particle.py:203-205).  The reference forms J as a
// 4-term FP32 dot product of O(X) terms whose cancellation leaves up to
// ~1e-5 relative noise on fine meshes; the closed form is exact.
template <typename T>
__device__ __forceinline__ void q4_invJ(T xi, T eta, T X0, T X1, T Y0, T Y1, T& ix, T& iy) {
  (void)xi;
  (void)eta;
  ix = T(2) / (X1 - X0);
  iy = T(2) / (Y1 - Y0);
}

// Physical gradient of corner k (element.py:791-818).
template <typename T>
__device__ __forceinline__ void q4_dNdx(int k, T xi, T eta, T ix, T iy, T& dx, T& dy) {
  T gx, gy;
  q4_dNdxi(k, xi, eta, gx, gy);
  dx = gx * ix;
  dy = gy * iy;
}

// 1D physical gradient (element.py:488-515): inv(J) = 2/(X1 - X0).
template <typename T> __device__ __forceinline__ T l1_inv(T X0, T X1) {
  return T(2) / (X1 - X0);
}
template <typename T> __device__ __forceinline__ T l1_dNdx(int k, T inv) {
  return (k == 0 ? T(-0.5) : T(0.5)) * inv;
}

// jnp.interp (functions.py:40-41) with end values held outside [x0, xn].
template <typename T>
__device__ __forceinline__ T interp(T t, const double* xs, const double* fs, int n) {
  if (n <= 0) return T(0);
  if (n == 1) return T(fs[0]);
  int i = 1;
  while (i < n - 1 && T(xs[i]) <= t) ++i;  // searchsorted(side='right'), clipped
  T x0 = T(xs[i - 1]), x1 = T(xs[i]);
  T f0 = T(fs[i - 1]), f1 = T(fs[i]);
  T dx = x1 - x0;
  T f = (dx == T(0)) ? f0 : f0 + ((t - x0) / dx) * (f1 - f0);
  if (t < T(xs[0])) f = T(fs[0]);
  if (t > T(xs[n - 1])) f = T(fs[n - 1]);
  return f;
}

struct FnTable {
  const int* kind;
  const int* off;
  const int* cnt;
  const double* xs;
  const double* fs;
};

template <typename T>
__device__ __forceinline__ T fn_value(const FnTable& f, int k, T t) {
  if (k < 0 || f.kind[k] == DMPM_FN_UNIT) return T(1);
  return interp<T>(t, f.xs + f.off[k], f.fs + f.off[k], f.cnt[k]);
}

}  // namespace dmpm
