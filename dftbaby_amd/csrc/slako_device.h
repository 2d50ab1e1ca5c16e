// Device functions shared by the Slater-Koster kernels (slako.hip, gradients.hip):
// the B-spline evaluation of DFTB/extensions/splines.f90 (splev_uniform with its
// interval rule, FITPACK's splder for first derivatives) and the s/p rules of
// slako_transformations_lmax2.f90 written as rotation-covariant tensors.
#pragma once
#include "common.h"

namespace slako_dev {

constexpr int kDeg = 3;   // cubic B-splines (splrep's default degree)

// B-spline basis and knot interval for one distance, as splev_uniform/fpbspl.
struct Basis {
  double h[kDeg + 1];
  int c0;   // first coefficient index (0-based)
};

__device__ inline Basis spline_basis(const double* __restrict__ t, int n, double x) {
  const int k1 = kDeg + 1, nk1 = n - k1;
  const double tb = t[k1 - 1], te = t[nk1];
  double arg;
  int l;   // 1-based knot interval, as in splines.f90
  if (x <= tb) {
    arg = tb;
    l = k1;
  } else if (x >= te) {
    arg = te;
    l = nk1;
  } else {
    arg = x;
    const double dt = t[k1 + 1] - t[k1];
    l = (int)((x - t[0]) / dt) + 1;
  }
  // splines.f90 moves l < k to k+1; l == k (x within three grid steps of the
  // first point) would read before the coefficient array there.  Distinct atoms
  // are never that close; clamp so the read stays in bounds.
  if (l < k1) l = k1;
  if (l > nk1) l = nk1;
  Basis b;
  b.h[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= kDeg; ++j) {
    double hh[kDeg];
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = b.h[i];
    b.h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      const double f = hh[i - 1] / (t[li - 1] - t[lj - 1]);
      b.h[i - 1] = b.h[i - 1] + f * (t[li - 1] - arg);
      b.h[i] = f * (arg - t[lj - 1]);
    }
  }
  b.c0 = l - k1;
  return b;
}

__device__ __forceinline__ double spline_value(const double* __restrict__ c, const Basis& b) {
  double s = 0.0;
#pragma unroll
  for (int j = 0; j <= kDeg; ++j) s = s + c[b.c0 + j] * b.h[j];
  return s;
}

// real p orbital m = -1, 0, 1 -> Cartesian y, z, x
__device__ __forceinline__ int pidx(int m) { return m == -1 ? 1 : (m == 0 ? 2 : 0); }

__device__ __forceinline__ double sum_other_sq(const double* e, int i) {
  const double x2 = e[0] * e[0], y2 = e[1] * e[1], z2 = e[2] * e[2];
  return i == 0 ? y2 + z2 : (i == 1 ? x2 + z2 : x2 + y2);
}

// <first|second> from the table values T[0..6] (ss-sigma 0, sp-sigma 2,
// ps-sigma 4, pp-pi 5, pp-sigma 6), e: unit vector first -> second.
__device__ inline double sk_rule(int l1, int m1, int l2, int m2, const double* e, const double* T) {
  if (l1 == 0 && l2 == 0) return T[0];
  if (l1 == 0) return e[pidx(m2)] * T[2];
  if (l2 == 0) return e[pidx(m1)] * T[4];
  const int i = pidx(m1), j = pidx(m2);
  if (i == j) return sum_other_sq(e, i) * T[5] + e[i] * e[i] * T[6];
  return e[i] * e[j] * (T[6] - T[5]);
}

// <first|r_M|second> (without the origin shift) from the dipole tables D[0..14]
// (indices 1, 3, 4, 6, 7, 8, 13, 14 carry data for s and p).
__device__ inline double sk_rule_dipole(int l1, int m1, int M, int l2, int m2, const double* e, const double* D) {
  if (l1 == 0 && l2 == 0) return e[M] * D[3];
  if (l1 == 0) {
    const int j = pidx(m2);
    if (M == j) return sum_other_sq(e, j) * D[1] + e[j] * e[j] * D[4];
    return e[M] * e[j] * (D[4] - D[1]);
  }
  if (l2 == 0) {
    const int i = pidx(m1);
    if (M == i) return sum_other_sq(e, i) * D[6] + e[i] * e[i] * D[13];
    return e[i] * e[M] * (D[13] - D[6]);
  }
  const int i = pidx(m1), j = pidx(m2);
  double v = (D[14] - D[7] - 2.0 * D[8]) * e[i] * e[M] * e[j];
  if (i == M) v += D[7] * e[j];
  if (M == j) v += D[8] * e[i];
  if (i == j) v += D[8] * e[M];
  return v;
}

// The KD+1 non-zero B-splines of degree KD at x for the 1-based knot interval l
// (de Boor-Cox recurrence, as fpbspl).
template <int KD>
__device__ inline void bspline_basis(const double* __restrict__ t, int l, double x, double* h) {
  h[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= KD; ++j) {
    double hh[KD];
#pragma unroll
    for (int i = 0; i < j; ++i) hh[i] = h[i];
    h[0] = 0.0;
#pragma unroll
    for (int i = 1; i <= j; ++i) {
      const int li = l + i, lj = li - j;
      const double f = hh[i - 1] / (t[li - 1] - t[lj - 1]);
      h[i - 1] = h[i - 1] + f * (t[li - 1] - x);
      h[i] = f * (x - t[lj - 1]);
    }
  }
}

// First derivative as FITPACK's splder (splder.f, called through
// splines.f90:97-115): the degree-2 spline with coefficients w (computed once on
// the host, w_i = 3 (c_{i+1} - c_i) / (t_{i+4} - t_{i+1})) on the knots t[1..n-2],
// x clamped to the table, the knot interval found by search (not by the rule of
// splev_uniform).
__device__ inline double spline_deriv(const double* __restrict__ t, int n, const double* __restrict__ w, double x) {
  const int k1 = kDeg + 1, nk1 = n - k1;
  const double tb = t[k1 - 1], te = t[nk1];
  const double arg = x < tb ? tb : (x > te ? te : x);
  const double* t1 = t + 1;
  constexpr int k2 = kDeg - 1;
  const int lmin = k2 + 1, lmax = (n - 2) - k2 - 1;
  const double dt = t[k1 + 1] - t[k1];
  int l = lmin + (int)((arg - t1[lmin - 1]) / dt);
  if (l < lmin) l = lmin;
  if (l > lmax) l = lmax;
  while (l > lmin && arg < t1[l - 1]) --l;
  while (l < lmax && arg >= t1[l]) ++l;
  double h[k2 + 1];
  bspline_basis<k2>(t1, l, arg, h);
  double s = 0.0;
#pragma unroll
  for (int j = 0; j <= k2; ++j) s = s + w[l - k2 - 1 + j] * h[j];
  return s;
}

// Gradient of the S / H0 rule with respect to the second atom's position:
// g = (sum_k dV/dT_k T_k') e + (1 - e e^T) grad_e V / r, e: unit vector
// first -> second.
__device__ inline void sk_rule_grad(int l1, int m1, int l2, int m2, const double* e, double r, const double* T,
                                    const double* Td, double* g) {
  double dT = 0.0, ge[3] = {0.0, 0.0, 0.0};
  if (l1 == 0 && l2 == 0) {
    dT = Td[0];
  } else if (l1 == 0) {
    const int j = pidx(m2);
    dT = e[j] * Td[2];
    ge[j] = T[2];
  } else if (l2 == 0) {
    const int i = pidx(m1);
    dT = e[i] * Td[4];
    ge[i] = T[4];
  } else {
    const int i = pidx(m1), j = pidx(m2);
    if (i == j) {
      dT = sum_other_sq(e, i) * Td[5] + e[i] * e[i] * Td[6];
#pragma unroll
      for (int k = 0; k < 3; ++k) ge[k] = 2.0 * e[k] * (k == i ? T[6] : T[5]);
    } else {
      dT = e[i] * e[j] * (Td[6] - Td[5]);
      ge[i] += e[j] * (T[6] - T[5]);
      ge[j] += e[i] * (T[6] - T[5]);
    }
  }
  const double edg = e[0] * ge[0] + e[1] * ge[1] + e[2] * ge[2];
#pragma unroll
  for (int c = 0; c < 3; ++c) g[c] = dT * e[c] + (ge[c] - edg * e[c]) / r;
}

}  // namespace slako_dev
