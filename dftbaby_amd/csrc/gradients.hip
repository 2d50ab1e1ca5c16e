// Nuclear-derivative contractions of the LC-TD-DFTB gradient.
//
// The reference (DFTB/ExcGradients.py:286-711) forms the derivative tensors
// dS/dR, dH0/dR (slako.f90:402-569) and d(gamma)/dR, d(gamma_lr)/dR
// (GammaApproximation.py:419-443) as [3 Nat][Norb][Norb] arrays and contracts
// them with dense matrices.  Every such contraction is linear in the tensor, so
// the host folds all of them into four Norb x Norb matrices (A_S, A_H against
// dS/dR and dH0/dR, B_g, B_lr against the two gamma derivatives, formed with
// GEMMs) and this kernel forms
//   grad[3C + c] = <A_S, dS/dR_Cc> + <A_H, dH0/dR_Cc> + <B_g, dgamma_AO/dR_Cc>
//                + <B_lr, dgamma_lr_AO/dR_Cc>
// with the derivative blocks evaluated on the fly from the Slater-Koster
// splines, never stored.  One workgroup per atom C; its threads walk the other
// atoms D, each pair contributes
//   * SK part: sum over mu on D, nu on C of (A[mu][nu] + A[nu][mu]) times the
//     derivative of the (mu, nu) element with respect to R_C, formed as
//     slako.f90's visit (mu, nu) forms it (lighter type first, values from
//     splev_uniform, derivatives from splder);
//   * gamma part: g'(R_CD) (R_C - R_D) / R_CD times the block sum over mu on C,
//     nu on D of (B[mu][nu] + B[nu][mu]), for gamma and gamma_lr (Gaussian
//     charge fluctuations, GammaApproximation.py:141-147, 234-238);
// and a block reduction writes the three components once (deterministic).
#include "common.h"
#include "dftbaby_amd.h"
#include "slako_device.h"

namespace {
using namespace slako_dev;

__device__ __forceinline__ double gauss_gamma_deriv(double c, double R) {
  return 2.0 * c / sqrt(M_PI) * exp(-(c * R) * (c * R)) / R - erf(c * R) / (R * R);
}

__device__ __forceinline__ double block_sym_sum(const double* __restrict__ B, int ld, int o1, int n1, int o2, int n2) {
  double s = 0.0;
  for (int a = 0; a < n1; ++a)
    for (int b = 0; b < n2; ++b)
      s += B[(size_t)(o1 + a) * ld + o2 + b] + B[(size_t)(o2 + b) * ld + o1 + a];
  return s;
}

__global__ void __launch_bounds__(256) grad_pairs_kernel(
    const double* __restrict__ pos, const int* __restrict__ atom_type, const int* __restrict__ atom_off,
    const int* __restrict__ orb_l, const int* __restrict__ orb_m, const double* __restrict__ sigma, double rlr,
    const double* __restrict__ knots, const double* __restrict__ coefs, const double* __restrict__ dcoefs, int npair,
    int npts, double cutoff, const double* __restrict__ AS, const double* __restrict__ AH,
    const double* __restrict__ BG, const double* __restrict__ BL, int nat, int norb, double* __restrict__ grad) {
  const int C = blockIdx.x;
  const int oC = atom_off[C], nC = atom_off[C + 1] - oC;
  const int tC = atom_type[C];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int D = threadIdx.x; D < nat; D += blockDim.x) {
    if (D == C) continue;
    const int oD = atom_off[D], nD = atom_off[D + 1] - oD;
    // ---- gamma, gamma_lr: d/dR_C of gamma(|R_C - R_D|)
    const double rx = pos[3 * C] - pos[3 * D], ry = pos[3 * C + 1] - pos[3 * D + 1],
                 rz = pos[3 * C + 2] - pos[3 * D + 2];
    const double R = sqrt(rx * rx + ry * ry + rz * rz);
    if (BG || BL) {
      const double s2 = sigma[C] * sigma[C] + sigma[D] * sigma[D];
      double f = 0.0;
      if (BG) f += gauss_gamma_deriv(1.0 / sqrt(2.0 * s2), R) * block_sym_sum(BG, norb, oC, nC, oD, nD);
      if (BL) f += gauss_gamma_deriv(1.0 / sqrt(2.0 * (s2 + 0.5 * rlr * rlr)), R) *
                   block_sym_sum(BL, norb, oC, nC, oD, nD);
      acc[0] += f * (rx / R);
      acc[1] += f * (ry / R);
      acc[2] += f * (rz / R);
    }
    if (!AS && !AH) continue;
    // ---- Slater-Koster part: elements (mu on D, nu on C), visit (mu, nu)
    const double xc = pos[3 * C] - pos[3 * D], yc = pos[3 * C + 1] - pos[3 * D + 1], zc = pos[3 * C + 2] - pos[3 * D + 2];
    const double r = sqrt(xc * xc + yc * yc + zc * zc);
    if (!(r < cutoff)) continue;
    const int tD = atom_type[D];
    const bool swap = tD > tC;
    const int tmin = swap ? tC : tD, tmax = swap ? tD : tC;
    const int ip = tmax * (tmax + 1) / 2 + tmin;
    if (ip >= npair) continue;
    double e[3] = {xc / r, yc / r, zc / r};            // D -> C
    double em[3] = {-e[0], -e[1], -e[2]};
    const double* t = knots + (size_t)ip * npts;
    const double* c = coefs + (size_t)ip * SLAKO_NTAB * npts;
    const double* w = dcoefs + (size_t)ip * 14 * npts;
    const Basis b = spline_basis(t, npts, r);
    double Ts[7], Th[7], Tsd[7], Thd[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const bool used = (k == 0 || k == 2 || k == 4 || k == 5 || k == 6);
      Ts[k] = used ? spline_value(c + (size_t)k * npts, b) : 0.0;
      Th[k] = used ? spline_value(c + (size_t)(7 + k) * npts, b) : 0.0;
      Tsd[k] = used ? spline_deriv(t, npts, w + (size_t)k * npts, r) : 0.0;
      Thd[k] = used ? spline_deriv(t, npts, w + (size_t)(7 + k) * npts, r) : 0.0;
    }
    for (int a = 0; a < nD; ++a) {
      const int mu = oD + a;
      for (int bb = 0; bb < nC; ++bb) {
        const int nu = oC + bb;
        double gs[3], gh[3];
        if (!swap) {
          sk_rule_grad(orb_l[mu], orb_m[mu], orb_l[nu], orb_m[nu], e, r, Ts, Tsd, gs);
          sk_rule_grad(orb_l[mu], orb_m[mu], orb_l[nu], orb_m[nu], e, r, Th, Thd, gh);
        } else {
          sk_rule_grad(orb_l[nu], orb_m[nu], orb_l[mu], orb_m[mu], em, r, Ts, Tsd, gs);
          sk_rule_grad(orb_l[nu], orb_m[nu], orb_l[mu], orb_m[mu], em, r, Th, Thd, gh);
#pragma unroll
          for (int q = 0; q < 3; ++q) { gs[q] = -gs[q]; gh[q] = -gh[q]; }
        }
        const double ws = AS ? AS[(size_t)mu * norb + nu] + AS[(size_t)nu * norb + mu] : 0.0;
        const double wh = AH ? AH[(size_t)mu * norb + nu] + AH[(size_t)nu * norb + mu] : 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] += ws * gs[q] + wh * gh[q];
      }
    }
  }
  // block reduction (fixed order: waves, then the wave partials in index order)
  __shared__ double part[3][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) part[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[q][k];
    grad[3 * C + q] = s;
  }
}

// Gradient of the repulsive energy (ExcGradients.py:181-210): grad[3C + c] =
// sum over D != C of Vrep'(R_CD) (R_C - R_D)_c / R_CD, Vrep' = 0 at and beyond
// the table end.  One workgroup per atom C, block reduction as above.
__global__ void __launch_bounds__(256) grad_vrep_kernel(const double* __restrict__ pos,
                                                        const int* __restrict__ atom_type,
                                                        const double* __restrict__ knots,
                                                        const double* __restrict__ dcoefs,
                                                        const double* __restrict__ dmax, int npair, int npts, int nat,
                                                        double* __restrict__ grad) {
  const int C = blockIdx.x;
  const int tC = atom_type[C];
  double acc[3] = {0.0, 0.0, 0.0};
  for (int D = threadIdx.x; D < nat; D += blockDim.x) {
    if (D == C) continue;
    const double rx = pos[3 * C] - pos[3 * D], ry = pos[3 * C + 1] - pos[3 * D + 1],
                 rz = pos[3 * C + 2] - pos[3 * D + 2];
    const double R = sqrt(rx * rx + ry * ry + rz * rz);
    const int tD = atom_type[D];
    const int tmin = tC < tD ? tC : tD, tmax = tC < tD ? tD : tC;
    const int ip = tmax * (tmax + 1) / 2 + tmin;
    if (ip >= npair || !(R < dmax[ip])) continue;
    const double dv = spline_deriv(knots + (size_t)ip * npts, npts, dcoefs + (size_t)ip * npts, R);
    acc[0] += dv * (rx / R);
    acc[1] += dv * (ry / R);
    acc[2] += dv * (rz / R);
  }
  __shared__ double part[3][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) part[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[q][k];
    grad[3 * C + q] = s;
  }
}

}  // namespace

extern "C" int dftb_grad_vrep(const double* pos, const int* atom_type, const double* knots, const double* dcoefs,
                              const double* dmax, int npair, int npts, int nat, double* grad, void* stream) {
  if (nat <= 0 || npair <= 0 || npts < 2 * (slako_dev::kDeg + 1) + 2) return DFTB_ERR_ARG;
  if (!pos || !atom_type || !knots || !dcoefs || !dmax || !grad) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  grad_vrep_kernel<<<nat, 256, 0, st>>>(pos, atom_type, knots, dcoefs, dmax, npair, npts, nat, grad);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_grad_pairs(const double* pos, const int* atom_type, const int* atom_off, const int* orb_l,
                               const int* orb_m, const double* sigma, double rlr, const double* knots,
                               const double* coefs, const double* dcoefs, int npair, int npts, double cutoff,
                               const double* AS, const double* AH, const double* BG, const double* BL, int nat,
                               int norb, double* grad, void* stream) {
  if (nat <= 0 || norb <= 0 || npair <= 0 || npts < 2 * (slako_dev::kDeg + 1) + 2) return DFTB_ERR_ARG;
  if (!pos || !atom_type || !atom_off || !orb_l || !orb_m || !grad) return DFTB_ERR_ARG;
  if ((AS || AH) && (!knots || !coefs || !dcoefs)) return DFTB_ERR_ARG;
  if ((BG || BL) && !sigma) return DFTB_ERR_ARG;
  if (!(cutoff > 0.0)) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  grad_pairs_kernel<<<nat, 256, 0, st>>>(pos, atom_type, atom_off, orb_l, orb_m, sigma, rlr, knots, coefs, dcoefs,
                                         npair, npts, cutoff, AS, AH, BG, BL, nat, norb, grad);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
