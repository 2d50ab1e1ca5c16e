// Slater-Koster assembly of the overlap S, the reference hamiltonian H0 and the
// dipole matrix D between valence orbitals (s and p), for the ground-state
// producer.  Restates DFTB/extensions/slako.f90:77-394 (H0andS, DipoleMatrix)
// with the B-spline evaluation of splines.f90:25-94 (splev_uniform), including
// its interval rule l = INT((x - t(1)) / dt) + 1 and its clamping at the table
// ends (see dftb_slako_matrices in include/dftbaby_amd.h).
//
// One thread per output element (row, col): it forms the lower-triangle element
// (max, min) of the pair, as the reference does before mirroring, so the two
// triangles are bit-identical and every row is written by consecutive threads.
// The tables (one knot vector per atom-type pair, 29 coefficient rows: S, H0,
// dipole) are a few hundred kB and stay in L2; the kernel is bound by its HBM
// writes (40 bytes per element with D).
#include "common.h"
#include "dftbaby_amd.h"

#include "slako_device.h"

namespace {
using namespace slako_dev;

__global__ void __launch_bounds__(256) slako_kernel(const double* __restrict__ pos, const int* __restrict__ orb_atom,
                                                    const int* __restrict__ orb_l, const int* __restrict__ orb_m,
                                                    const int* __restrict__ atom_type,
                                                    const double* __restrict__ orben, int norb,
                                                    const double* __restrict__ knots,
                                                    const double* __restrict__ coefs, int npair, int npts,
                                                    double cutoff, double* __restrict__ S, double* __restrict__ H,
                                                    double* __restrict__ Dm) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (col >= norb) return;
  const int hi = row > col ? row : col, lo = row > col ? col : row;
  const int A = orb_atom[hi], B = orb_atom[lo];
  const size_t o = (size_t)row * norb + col;
  const size_t plane = (size_t)norb * norb;
  const double xc = pos[3 * B] - pos[3 * A], yc = pos[3 * B + 1] - pos[3 * A + 1], zc = pos[3 * B + 2] - pos[3 * A + 2];
  const double r = sqrt(xc * xc + yc * yc + zc * zc);
  double sv = 0.0, hv = 0.0, dv[3] = {0.0, 0.0, 0.0};
  if (A == B || r < cutoff) {
    const int tA = atom_type[A], tB = atom_type[B];
    const bool swap = tA > tB;     // the orbital on the lighter type goes first
    const int first = swap ? lo : hi, second = swap ? hi : lo;
    const int tmin = swap ? tB : tA, tmax = swap ? tA : tB;
    const int ip = tmax * (tmax + 1) / 2 + tmin;
    const int l1 = orb_l[first], m1 = orb_m[first], l2 = orb_l[second], m2 = orb_m[second];
    double e[3];
    double rr;
    if (A == B) {
      e[0] = 0.0; e[1] = 0.0; e[2] = 1.0;
      rr = 0.0;
    } else {
      e[0] = xc / r; e[1] = yc / r; e[2] = zc / r;
      if (swap) { e[0] = -e[0]; e[1] = -e[1]; e[2] = -e[2]; }
      rr = r;
    }
    if (ip < npair && l1 <= 1 && l2 <= 1) {
      const double* t = knots + (size_t)ip * npts;
      const double* c = coefs + (size_t)ip * SLAKO_NTAB * npts;
      const Basis b = spline_basis(t, npts, rr);
      if (A == B) {
        if (hi == lo) { sv = 1.0; hv = orben[hi]; }
      } else {
        double Ts[7], Th[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) { Ts[k] = 0.0; Th[k] = 0.0; }
        // only the tables the rule of (l1, l2) reads
        const int need0 = (l1 == 0 && l2 == 0) ? 0 : (l1 == 0 ? 2 : (l2 == 0 ? 4 : 5));
        const int need1 = (l1 == 1 && l2 == 1) ? 6 : need0;
        Ts[need0] = spline_value(c + (size_t)need0 * npts, b);
        Th[need0] = spline_value(c + (size_t)(7 + need0) * npts, b);
        Ts[need1] = spline_value(c + (size_t)need1 * npts, b);
        Th[need1] = spline_value(c + (size_t)(7 + need1) * npts, b);
        sv = sk_rule(l1, m1, l2, m2, e, Ts);
        hv = sk_rule(l1, m1, l2, m2, e, Th);
      }
      if (Dm) {
        double Dt[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) Dt[k] = 0.0;
        if (l1 == 0 && l2 == 0) {
          Dt[3] = spline_value(c + (size_t)(14 + 3) * npts, b);
        } else if (l1 == 0) {
          Dt[1] = spline_value(c + (size_t)(14 + 1) * npts, b);
          Dt[4] = spline_value(c + (size_t)(14 + 4) * npts, b);
        } else if (l2 == 0) {
          Dt[6] = spline_value(c + (size_t)(14 + 6) * npts, b);
          Dt[13] = spline_value(c + (size_t)(14 + 13) * npts, b);
        } else {
          Dt[7] = spline_value(c + (size_t)(14 + 7) * npts, b);
          Dt[8] = spline_value(c + (size_t)(14 + 8) * npts, b);
          Dt[14] = spline_value(c + (size_t)(14 + 14) * npts, b);
        }
        const int G = orb_atom[second];
#pragma unroll
        for (int M = 0; M < 3; ++M) dv[M] = sk_rule_dipole(l1, m1, M, l2, m2, e, Dt) + pos[3 * G + M] * sv;
      }
    } else {
      sv = hv = dv[0] = dv[1] = dv[2] = __longlong_as_double(0x7ff8000000000000LL);   // unsupported input
    }
  }
  S[o] = sv;
  if (H) H[o] = hv;
  if (Dm) {
    Dm[o] = dv[0];
    Dm[plane + o] = dv[1];
    Dm[2 * plane + o] = dv[2];
  }
}

// Overlap between the orbitals of two geometries of one molecule, as
// slako.f90:576-687 (overlap12) forms it: element (mu, nu), mu on atom i of
// geometry 1, nu on atom j of geometry 2, direction from R1_i to R2_j, the
// orbital on the lighter type first; no proximity cutoff and no mirroring.
__global__ void __launch_bounds__(256) overlap12_kernel(const double* __restrict__ pos1,
                                                        const double* __restrict__ pos2,
                                                        const int* __restrict__ orb_atom,
                                                        const int* __restrict__ orb_l, const int* __restrict__ orb_m,
                                                        const int* __restrict__ atom_type, int norb,
                                                        const double* __restrict__ knots,
                                                        const double* __restrict__ coefs, int npair, int npts,
                                                        double* __restrict__ S) {
  const int nu = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (nu >= norb) return;
  const int i = orb_atom[mu], j = orb_atom[nu];
  const double xc = pos2[3 * j] - pos1[3 * i], yc = pos2[3 * j + 1] - pos1[3 * i + 1],
               zc = pos2[3 * j + 2] - pos1[3 * i + 2];
  const double r = sqrt(xc * xc + yc * yc + zc * zc);
  double e[3];
  if (r > 0.0) {
    e[0] = xc / r; e[1] = yc / r; e[2] = zc / r;
  } else {
    e[0] = 0.0; e[1] = 0.0; e[2] = 1.0;
  }
  const int ti = atom_type[i], tj = atom_type[j];
  const bool swap = ti > tj;
  const int tmin = swap ? tj : ti, tmax = swap ? ti : tj;
  const int ip = tmax * (tmax + 1) / 2 + tmin;
  double v = __longlong_as_double(0x7ff8000000000000LL);
  const int first = swap ? nu : mu, second = swap ? mu : nu;
  const int l1 = orb_l[first], m1 = orb_m[first], l2 = orb_l[second], m2 = orb_m[second];
  if (ip < npair && l1 <= 1 && l2 <= 1) {
    if (swap) { e[0] = -e[0]; e[1] = -e[1]; e[2] = -e[2]; }
    const double* t = knots + (size_t)ip * npts;
    const double* c = coefs + (size_t)ip * SLAKO_NTAB * npts;
    const Basis b = spline_basis(t, npts, r);
    double T[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) T[k] = 0.0;
    const int need0 = (l1 == 0 && l2 == 0) ? 0 : (l1 == 0 ? 2 : (l2 == 0 ? 4 : 5));
    const int need1 = (l1 == 1 && l2 == 1) ? 6 : need0;
    T[need0] = spline_value(c + (size_t)need0 * npts, b);
    T[need1] = spline_value(c + (size_t)need1 * npts, b);
    v = sk_rule(l1, m1, l2, m2, e, T);
  }
  S[(size_t)mu * norb + nu] = v;
}

}  // namespace

extern "C" int dftb_slako_overlap12(const double* pos1, const double* pos2, const int* orb_atom, const int* orb_l,
                                    const int* orb_m, const int* atom_type, int nat, int norb, const double* knots,
                                    const double* coefs, int npair, int npts, double* S, void* stream) {
  if (nat <= 0 || norb <= 0 || npair <= 0 || npts < 2 * (slako_dev::kDeg + 1) + 2) return DFTB_ERR_ARG;
  if (!pos1 || !pos2 || !orb_atom || !orb_l || !orb_m || !atom_type || !knots || !coefs || !S) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(ceil_div(norb, 256), norb);
  overlap12_kernel<<<grid, 256, 0, st>>>(pos1, pos2, orb_atom, orb_l, orb_m, atom_type, norb, knots, coefs, npair,
                                         npts, S);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_slako_matrices(const double* pos, const int* orb_atom, const int* orb_l, const int* orb_m,
                                   const int* atom_type, const double* orben, int nat, int norb,
                                   const double* knots, const double* coefs, int npair, int npts, double cutoff,
                                   double* S, double* H0, double* D, void* stream) {
  if (nat <= 0 || norb <= 0 || npair <= 0 || npts < 2 * (slako_dev::kDeg + 1) + 2) return DFTB_ERR_ARG;
  if (!pos || !orb_atom || !orb_l || !orb_m || !atom_type || !orben || !knots || !coefs || !S) return DFTB_ERR_ARG;
  if (!(cutoff > 0.0)) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(ceil_div(norb, 256), norb);
  slako_kernel<<<grid, 256, 0, st>>>(pos, orb_atom, orb_l, orb_m, atom_type, orben, norb, knots, coefs, npair, npts,
                                     cutoff, S, H0, D);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
