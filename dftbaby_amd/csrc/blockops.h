// Block-level (one workgroup) FP64 building blocks shared by the single-molecule
// kernels and the persistent ensemble kernel.
#pragma once
#include <math.h>
#include "common.h"

#define EIG_SMALL_CLUSTER 8

__host__ __device__ static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct EigLayout {
  size_t W, d, e, tau, z, scratch, coef, total;
};
__host__ __device__ static EigLayout eig_layout(int l, int nev) {
  EigLayout L;
  size_t off = 0;
  L.W = off; off += al256((size_t)l * l * 8);
  L.d = off; off += al256((size_t)l * 8);
  L.e = off; off += al256((size_t)l * 8);
  L.tau = off; off += al256((size_t)l * 8);
  L.z = off; off += al256((size_t)l * nev * 8);
  L.scratch = off; off += al256((size_t)6 * l * nev * 8);
  L.coef = off; off += al256((size_t)nev * 8);
  L.total = off;
  return L;
}


__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ double blk_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;
}

// Sturm count: number of eigenvalues of the tridiagonal (d, e) below x.
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) int lds_i32;

// d, e in LDS (address-space-3 pointers: ds_read instead of flat loads)
__device__ int sturm_count(const lds_f64* d, const lds_f64* e, int l, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0) ++c;
  for (int i = 1; i < l; ++i) {
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
  }
  return c;
}

__device__ __forceinline__ double hash_unit(unsigned int a, unsigned int b) {
  unsigned int h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
  return ((double)(h & 0xFFFFFF) / (double)0xFFFFFF) * 2.0 - 1.0;
}

// LU factorisation of T - lam I with partial pivoting between rows k and k+1
// (u0r: reciprocals of the pivots, u1/u2: the two superdiagonals of U, mlt: the
// multipliers, piv: 1 where rows were interchanged).
__device__ void tri_factor(const double* __restrict__ d, const double* __restrict__ e, int l, double lam,
                           double tiny, double* __restrict__ u0r, double* __restrict__ u1, double* __restrict__ u2,
                           double* __restrict__ mlt, int* __restrict__ piv) {
  double r0 = d[0] - lam, r1 = (l > 1) ? e[0] : 0.0;
  for (int k = 0; k < l - 1; ++k) {
    const double sub = e[k];
    const double a1 = d[k + 1] - lam;
    const double b1 = (k + 1 < l - 1) ? e[k + 1] : 0.0;
    if (fabs(r0) >= fabs(sub)) {
      if (fabs(r0) < tiny) r0 = (r0 >= 0 ? tiny : -tiny);
      const double m = sub / r0;
      u0r[k] = 1.0 / r0; u1[k] = r1; u2[k] = 0.0; mlt[k] = m; piv[k] = 0;
      r0 = a1 - m * r1;
      r1 = b1;
    } else {
      const double m = r0 / sub;
      u0r[k] = 1.0 / sub; u1[k] = a1; u2[k] = b1; mlt[k] = m; piv[k] = 1;
      r0 = r1 - m * a1;
      r1 = -m * b1;
    }
  }
  if (fabs(r0) < tiny) r0 = (r0 >= 0 ? tiny : -tiny);
  u0r[l - 1] = 1.0 / r0;
  u1[l - 1] = 0.0;
  u2[l - 1] = 0.0;
}

// x <- (T - lam I)^{-1} x with the factors of tri_factor.  The entries being
// eliminated are carried in registers; the operands of TRI_B consecutive steps
// are loaded together, without bounds tests (full blocks first, then the
// remainder one step at a time), so one memory latency is paid per block of
// steps instead of per step.
#define TRI_B 8
template <typename D, typename I>
__device__ void tri_apply(int l, const D* __restrict__ u0r, const D* __restrict__ u1, const D* __restrict__ u2,
                          const D* __restrict__ mlt, const I* __restrict__ piv, D* __restrict__ x) {
  double c = x[0];
  int k = 0;
  for (; k + TRI_B <= l - 1; k += TRI_B) {
    double nx[TRI_B], m[TRI_B];
    int pv[TRI_B];
#pragma unroll
    for (int u = 0; u < TRI_B; ++u) {
      nx[u] = x[k + u + 1];
      m[u] = mlt[k + u];
      pv[u] = piv[k + u];
    }
#pragma unroll
    for (int u = 0; u < TRI_B; ++u) {
      const double keep = pv[u] ? nx[u] : c;
      c = pv[u] ? c - m[u] * nx[u] : nx[u] - m[u] * c;
      x[k + u] = keep;
    }
  }
  for (; k < l - 1; ++k) {
    const double nx = x[k + 1], m = mlt[k];
    const double keep = piv[k] ? nx : c;
    c = piv[k] ? c - m * nx : nx - m * c;
    x[k] = keep;
  }
  x[l - 1] = c;
  double x1 = 0.0, x2 = 0.0;
  k = l - 1;
  for (; k - TRI_B + 1 >= 0; k -= TRI_B) {
    double xv[TRI_B], a1[TRI_B], a2[TRI_B], r0[TRI_B];
#pragma unroll
    for (int u = 0; u < TRI_B; ++u) {
      xv[u] = x[k - u];
      a1[u] = u1[k - u];
      a2[u] = u2[k - u];
      r0[u] = u0r[k - u];
    }
#pragma unroll
    for (int u = 0; u < TRI_B; ++u) {
      const double sv = (xv[u] - a1[u] * x1 - a2[u] * x2) * r0[u];
      x[k - u] = sv;
      x2 = x1;
      x1 = sv;
    }
  }
  for (; k >= 0; --k) {
    const double sv = (x[k] - u1[k] * x1 - u2[k] * x2) * u0r[k];
    x[k] = sv;
    x2 = x1;
    x1 = sv;
  }
}


// Lowest nev eigenpairs of the symmetric l x l matrix A (leading dim ld), one
// workgroup: Householder tridiagonalisation, Sturm bisection, inverse
// iteration (clusters re-orthogonalised), back-transformation.  sv, sp: LDS
// scratch of cap >= l doubles each; sh: >= blockDim/64 doubles; s_bnd: 2.
__device__ __noinline__ void syevx_tail(int l, int nev, double* __restrict__ evals, double* __restrict__ evecs,
                                        char* __restrict__ ws, int* __restrict__ info, double* sv, double* sp,
                                        double* sh, double* s_bnd, int cap);

__device__ __noinline__ void syevx_block(const double* __restrict__ A, int l, int ld, int nev,
                                         double* __restrict__ evals, double* __restrict__ evecs,
                                         char* __restrict__ ws, int* __restrict__ info, double* sv, double* sp,
                                         double* sh, double* s_bnd, int cap) {
  EigLayout Lo = eig_layout(l, nev);
  double* W = (double*)(ws + Lo.W);
  double* d = (double*)(ws + Lo.d);
  double* e = (double*)(ws + Lo.e);
  double* tau = (double*)(ws + Lo.tau);
  double* Z = (double*)(ws + Lo.z);
  double* scr = (double*)(ws + Lo.scratch);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;

  for (long i = tid; i < (long)l * l; i += blockDim.x) {
    long r = i / l, c = i % l;
    W[i] = 0.5 * (A[r * ld + c] + A[c * ld + r]);
  }
  __syncthreads();

  // ---- 1. Householder tridiagonalisation
  for (int j = 0; j + 2 < l; ++j) {
    const int n2 = l - j - 1;
    double part = 0.0;
    for (int r = tid; r < n2; r += blockDim.x) {
      double x = W[(long)(j + 1 + r) * l + j];
      sv[r] = x;
      part += x * x;
    }
    const double alpha2 = blk_sum(part, sh);
    const double x0 = sv[0];
    __syncthreads();
    if (alpha2 <= 1e-300) {
      if (tid == 0) { tau[j] = 0.0; e[j] = 0.0; d[j] = W[(long)j * l + j]; }
      for (int r = tid; r < n2; r += blockDim.x) W[(long)(j + 1 + r) * l + j] = 0.0;
      __syncthreads();
      continue;
    }
    const double alpha = sqrt(alpha2);
    const double beta = (x0 >= 0.0) ? -alpha : alpha;
    const double v0 = x0 - beta;
    const double vtv = alpha2 - x0 * x0 + v0 * v0;
    const double t = 2.0 / vtv;
    if (tid == 0) { sv[0] = v0; tau[j] = t; e[j] = beta; d[j] = W[(long)j * l + j]; }
    __syncthreads();
    // p = t * A22 v  (one wave per row, lanes across columns)
    for (int r = wv; r < n2; r += nw) {
      const double* row = W + (long)(j + 1 + r) * l + (j + 1);
      double s = 0.0;
      for (int c = lane; c < n2; c += 64) s += row[c] * sv[c];
      s = wave_sum(s);
      if (lane == 0) sp[r] = t * s;
    }
    __syncthreads();
    part = 0.0;
    for (int r = tid; r < n2; r += blockDim.x) part += sp[r] * sv[r];
    const double K = 0.5 * t * blk_sum(part, sh);
    for (int r = tid; r < n2; r += blockDim.x) sp[r] -= K * sv[r];
    __syncthreads();
    for (long i = tid; i < (long)n2 * n2; i += blockDim.x) {
      int r = (int)(i / n2), c = (int)(i % n2);
      W[(long)(j + 1 + r) * l + (j + 1 + c)] -= sv[r] * sp[c] + sp[r] * sv[c];
    }
    for (int r = tid; r < n2; r += blockDim.x) W[(long)(j + 1 + r) * l + j] = sv[r];
    __syncthreads();
  }
  syevx_tail(l, nev, evals, evecs, ws, info, sv, sp, sh, s_bnd, cap);
}

// Steps after the Householder reduction of W (reflectors below the first
// subdiagonal, d[0..l-3], e[0..l-3], tau set): last 2x2 block, bisection for
// the nev lowest eigenvalues and inverse iteration into Z (l x nev, row-major).
// sv, sp: LDS scratch of cap >= l doubles each (they hold d and e).
__device__ __noinline__ void syevx_tridiag_vectors(int l, int nev, double* __restrict__ evals,
                                                   char* __restrict__ ws, double* sv, double* sp, double* sh,
                                                   double* s_bnd, int cap) {
  EigLayout Lo = eig_layout(l, nev);
  double* W = (double*)(ws + Lo.W);
  double* d = (double*)(ws + Lo.d);
  double* e = (double*)(ws + Lo.e);
  double* Z = (double*)(ws + Lo.z);
  double* scr = (double*)(ws + Lo.scratch);
  const int tid = threadIdx.x;
  if (tid == 0) {
    if (l >= 2) {
      d[l - 2] = W[(long)(l - 2) * l + (l - 2)];
      e[l - 2] = W[(long)(l - 1) * l + (l - 2)];
    }
    d[l - 1] = W[(long)(l - 1) * l + (l - 1)];
    if (l >= 1) e[l - 1] = 0.0;
    // Gershgorin bounds
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < l; ++i) {
      double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < l - 1 ? fabs(e[i]) : 0.0);
      lo = fmin(lo, d[i] - r);
      hi = fmax(hi, d[i] + r);
    }
    s_bnd[0] = lo;
    s_bnd[1] = hi;
  }
  __syncthreads();
  for (int i = tid; i < l; i += blockDim.x) {
    sv[i] = d[i];
    sp[i] = e[i];
  }
  __syncthreads();
  const double* dl = sv;
  const double* el = sp;
  const lds_f64* ldl = (const lds_f64*)sv;   // the same LDS buffers, address-space typed
  const lds_f64* lel = (const lds_f64*)sp;
  const double lo0 = s_bnd[0], hi0 = s_bnd[1];
  const double tnorm = fmax(fabs(lo0), fabs(hi0));
  double emax2 = 0.0;
  for (int i = 0; i + 1 < l; ++i) emax2 = fmax(emax2, el[i] * el[i]);
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  const double eps = 1.1102230246251565e-16;

  // ---- 2. the nev lowest eigenvalues by Sturm counts.  With P = blockDim / nev >= 2
  // points per eigenvalue and round (multisection), every round narrows each
  // interval to one of its P + 1 parts; otherwise plain bisection per thread.
  // Both stop at b - a <= 2 eps max(|a|, |b|) + pivmin.
  // (the counts live in the 6 l nev doubles of scratch: P <= 12 l - 6 keeps them inside)
  const int P = min((int)blockDim.x / nev, 12 * l - 6);
  if (P < 2) {
    for (int m = tid; m < nev; m += blockDim.x) {
      double a = lo0 - 2.0 * eps * tnorm - pivmin, b = hi0 + 2.0 * eps * tnorm + pivmin;
      for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (a + b);
        if (sturm_count(ldl, lel, l, mid, pivmin) > m) b = mid; else a = mid;
        if (b - a <= 2.0 * eps * fmax(fabs(a), fabs(b)) + pivmin) break;
      }
      evals[m] = 0.5 * (a + b);
    }
  } else {
    double* ivl = scr;                          // [nev][2] intervals
    int* cnt = (int*)(scr + 2 * nev);           // [nev][P] counts
    int* flag = cnt + nev * P;                  // [nev] converged, then [1] all converged
    for (int m = tid; m < nev; m += blockDim.x) {
      ivl[2 * m] = lo0 - 2.0 * eps * tnorm - pivmin;
      ivl[2 * m + 1] = hi0 + 2.0 * eps * tnorm + pivmin;
      flag[m] = 0;
    }
    __syncthreads();
    for (int round = 0; round < 200; ++round) {
      if (tid < nev * P) {
        const int m = tid / P, k = tid % P;
        if (!flag[m]) {
          const double a = ivl[2 * m], b = ivl[2 * m + 1];
          const double x = a + (b - a) * (double)(k + 1) / (double)(P + 1);
          cnt[m * P + k] = sturm_count(ldl, lel, l, x, pivmin);
        }
      }
      __syncthreads();
      for (int m = tid; m < nev; m += blockDim.x) {
        if (flag[m]) continue;
        const double a = ivl[2 * m], b = ivl[2 * m + 1];
        double na = a, nb = b;
        for (int k = 0; k < P; ++k) {
          const double x = a + (b - a) * (double)(k + 1) / (double)(P + 1);
          if (cnt[m * P + k] > m) {
            nb = x;
            break;
          }
          na = x;
        }
        ivl[2 * m] = na;
        ivl[2 * m + 1] = nb;
        if (nb - na <= 2.0 * eps * fmax(fabs(na), fabs(nb)) + pivmin) flag[m] = 1;
      }
      __syncthreads();
      if (tid == 0) {
        int all = 1;
        for (int m = 0; m < nev; ++m) all &= flag[m];
        flag[nev] = all;
      }
      __syncthreads();
      if (flag[nev]) break;
    }
    for (int m = tid; m < nev; m += blockDim.x) evals[m] = 0.5 * (ivl[2 * m] + ivl[2 * m + 1]);
  }
  __syncthreads();

  // ---- 3. inverse iteration.  Clusters (gap < 1e-3 ||T||) of at most EIG_SMALL_CLUSTER
  // members are handled by the thread owning their first member; larger clusters
  // are processed by the whole workgroup with classical Gram-Schmidt (twice)
  // against the earlier members in every iteration.  Member m of a cluster
  // starting at m0 uses the shift evals[m] + 10 (m - m0) eps ||T||; its LU
  // factors (6 l doubles at scr + 6 l m: 1/u0, u1, u2, multipliers, pivots, x)
  // are formed once, for all members in parallel.
  const double ortol = 1.0e-3 * tnorm;
  const double tiny = eps * fmax(tnorm, 1e-300);
  double* coef = (double*)(ws + Lo.coef);
  for (int m = tid; m < nev; m += blockDim.x) {
    int m0 = m;
    while (m0 > 0 && evals[m0] - evals[m0 - 1] < ortol) --m0;
    const double lam = evals[m] + (double)(m - m0) * 10.0 * tiny;
    double* base = scr + (size_t)6 * l * m;
    tri_factor(dl, el, l, lam, tiny, base, base + l, base + 2 * l, base + 3 * l, (int*)(base + 4 * l));
  }
  __syncthreads();
  for (int m0 = tid; m0 < nev; m0 += blockDim.x) {
    if (m0 > 0 && evals[m0] - evals[m0 - 1] < ortol) continue;  // not a cluster head
    int m1 = m0 + 1;
    while (m1 < nev && evals[m1] - evals[m1 - 1] < ortol) ++m1;
    if (m1 - m0 > EIG_SMALL_CLUSTER) continue;
    for (int m = m0; m < m1; ++m) {
      double* base = scr + (size_t)6 * l * m;
      double* x = base + 5 * l;
      for (int i = 0; i < l; ++i) x[i] = hash_unit((unsigned)m, (unsigned)i);
      for (int it = 0; it < 4; ++it) {
        tri_apply(l, base, base + l, base + 2 * l, base + 3 * l, (const int*)(base + 4 * l), x);
        for (int mm = m0; mm < m; ++mm) {
          double s = 0.0;
          for (int i = 0; i < l; ++i) s += Z[(long)i * nev + mm] * x[i];
          for (int i = 0; i < l; ++i) x[i] -= s * Z[(long)i * nev + mm];
        }
        double nrm = 0.0;
        for (int i = 0; i < l; ++i) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        if (!(nrm > 0.0)) nrm = 1.0;
        for (int i = 0; i < l; ++i) x[i] /= nrm;
      }
      for (int i = 0; i < l; ++i) Z[(long)i * nev + m] = x[i];
    }
  }
  __syncthreads();
  // large clusters, cooperatively.  The member's factors and iterate are staged in
  // the LDS buffers (free once the factors exist) when 4 l doubles fit, so the
  // serial solve runs on LDS; the Gram-Schmidt coefficients take one wave each.
  const int lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const bool stage = 4 * l <= cap;
  for (int m0 = 0; m0 < nev;) {
    int m1 = m0 + 1;
    while (m1 < nev && evals[m1] - evals[m1 - 1] < ortol) ++m1;
    if (m1 - m0 > EIG_SMALL_CLUSTER) {
      for (int m = m0; m < m1; ++m) {
        double* base = scr + (size_t)6 * l * m;
        const double *fu0 = base, *fu1 = base + l, *fu2 = base + 2 * l, *fml = base + 3 * l;
        const int* fpv = (const int*)(base + 4 * l);
        double* x = base + 5 * l;
        if (stage) {
          for (int i = tid; i < l; i += blockDim.x) {
            sv[i] = base[i];
            sv[l + i] = base[l + i];
            sv[2 * l + i] = base[2 * l + i];
            sv[3 * l + i] = base[3 * l + i];
            ((int*)(sp + l))[i] = ((const int*)(base + 4 * l))[i];
          }
          fu0 = sv; fu1 = sv + l; fu2 = sv + 2 * l; fml = sv + 3 * l;
          fpv = (const int*)(sp + l);
          x = sp;
        }
        for (int i = tid; i < l; i += blockDim.x) x[i] = hash_unit((unsigned)m, (unsigned)i);
        __syncthreads();
        for (int it = 0; it < 3; ++it) {
          if (tid == 0) {
            if (stage)
              tri_apply(l, (const lds_f64*)fu0, (const lds_f64*)fu1, (const lds_f64*)fu2, (const lds_f64*)fml,
                        (const lds_i32*)fpv, (lds_f64*)x);
            else
              tri_apply(l, fu0, fu1, fu2, fml, fpv, x);
          }
          __syncthreads();
          for (int pass = 0; pass < 2 && m > m0; ++pass) {
            for (int mm = m0 + wv; mm < m; mm += nw) {
              double s = 0.0;
              for (int i = lane; i < l; i += 64) s += Z[(long)i * nev + mm] * x[i];
              s = wave_sum(s);
              if (lane == 0) coef[mm] = s;
            }
            __syncthreads();
            for (int i = tid; i < l; i += blockDim.x) {
              double s = x[i];
              for (int mm = m0; mm < m; ++mm) s -= coef[mm] * Z[(long)i * nev + mm];
              x[i] = s;
            }
            __syncthreads();
          }
          double part = 0.0;
          for (int i = tid; i < l; i += blockDim.x) part += x[i] * x[i];
          double nrm = sqrt(blk_sum(part, sh));
          if (!(nrm > 0.0)) nrm = 1.0;
          for (int i = tid; i < l; i += blockDim.x) x[i] /= nrm;
          __syncthreads();
        }
        for (int i = tid; i < l; i += blockDim.x) Z[(long)i * nev + m] = x[i];
        __syncthreads();
      }
    }
    m0 = m1;
  }
  __syncthreads();
}

// ---- 4. back-transformation z <- H_0 ... H_{l-3} z in one workgroup (reflectors
// in the columns of W below the subdiagonal), evecs = Z.
__device__ __noinline__ void syevx_back_block(int l, int nev, double* __restrict__ evecs, char* __restrict__ ws,
                                              double* sv) {
  EigLayout Lo = eig_layout(l, nev);
  const double* W = (const double*)(ws + Lo.W);
  const double* tau = (const double*)(ws + Lo.tau);
  double* Z = (double*)(ws + Lo.z);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int j = l - 3; j >= 0; --j) {
    const int n2 = l - j - 1;
    const double t = tau[j];
    if (t == 0.0) continue;
    for (int r = tid; r < n2; r += blockDim.x) sv[r] = W[(long)(j + 1 + r) * l + j];
    __syncthreads();
    for (int n = wv; n < nev; n += nw) {
      double s = 0.0;
      for (int r = lane; r < n2; r += 64) s += sv[r] * Z[(long)(j + 1 + r) * nev + n];
      s = t * wave_sum(s);
      for (int r = lane; r < n2; r += 64) Z[(long)(j + 1 + r) * nev + n] -= s * sv[r];
    }
    __syncthreads();
  }
  for (long i = tid; i < (long)l * nev; i += blockDim.x) evecs[i] = Z[i];
}

__device__ __noinline__ void syevx_tail(int l, int nev, double* __restrict__ evals, double* __restrict__ evecs,
                                        char* __restrict__ ws, int* __restrict__ info, double* sv, double* sp,
                                        double* sh, double* s_bnd, int cap) {
  syevx_tridiag_vectors(l, nev, evals, ws, sv, sp, sh, s_bnd, cap);
  syevx_back_block(l, nev, evecs, ws, sv);
}
