// Persistent ensemble solver: many molecules of uneven size, one launch.
//
// Each workgroup runs the complete non-Hermitian Davidson loop of
// DFTB/Solver.py:637-917 for one molecule on the device (no host round trip
// per iteration): MFMA matvecs on the molecule's own charges, subspace
// matrices, Newton-Schulz sqrtm(Mm), the lowest-root eigensolver, residuals,
// the Solver.py selection rule for new vectors, preconditioning and CGS2 +
// Cholesky-QR orthonormalisation.  The host orders molecules largest first so
// the big ones start first.  Everything is FP64 except the exchange
// contractions in precision 1 (FP32 operands and intermediates, FP32
// accumulation), which read FP32 copies of the charges.
#include "common.h"
#include "blockops.h"
#include "../../include/dftbaby_amd.h"

#define ENS_THREADS 512
#define ENS_MAXL 1024   // largest subspace (LDS scratch of the eigensolver)
#define ENS_MAXK 32     // largest number of states

// --------------------------------------------------------------------------
// block-level MFMA GEMM with direct (L2/L1-cached) operand loads
//   C = alpha op(A) op(B) + beta C; optional column remap of the output:
//   col -> (col / cdiv) * cstr + row * ldc + col % cdiv.
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T ld_op(const T* X, long ld, int trans, long r, long c) {
  return trans ? X[c * ld + r] : X[r * ld + c];
}

template <typename T, typename TC>
__device__ __noinline__ void bgemm(int M, int N, int K, const T* __restrict__ A, long lda, int at,
                                   const T* __restrict__ B, long ldb, int bt, TC* __restrict__ C, long ldc,
                                   double alpha, double beta, long cdiv, long cstr) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int tm = (M + 31) / 32, tn = (N + 31) / 32;
  constexpr bool F64 = sizeof(T) == 8;
  for (int t = wv; t < tm * tn; t += nw) {
    const int m0 = (t / tn) * 32, n0 = (t % tn) * 32;
    d4_t acc[2][2];
    f4_t accf[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc[i][j] = (d4_t)(0.0);
        accf[i][j] = (f4_t)(0.0f);
      }
    const int mr0 = m0 + (lane & 15), nr0 = n0 + (lane & 15);
    for (int k0 = 0; k0 < K; k0 += 16) {
      T a[4][2], b[4][2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kk = k0 + 4 * u + (lane >> 4);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int m = mr0 + 16 * i;
          a[u][i] = (m < M && kk < K) ? ld_op(A, lda, at, m, kk) : (T)0;
          const int n = nr0 + 16 * i;
          b[u][i] = (n < N && kk < K) ? ld_op(B, ldb, bt, kk, n) : (T)0;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if constexpr (F64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[u][i], (double)b[u][j], acc[i][j], 0, 0, 0);
            else
              accf[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a[u][i], (float)b[u][j], accf[i][j], 0, 0, 0);
          }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // v_mfma_f64_16x16x4f64 holds rows (lane/16) + 4r, the f32 form rows 4(lane/16) + r
          const int rr = F64 ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
          const long row = m0 + 16 * i + rr;
          const long col = n0 + 16 * j + (lane & 15);
          if (row < M && col < N) {
            const long off = cdiv > 0 ? (col / cdiv) * cstr + row * ldc + col % cdiv : row * ldc + col;
            double v = alpha * (F64 ? (double)acc[i][j][r] : (double)accf[i][j][r]);
            if (beta != 0.0) v += beta * (double)C[off];
            C[off] = (TC)v;
          }
        }
  }
  __syncthreads();
}

template <typename T, typename TC>
__device__ __forceinline__ void bg(int M, int N, int K, const T* A, long lda, int at, const T* B, long ldb, int bt,
                                   TC* C, long ldc, double alpha = 1.0, double beta = 0.0) {
  bgemm<T, TC>(M, N, K, A, lda, at, B, ldb, bt, C, ldc, alpha, beta, 0, 0);
}

__device__ double bsum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

__device__ double bmax(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s = fmax(s, sh[i]);
  __syncthreads();
  return s;
}

// --------------------------------------------------------------------------
// per-molecule workspace layout (host and device agree)
// --------------------------------------------------------------------------
struct EnsLayout {
  size_t TT, X1, X2, tv, gv, Sq, Sinv, H, Y2, Z2, P, Tns, eig, Tbv, Rb, Lb, R, L, PR, ML, W, nrm, w2, G, Lc, Li, ref,
      Qt, v32, keep, total;
};

__host__ __device__ static long ens_mmax(const EnsMolecule& m) {
  long a = 2L * m.k, b = m.nguess;
  return a > b ? a : b;
}

__host__ __device__ static EnsLayout ens_layout(const EnsMolecule& m) {
  EnsLayout L;
  const long N = (long)m.opad * m.vpad;
  const long rows = m.opad > m.vpad ? m.opad : m.vpad;
  const long mm = ens_mmax(m);
  const long cap = m.cap, k = m.k;
  size_t o = 0;
  size_t* fields[] = {&L.TT, &L.X1, &L.X2, &L.tv, &L.gv, &L.Sq, &L.Sinv, &L.H, &L.Y2, &L.Z2, &L.P, &L.Tns, &L.eig,
                      &L.Tbv, &L.Rb, &L.Lb, &L.R, &L.L, &L.PR, &L.ML, &L.W, &L.nrm, &L.w2, &L.G, &L.Lc, &L.Li,
                      &L.ref, &L.Qt, &L.v32, &L.keep};
  const size_t sizes[] = {(size_t)m.nat * rows * m.vpad * 8, (size_t)mm * N * 8, (size_t)mm * N * 8,
                          (size_t)mm * m.nat * 8, (size_t)mm * m.nat * 8, (size_t)cap * cap * 8,
                          (size_t)cap * cap * 8, (size_t)cap * cap * 8, (size_t)cap * cap * 8,
                          (size_t)cap * cap * 8, (size_t)cap * cap * 8, (size_t)cap * cap * 8,
                          eig_layout((int)cap, (int)k).total, (size_t)cap * k * 8, (size_t)cap * k * 8,
                          (size_t)cap * k * 8, (size_t)k * N * 8, (size_t)k * N * 8, (size_t)k * N * 8,
                          (size_t)k * N * 8, (size_t)2 * k * N * 8, (size_t)2 * k * 8, (size_t)k * 8,
                          (size_t)mm * mm * 8, (size_t)mm * mm * 8, (size_t)mm * mm * 8, (size_t)mm * 8,
                          (size_t)mm * N * 8, (size_t)mm * N * 4, (size_t)mm * 4};
  for (int i = 0; i < 30; ++i) {
    *fields[i] = o;
    o += al256(sizes[i]);
  }
  L.total = o;
  return L;
}

extern "C" size_t dftb_ensemble_workspace(const EnsMolecule* m) { return ens_layout(*m).total; }

// --------------------------------------------------------------------------
// matvec of m vectors V (FP64, [m][N]) -> outP = (A+B)V, outM = (A-B)V
// --------------------------------------------------------------------------
template <typename T>
__device__ void ens_matvec(const EnsMolecule& Mo, char* ws, const EnsLayout& L, const double* Vd, int m, double* outP,
                           double* outM) {
  const long N = (long)Mo.opad * Mo.vpad, nat = Mo.nat, op = Mo.opad, vp = Mo.vpad;
  double* tv = (double*)(ws + L.tv);
  double* gv = (double*)(ws + L.gv);
  double* X1 = (double*)(ws + L.X1);
  double* X2 = (double*)(ws + L.X2);
  const long total = (long)m * N;
  // Coulomb: t = V q_c^T, g = t gamma, outP = 4 g q_c  (FP64)
  bg<double, double>(m, (int)nat, (int)N, Vd, N, 0, Mo.q_c, N, 1, tv, nat);
  bg<double, double>(m, (int)nat, (int)nat, tv, nat, 0, Mo.gamma, nat, 0, gv, nat);
  bg<double, double>(m, (int)N, (int)nat, gv, nat, 0, Mo.q_c, N, 0, outP, N, 4.0, 0.0);
  if (Mo.lc) {
    for (long i = threadIdx.x; i < total; i += blockDim.x) {
      X1[i] = 0.0;
      X2[i] = 0.0;
    }
    const T* V = (const T*)Vd;
    if (sizeof(T) == 4) {
      float* v32 = (float*)(ws + L.v32);
      for (long i = threadIdx.x; i < total; i += blockDim.x) v32[i] = (float)Vd[i];
      V = (const T*)v32;
    }
    __syncthreads();
    const T* qoo = (const T*)(sizeof(T) == 4 ? Mo.q32[0] : Mo.q_oo);
    const T* qov = (const T*)(sizeof(T) == 4 ? Mo.q32[1] : Mo.q_ov);
    const T* qfov = (const T*)(sizeof(T) == 4 ? Mo.q32[2] : Mo.qf_ov);
    const T* qfvv = (const T*)(sizeof(T) == 4 ? Mo.q32[3] : Mo.qf_vv);
    T* TT = (T*)(ws + L.TT);
    for (int n = 0; n < m; ++n) {
      const T* v = V + (long)n * N;
      // K1 expand: T[A][j][a] = sum_b V[j][b] qf_vv[b][A][a]
      bgemm<T, T>((int)op, (int)(nat * vp), (int)vp, v, vp, 0, qfvv, nat * vp, 0, TT, vp, 1.0, 0.0, vp, op * vp);
      // K1 contract: X1[i][a] = sum_{A,j} q_oo[i][A][j] T[A][j][a]
      bgemm<T, double>(Mo.nocc, (int)vp, (int)(nat * op), qoo, nat * op, 0, TT, vp, 0, X1 + (long)n * N, vp, 1.0, 0.0,
                       0, 0);
      // K2 expand: U[A][b][a] = sum_j V[j][b] qf_ov[j][A][a]
      bgemm<T, T>((int)vp, (int)(nat * vp), Mo.nocc, v, vp, 1, qfov, nat * vp, 0, TT, vp, 1.0, 0.0, vp, vp * vp);
      // K2 contract: X2[i][a] = sum_{A,b} q_ov[i][A][b] U[A][b][a]
      bgemm<T, double>(Mo.nocc, (int)vp, (int)(nat * vp), qov, nat * vp, 0, TT, vp, 0, X2 + (long)n * N, vp, 1.0, 0.0,
                       0, 0);
    }
  }
  for (long i = threadIdx.x; i < total; i += blockDim.x) {
    const long e = i % N;
    const double ov = Mo.omega[e] * Vd[i];
    const double k1 = Mo.lc ? X1[i] : 0.0, k2 = Mo.lc ? X2[i] : 0.0;
    outP[i] = ov + outP[i] - k1 - k2;
    outM[i] = ov - k1 + k2;
  }
  __syncthreads();
}

// --------------------------------------------------------------------------
// Newton-Schulz sqrt of the l x l block of A (ld): S = A^{1/2}, Si = A^{-1/2}
// returns 0 on success, nonzero if not converged (A not positive definite)
// --------------------------------------------------------------------------
__device__ int ens_sqrtm(const double* A, int l, int ld, double* S, double* Si, double* T, double* P, double* Y2,
                         double* Z2, double* sh) {
  double fro = 0.0;
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    double a = A[(i / l) * ld + i % l];
    fro += a * a;
  }
  const double c = fmax(sqrt(bsum(fro, sh)), 1e-300);
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    S[i] = 0.5 * (A[(long)r * ld + q] + A[(long)q * ld + r]) / c;
    Si[i] = (r == q) ? 1.0 : 0.0;
  }
  __syncthreads();
  double err = 1.0;
  int ok = 0;
  for (int it = 0; it < 100; ++it) {
    bg<double, double>(l, l, l, Si, l, 0, S, l, 0, P, l);
    double m = 0.0;
    for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
      int r = (int)(i / l), q = (int)(i % l);
      double id = (r == q) ? 1.0 : 0.0;
      m = fmax(m, fabs(P[i] - id));
      T[i] = 0.5 * (3.0 * id - P[i]);
    }
    err = bmax(m, sh);
    if (!(err == err) || err > 1e6) break;
    if (err < 1e-14 * (l > 16 ? l : 16)) {
      ok = 1;
      break;
    }
    bg<double, double>(l, l, l, S, l, 0, T, l, 0, Y2, l);
    bg<double, double>(l, l, l, T, l, 0, Si, l, 0, Z2, l);
    for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
      S[i] = Y2[i];
      Si[i] = Z2[i];
    }
    __syncthreads();
  }
  if (!ok && err < 1e-10) ok = 1;
  const double f = sqrt(c);
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    S[i] *= f;
    Si[i] /= f;
  }
  __syncthreads();
  return ok ? 0 : 1;
}

// --------------------------------------------------------------------------
// orthonormalise rows Q[0..m) (N) against B[0..l) and among themselves;
// dependent rows dropped (compacted); returns the number kept
// --------------------------------------------------------------------------
__device__ int ens_orthonormalize(const double* B, int l, double* Q, int m, long N, double tol, double* Cc, double* G,
                                  double* Lc, double* Li, double* ref, double* Qt, int* keep, double* sh) {
  for (int t = 0; t < m; ++t) {
    double s = 0.0;
    for (long i = threadIdx.x; i < N; i += blockDim.x) s += Q[(long)t * N + i] * Q[(long)t * N + i];
    s = bsum(s, sh);
    if (threadIdx.x == 0) {
      ref[t] = sqrt(s);
      keep[t] = 1;
    }
  }
  __syncthreads();
  for (int pass = 0; pass < 2 && l > 0; ++pass) {
    bg<double, double>(l, m, (int)N, B, N, 0, Q, N, 1, Cc, m);            // Cc[p][t] = <B_p, Q_t>
    bg<double, double>(m, (int)N, l, Cc, m, 1, B, N, 0, Q, N, -1.0, 1.0);  // Q -= Cc^T B
  }
  for (int pass = 0; pass < 2; ++pass) {
    bg<double, double>(m, m, (int)N, Q, N, 0, Q, N, 1, G, m);
    if (threadIdx.x == 0) {
      for (int i = 0; i < m * m; ++i) {
        Lc[i] = 0.0;
        Li[i] = 0.0;
      }
      for (int j = 0; j < m; ++j) {
        double d = G[j * m + j];
        for (int q = 0; q < j; ++q) d -= Lc[j * m + q] * Lc[j * m + q];
        const double r = pass == 0 ? ref[j] * ref[j] : 1.0;
        const double t2 = pass == 0 ? tol * tol : 1e-24;
        const int ok = keep[j] && d > t2 * r && d > 1e-12 * G[j * m + j] && d > 0.0;
        keep[j] = ok;
        Lc[j * m + j] = ok ? sqrt(d) : 1.0;
        for (int i = j + 1; i < m; ++i) {
          double s = 0.0;
          if (ok) {
            s = G[i * m + j];
            for (int q = 0; q < j; ++q) s -= Lc[i * m + q] * Lc[j * m + q];
            s /= Lc[j * m + j];
          }
          Lc[i * m + j] = s;
        }
      }
      for (int c = 0; c < m; ++c)
        for (int i = c; i < m; ++i) {
          double s = (i == c) ? 1.0 : 0.0;
          for (int q = c; q < i; ++q) s -= Lc[i * m + q] * Li[q * m + c];
          Li[i * m + c] = s / Lc[i * m + i];
        }
      for (int i = 0; i < m; ++i)
        if (!keep[i])
          for (int c = 0; c < m; ++c) Li[i * m + c] = 0.0;
    }
    __syncthreads();
    bg<double, double>(m, (int)N, m, Li, m, 0, Q, N, 0, Qt, N);
    for (long i = threadIdx.x; i < (long)m * N; i += blockDim.x) Q[i] = Qt[i];
    __syncthreads();
  }
  // compact kept rows to the front
  int kept = 0;
  for (int t = 0; t < m; ++t) {
    if (keep[t]) {
      if (kept != t)
        for (long i = threadIdx.x; i < N; i += blockDim.x) Q[(long)kept * N + i] = Q[(long)t * N + i];
      __syncthreads();
      ++kept;
    }
  }
  return kept;
}

// --------------------------------------------------------------------------
// the per-molecule Davidson
// --------------------------------------------------------------------------
template <typename T>
__device__ void ens_davidson(const EnsMolecule& Mo, double conv, int maxiter, double* sh, double* sv, double* sp,
                             double* bnd, int* s_sel) {
  char* ws = (char*)Mo.work;
  const EnsLayout L = ens_layout(Mo);
  const long N = (long)Mo.opad * Mo.vpad;
  const int k = Mo.k, cap = Mo.cap;
  const long kmax = (long)Mo.nocc * Mo.nvirt;
  double* bs = Mo.bs;
  double* bp = Mo.bp;
  double* bm = Mo.bm;
  double* Mp = Mo.Mp;
  double* Mm = Mo.Mm;
  double* Sq = (double*)(ws + L.Sq);
  double* Sinv = (double*)(ws + L.Sinv);
  double* H = (double*)(ws + L.H);
  double* Tb = (double*)(ws + L.Tbv);
  double* Rb = (double*)(ws + L.Rb);
  double* Lb = (double*)(ws + L.Lb);
  double* R = (double*)(ws + L.R);
  double* Lv = (double*)(ws + L.L);
  double* PR = (double*)(ws + L.PR);
  double* ML = (double*)(ws + L.ML);
  double* W = (double*)(ws + L.W);
  double* nrm = (double*)(ws + L.nrm);
  double* w2 = (double*)(ws + L.w2);
  double* w = Mo.w;
  int* status = Mo.status;
  // initial guess: unit vectors at the host-sorted lowest-omega indices
  int l = Mo.nguess;
  for (long i = threadIdx.x; i < (long)l * N; i += blockDim.x) bs[i] = 0.0;
  __syncthreads();
  for (int t = threadIdx.x; t < l; t += blockDim.x) bs[(long)t * N + Mo.guess[t]] = 1.0;
  __syncthreads();
  int new0 = 0;
  int code = 0, it = 0;
  for (it = 0; it < maxiter; ++it) {
    const int m = l - new0;
    if (m > 0) {
      ens_matvec<T>(Mo, ws, L, bs + (long)new0 * N, m, bp + (long)new0 * N, bm + (long)new0 * N);
      // new columns/rows of M+- = bs^T (A+-B) bs
      bg<double, double>(l, m, (int)N, bs, N, 0, bp + (long)new0 * N, N, 1, Mp + new0, cap);
      bg<double, double>(l, m, (int)N, bs, N, 0, bm + (long)new0 * N, N, 1, Mm + new0, cap);
      if (new0 > 0) {
        bg<double, double>(m, new0, (int)N, bs + (long)new0 * N, N, 0, bp, N, 1, Mp + (long)new0 * cap, cap);
        bg<double, double>(m, new0, (int)N, bs + (long)new0 * N, N, 0, bm, N, 1, Mm + (long)new0 * cap, cap);
      }
    }
    if (sizeof(T) == 4) {
      // FP32 products: use the symmetric parts of M+-
      for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
        int r = (int)(i / l), c = (int)(i % l);
        if (c > r) {
          double a = 0.5 * (Mp[(long)r * cap + c] + Mp[(long)c * cap + r]);
          Mp[(long)r * cap + c] = a;
          Mp[(long)c * cap + r] = a;
          double b = 0.5 * (Mm[(long)r * cap + c] + Mm[(long)c * cap + r]);
          Mm[(long)r * cap + c] = b;
          Mm[(long)c * cap + r] = b;
        }
      }
      __syncthreads();
    }
    // subspace step (Solver.py:759-798)
    if (ens_sqrtm(Mm, l, cap, Sq, Sinv, (double*)(ws + L.Tns), (double*)(ws + L.P), (double*)(ws + L.Y2),
                  (double*)(ws + L.Z2), sh)) {
      code = -3;
      break;
    }
    bg<double, double>(l, l, l, Mp, cap, 0, Sq, l, 0, (double*)(ws + L.P), l);
    bg<double, double>(l, l, l, Sq, l, 0, (double*)(ws + L.P), l, 0, H, l);
    double he = 0.0;
    for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
      int r = (int)(i / l), c = (int)(i % l);
      he += fabs(H[(long)r * l + c] - H[(long)c * l + r]);
    }
    he = bsum(he, sh);
    if (sizeof(T) == 8 && he > 1.0e-10) {
      code = -4;
      break;
    }
    const int nev = k < l ? k : l;
    syevx_block(H, l, l, nev, w2, Tb, ws + L.eig, nullptr, sv, sp, sh, bnd, ENS_MAXL);
    __syncthreads();
    int bad = 0;
    for (int n = threadIdx.x; n < nev; n += blockDim.x)
      if (!(w2[n] > 0.0)) bad = 1;
    if (bsum((double)bad, sh) > 0.0) {
      code = -3;
      break;
    }
    for (int n = threadIdx.x; n < nev; n += blockDim.x) w[n] = sqrt(w2[n]);
    __syncthreads();
    bg<double, double>(l, nev, l, Sq, l, 0, Tb, nev, 0, Rb, nev);
    for (long i = threadIdx.x; i < (long)l * nev; i += blockDim.x) Rb[i] /= sqrt(w[i % nev]);
    __syncthreads();
    bg<double, double>(l, nev, l, Mp, cap, 0, Rb, nev, 0, Lb, nev);
    for (long i = threadIdx.x; i < (long)l * nev; i += blockDim.x) Lb[i] /= w[i % nev];
    __syncthreads();
    // R, L and their products
    bg<double, double>(nev, (int)N, l, Rb, nev, 1, bs, N, 0, R, N);
    bg<double, double>(nev, (int)N, l, Lb, nev, 1, bs, N, 0, Lv, N);
    bg<double, double>(nev, (int)N, l, Rb, nev, 1, bp, N, 0, PR, N);
    bg<double, double>(nev, (int)N, l, Lb, nev, 1, bm, N, 0, ML, N);
    // residuals WL = (A+B)R - w L, WR = (A-B)L - w R
    for (int n = 0; n < nev; ++n) {
      double sL = 0.0, sR = 0.0;
      for (long i = threadIdx.x; i < N; i += blockDim.x) {
        double a = PR[(long)n * N + i] - w[n] * Lv[(long)n * N + i];
        double b = ML[(long)n * N + i] - w[n] * R[(long)n * N + i];
        W[(long)n * N + i] = a;
        W[(long)(k + n) * N + i] = b;
        sL += a * a;
        sR += b * b;
      }
      sL = bsum(sL, sh);
      sR = bsum(sR, sh);
      if (threadIdx.x == 0) {
        nrm[n] = sqrt(sL);
        nrm[k + n] = sqrt(sR);
      }
    }
    __syncthreads();
    // convergence and selection (thread 0; Solver.py:837-884)
    if (threadIdx.x == 0) {
      int all = 1;
      for (int n = 0; n < nev; ++n)
        if (!(nrm[n] + nrm[k + n] < conv)) all = 0;
      s_sel[0] = (all && it > 0) ? -1 : 0;
      if (!(all && it > 0)) {
        const long dkmax = (kmax - l < 2 * k) ? kmax - l : 2 * k;
        const double eps = 0.01 * conv;
        int ncL = 0, ncR = 0;
        for (int n = 0; n < nev; ++n) {
          ncL += nrm[n] > eps;
          ncR += nrm[k + n] > eps;
        }
        int dkR = (int)((dkmax + 1) / 2);
        if (dkR > ncL) dkR = ncL;
        int dkL = (int)(dkmax - dkR);
        if (dkL > ncR) dkL = ncR;
        const int dk = dkR + dkL;
        int nb = 0;
        for (int n = 0; n < nev && nb < dk; ++n)
          if (nrm[n] > eps) s_sel[1 + nb++] = n;
        for (int n = 0; n < nev && nb < dk; ++n)
          if (nrm[k + n] > eps) s_sel[1 + nb++] = k + n;
        if (l + nb > cap) nb = -2;
        s_sel[0] = nb;
      }
    }
    __syncthreads();
    const int dk = s_sel[0];
    if (dk == -1) {
      code = 1;
      break;
    }
    if (dk < 0) {
      code = -6;
      break;
    }
    if (dk == 0) {
      new0 = l;
      continue;
    }
    // precondition into bs[l .. l+dk)
    double* Q = bs + (long)l * N;
    for (long i = threadIdx.x; i < (long)dk * N; i += blockDim.x) {
      const int t = (int)(i / N);
      const long e = i - (long)t * N;
      const int src = s_sel[1 + t];
      const double om = Mo.omega[e];
      double d = w[src % k] - om;
      if (fabs(d) < 1.0e-6) d = 1.0e-6 * om;
      Q[i] = W[(long)src * N + e] / d;
    }
    __syncthreads();
    const int kept = ens_orthonormalize(bs, l, Q, dk, N, 1.0e-8, (double*)(ws + L.Sq), (double*)(ws + L.G),
                                        (double*)(ws + L.Lc), (double*)(ws + L.Li), (double*)(ws + L.ref),
                                        (double*)(ws + L.Qt), (int*)(ws + L.keep), sh);
    new0 = l;
    l += kept;
  }
  // outputs: C = bs Tb, X+Y = R, X-Y = L
  if (code == 1) {
    const int nev = k < l ? k : l;
    bg<double, double>(nev, (int)N, l, Tb, nev, 1, bs, N, 0, Mo.C, N);
    for (long i = threadIdx.x; i < (long)nev * N; i += blockDim.x) {
      Mo.XpY[i] = R[i];
      Mo.XmY[i] = Lv[i];
    }
    for (int n = threadIdx.x; n < nev; n += blockDim.x) Mo.resid[n] = nrm[n] + nrm[k + n];
  }
  if (threadIdx.x == 0) {
    status[0] = it + 1;
    status[1] = code;
    status[2] = l;
  }
  __syncthreads();
}

__global__ __launch_bounds__(ENS_THREADS) void ensemble_kernel(const EnsMolecule* __restrict__ mols, int nmol,
                                                               double conv, int maxiter) {
  __shared__ double sh[ENS_THREADS / 64];
  __shared__ double sv[ENS_MAXL];
  __shared__ double sp[ENS_MAXL];
  __shared__ double bnd[2];
  __shared__ int s_sel[1 + 2 * ENS_MAXK];
  const int mi = blockIdx.x;
  if (mi >= nmol) return;
  const EnsMolecule& Mo = mols[mi];
  if (Mo.precision == 1)
    ens_davidson<float>(Mo, conv, maxiter, sh, sv, sp, bnd, s_sel);
  else
    ens_davidson<double>(Mo, conv, maxiter, sh, sv, sp, bnd, s_sel);
}

extern "C" int dftb_ensemble_solve(const EnsMolecule* mols_dev, int nmol, double conv, int maxiter, void* stream) {
  if (!mols_dev || nmol <= 0 || maxiter <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(ensemble_kernel, dim3(nmol), dim3(ENS_THREADS), 0, (hipStream_t)stream, mols_dev, nmol, conv,
                     maxiter);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
