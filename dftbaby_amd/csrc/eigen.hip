// (c) FP64 subspace eigensolver and the Davidson subspace steps.
//
// The reference solves the l x l subspace problems with LAPACK
// (scipy.linalg.sqrtm + numpy.linalg.eigh, DFTB/Solver.py:762-772, and eigh at
// Solver.py:484).  Only the lowest roots are used, so the engine computes the
// lowest nev eigenpairs on the GPU:
//   1. Householder reduction to tridiagonal form (one workgroup; the matrix
//      stays L2-resident, the Householder vector and the product vector in LDS),
//   2. Sturm-sequence bisection for each wanted eigenvalue (one thread each),
//   3. inverse iteration with partial-pivot tridiagonal LU, vectors of close
//      eigenvalues (gap < 1e-3 ||T||) re-orthogonalised within their cluster,
//   4. back-transformation with the stored reflectors.
// The non-Hermitian step follows Solver.py:759-798: Mmsq = sqrtm(Mm) (coupled
// Newton-Schulz iteration on MFMA GEMMs), Mh = Mmsq Mp Mmsq, the lowest nev
// eigenpairs of Mh, Rb = Mmsq Tb / sqrt(w), Lb = Mp Rb / w.
#include <math.h>
#include "common.h"
#include "gemm.h"
#include "blockops.h"

#define EIG_THREADS 512
#define EIG_MAXL 4096

extern "C" size_t dftb_syevx_workspace(int l, int nev) { return eig_layout(l, nev).total; }

__global__ __launch_bounds__(EIG_THREADS) void syevx_kernel(const double* __restrict__ A, int l, int ld, int nev,
                                                             double* __restrict__ evals, double* __restrict__ evecs,
                                                             char* __restrict__ ws, int* __restrict__ info) {
  __shared__ double sv[EIG_MAXL];
  __shared__ double sp[EIG_MAXL];
  __shared__ double sh[EIG_THREADS / 64];
  __shared__ double s_bnd[2];
  syevx_block(A, l, ld, nev, evals, evecs, ws, info, sv, sp, sh, s_bnd, EIG_MAXL);
}

// ---- multi-workgroup Householder reduction (large l) -----------------------
// Column j is reduced by two launches over all workgroups: hh_mv forms
// p = t A22 v (one wave per row), hh_up applies A22 -= v w^T + w v^T with
// w = p - (t/2)(p.v) v (rows spread over waves) and stores v in column j.  Every
// workgroup rebuilds v from the still unmodified column j in LDS, and the kernel
// boundary orders the updates, so no grid-wide synchronisation is needed.
#define HH_THREADS 512

__global__ __launch_bounds__(256) void sym_copy_kernel(const double* __restrict__ A, int l, int ld, double* __restrict__ W) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    const long r = i / l, c = i % l;
    W[i] = 0.5 * (A[r * ld + c] + A[c * ld + r]);
  }
}

// Householder vector of column j in LDS: returns t (0 when the column is already
// zero below the subdiagonal), beta via *beta_out; sv[0..n2) = v.
__device__ double hh_vector(const double* __restrict__ W, int l, int j, double* sv, double* sh, double* beta_out) {
  const int n2 = l - j - 1, tid = threadIdx.x;
  double part = 0.0;
  for (int r = tid; r < n2; r += blockDim.x) {
    const double x = W[(long)(j + 1 + r) * l + j];
    sv[r] = x;
    part += x * x;
  }
  const double alpha2 = blk_sum(part, sh);
  const double x0 = sv[0];
  __syncthreads();
  if (alpha2 <= 1e-300) {
    *beta_out = 0.0;
    return 0.0;
  }
  const double alpha = sqrt(alpha2);
  const double beta = (x0 >= 0.0) ? -alpha : alpha;
  const double v0 = x0 - beta;
  const double vtv = alpha2 - x0 * x0 + v0 * v0;
  __syncthreads();
  if (tid == 0) sv[0] = v0;
  __syncthreads();
  *beta_out = beta;
  return 2.0 / vtv;
}

__global__ __launch_bounds__(HH_THREADS) void hh_mv_kernel(const double* __restrict__ W, int l, int j,
                                                           double* __restrict__ P) {
  __shared__ double sv[EIG_MAXL];
  __shared__ double sh[HH_THREADS / 64];
  const int n2 = l - j - 1;
  double beta;
  const double t = hh_vector(W, l, j, sv, sh, &beta);
  if (t == 0.0) return;
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int r = blockIdx.x * nw + (threadIdx.x >> 6); r < n2; r += gridDim.x * nw) {
    const double* row = W + (long)(j + 1 + r) * l + (j + 1);
    double s = 0.0;
    for (int c = lane; c < n2; c += 64) s += row[c] * sv[c];
    s = wave_sum(s);
    if (lane == 0) P[r] = t * s;
  }
}

// Column j keeps its original entries while the launch runs (other workgroups
// rebuild v from it): rows r >= 1 of v equal them, and v_0 goes to hv0[j]; the
// reflectors' first entries are written back after the last column.
__global__ __launch_bounds__(HH_THREADS) void hh_up_kernel(double* __restrict__ W, int l, int j,
                                                           const double* __restrict__ P, double* __restrict__ d,
                                                           double* __restrict__ e, double* __restrict__ tau,
                                                           double* __restrict__ hv0) {
  __shared__ double sv[EIG_MAXL];
  __shared__ double sw[EIG_MAXL];
  __shared__ double sh[HH_THREADS / 64];
  const int n2 = l - j - 1, tid = threadIdx.x;
  double beta;
  const double t = hh_vector(W, l, j, sv, sh, &beta);
  if (blockIdx.x == 0 && tid == 0) {
    tau[j] = t;
    e[j] = beta;
    d[j] = W[(long)j * l + j];
    hv0[j] = (t == 0.0) ? 0.0 : sv[0];
  }
  const int lane = tid & 63, nw = blockDim.x >> 6;
  if (t == 0.0) {
    for (int r = 1 + blockIdx.x * blockDim.x + tid; r < n2; r += gridDim.x * blockDim.x)
      W[(long)(j + 1 + r) * l + j] = 0.0;
    return;
  }
  double part = 0.0;
  for (int r = tid; r < n2; r += blockDim.x) {
    const double pr = P[r];
    sw[r] = pr;
    part += pr * sv[r];
  }
  const double K = 0.5 * t * blk_sum(part, sh);
  for (int r = tid; r < n2; r += blockDim.x) sw[r] -= K * sv[r];
  __syncthreads();
  for (int r = blockIdx.x * nw + (tid >> 6); r < n2; r += gridDim.x * nw) {
    double* row = W + (long)(j + 1 + r) * l + (j + 1);
    const double vr = sv[r], wr = sw[r];
    for (int c = lane; c < n2; c += 64) row[c] -= vr * sw[c] + wr * sv[c];
  }
}

// Writes the reflectors' first entries back into column j and copies every
// reflector into row j of the upper triangle (W[j][j+1+r] = v_r), where the
// back-transformation reads them with unit stride.  Reads touch only the strict
// lower triangle below the subdiagonal, writes the subdiagonal and the upper
// triangle, so the launch has no read/write overlap.
__global__ void hh_rows_kernel(double* __restrict__ W, int l, const double* __restrict__ hv0) {
  const long total = (long)l * l;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int j = (int)(idx / l), i = (int)(idx % l);
    if (j + 2 >= l || i <= j) continue;
    if (i == j + 1) {
      W[(long)(j + 1) * l + j] = hv0[j];
      W[(long)j * l + i] = hv0[j];
    } else {
      W[(long)j * l + i] = W[(long)i * l + j];
    }
  }
}

__global__ __launch_bounds__(EIG_THREADS) void syevx_vectors_kernel(int l, int nev, double* __restrict__ evals,
                                                                     char* __restrict__ ws) {
  __shared__ double sv[EIG_MAXL];
  __shared__ double sp[EIG_MAXL];
  __shared__ double sh[EIG_THREADS / 64];
  __shared__ double s_bnd[2];
  syevx_tridiag_vectors(l, nev, evals, ws, sv, sp, sh, s_bnd, EIG_MAXL);
}

__global__ __launch_bounds__(EIG_THREADS) void syevx_back_kernel(int l, int nev, double* __restrict__ evecs,
                                                                  char* __restrict__ ws) {
  __shared__ double sv[EIG_MAXL];
  syevx_back_block(l, nev, evecs, ws, sv);
}

// Back-transformation z <- H_0 ... H_{l-3} z with one wavefront per eigenvector:
// entry i of the vector lives in lane i % 64, slot i / 64 (registers), reflector
// j is read from row j of the upper triangle, and the next reflector's loads are
// issued before the current one's reduction.  R = ceil(l / 64) slots.
template <int R>
__global__ __launch_bounds__(64) void hh_back_kernel(const double* __restrict__ W, int l, int nev,
                                                     const double* __restrict__ tau, const double* __restrict__ Z,
                                                     double* __restrict__ evecs) {
  const int m = blockIdx.x, lane = threadIdx.x;
  double z[R], vn[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    z[q] = (i < l) ? Z[(long)i * nev + m] : 0.0;
  }
  int j = l - 3;
  double tn = 0.0;
  if (j >= 0) {
    tn = tau[j];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      vn[q] = (i > j && i < l) ? W[(long)j * l + i] : 0.0;
    }
  }
  for (; j >= 0; --j) {
    double v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = vn[q];
    const double t = tn;
    if (j > 0) {
      tn = tau[j - 1];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        vn[q] = (i > j - 1 && i < l) ? W[(long)(j - 1) * l + i] : 0.0;
      }
    }
    if (t == 0.0) continue;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) s += v[q] * z[q];
    s = t * wave_sum(s);
#pragma unroll
    for (int q = 0; q < R; ++q) z[q] -= s * v[q];
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    if (i < l) evecs[(long)i * nev + m] = z[q];
  }
}

extern "C" int dftb_syevx(const double* A, int l, int ld, int nev, double* evals, double* evecs, void* workspace,
                          size_t ws_bytes, void* stream) {
  if (l <= 0 || nev <= 0 || nev > l || l > EIG_MAXL || ld < l || !A || !evals || !evecs || !workspace)
    return DFTB_ERR_ARG;
  EigLayout Lo = eig_layout(l, nev);
  if (ws_bytes < Lo.total) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  if (l < 160) {
    hipLaunchKernelGGL(syevx_kernel, dim3(1), dim3(EIG_THREADS), 0, st, A, l, ld, nev, evals, evecs, ws,
                       (int*)nullptr);
    DFTB_CHECK_LAUNCH();
    return DFTB_OK;
  }
  double* W = (double*)(ws + Lo.W);
  double* d = (double*)(ws + Lo.d);
  double* e = (double*)(ws + Lo.e);
  double* tau = (double*)(ws + Lo.tau);
  double* P = (double*)(ws + Lo.scratch);   // product vector (the tail's scratch is not yet in use)
  double* hv0 = P + l;                       // first entries of the reflectors
  hipLaunchKernelGGL(sym_copy_kernel, dim3(ceil_div((long)l * l, 256) < 4096 ? ceil_div((long)l * l, 256) : 4096),
                     dim3(256), 0, st, A, l, ld, W);
  for (int j = 0; j + 2 < l; ++j) {
    const int n2 = l - j - 1;
    const int waves = HH_THREADS / 64;
    int g = (n2 + waves - 1) / waves;       // one wave per row
    if (g > 256) g = 256;
    hipLaunchKernelGGL(hh_mv_kernel, dim3(g), dim3(HH_THREADS), 0, st, (const double*)W, l, j, P);
    hipLaunchKernelGGL(hh_up_kernel, dim3(g), dim3(HH_THREADS), 0, st, W, l, j, (const double*)P, d, e, tau, hv0);
  }
  hipLaunchKernelGGL(hh_rows_kernel, dim3(ceil_div((long)l * l, 256) < 4096 ? ceil_div((long)l * l, 256) : 4096),
                     dim3(256), 0, st, W, l, (const double*)hv0);
  hipLaunchKernelGGL(syevx_vectors_kernel, dim3(1), dim3(EIG_THREADS), 0, st, l, nev, evals, ws);
  const double* Z = (const double*)(ws + Lo.z);
  const int R = (l + 63) / 64;
  if (R <= 4)
    hipLaunchKernelGGL(hh_back_kernel<4>, dim3(nev), dim3(64), 0, st, (const double*)W, l, nev, (const double*)tau, Z, evecs);
  else if (R <= 8)
    hipLaunchKernelGGL(hh_back_kernel<8>, dim3(nev), dim3(64), 0, st, (const double*)W, l, nev, (const double*)tau, Z, evecs);
  else if (R <= 16)
    hipLaunchKernelGGL(hh_back_kernel<16>, dim3(nev), dim3(64), 0, st, (const double*)W, l, nev, (const double*)tau, Z, evecs);
  else if (R <= 32)
    hipLaunchKernelGGL(hh_back_kernel<32>, dim3(nev), dim3(64), 0, st, (const double*)W, l, nev, (const double*)tau, Z, evecs);
  else
    hipLaunchKernelGGL(syevx_back_kernel, dim3(1), dim3(EIG_THREADS), 0, st, l, nev, evecs, ws);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// ---------------------------------------------------------------------------
// Subspace steps
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Matrix square root by the coupled Newton-Schulz iteration (all MFMA GEMMs):
//   Y0 = A/c, Z0 = I;  T = (3I - Z Y)/2;  Y <- Y T;  Z <- T Z
// converges to Y = (A/c)^{1/2}, Z = (A/c)^{-1/2} for SPD A with spectrum in (0, c].
// Replaces scipy.linalg.sqrtm(Mm) of Solver.py:762.
// ---------------------------------------------------------------------------
__global__ void ns_init_kernel(const double* A, int l, int ld, double* Y, double* Z, const double* scale) {
  const double c = scale[0];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    Y[i] = 0.5 * (A[(long)r * ld + q] + A[(long)q * ld + r]) / c;
    Z[i] = (r == q) ? 1.0 : 0.0;
  }
}

// scale[0] = Frobenius norm of the symmetric part (>= largest eigenvalue)
__global__ __launch_bounds__(256) void fro_kernel(const double* A, int l, int ld, double* scale) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    double a = A[(long)r * ld + q];
    s += a * a;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) scale[0] = fmax(sqrt(sh[0] + sh[1] + sh[2] + sh[3]), 1e-300);
}

// T = (3I - P)/2 ; err[0] = max |P - I|
__global__ __launch_bounds__(1024) void ns_t_kernel(const double* P, int l, double* T, double* err) {
  __shared__ double sh[16];
  double m = 0.0;
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    double id = (r == q) ? 1.0 : 0.0;
    double p = P[i];
    m = fmax(m, fabs(p - id));
    T[i] = 0.5 * (3.0 * id - p);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) mm = fmax(mm, sh[i]);
    err[0] = mm;
  }
}

// out = X * sqrt(c) (sqrt) or X / sqrt(c) (inverse sqrt)
__global__ void ns_finish_kernel(double* X, long n, const double* scale, int inverse) {
  const double f = inverse ? 1.0 / sqrt(scale[0]) : sqrt(scale[0]);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) X[i] *= f;
}

static GemmArgs small_gemm(const double* A, int lda, bool at, const double* B, int ldb, double* C, int ldc, int M,
                           int N, int K) {
  GemmArgs g = gemm_args_default();
  g.A = A; g.lda = lda; g.a_trans = at;
  g.B = B; g.ldb = ldb;
  g.C = C; g.ldc = ldc;
  g.M = M; g.N = N; g.K = K; g.tile = 64;
  return g;
}

static size_t ns_workspace(int l) { return al256((size_t)l * l * 8) * 5 + 256 * 2; }

// Y <- A^{1/2}, Z <- A^{-1/2}; iters[0] = iterations used; returns DFTB_ERR_NOT_SPD if not converged.
static int sqrtm_ns(const double* A, int l, int ld, double* Y, double* Z, char* ws, hipStream_t st, int* iters_out,
                    double* err_out) {
  double* T = (double*)ws;
  double* P = (double*)(ws + al256((size_t)l * l * 8));
  double* Y2 = (double*)(ws + 2 * al256((size_t)l * l * 8));
  double* Z2 = (double*)(ws + 3 * al256((size_t)l * l * 8));
  double* scal = (double*)(ws + 5 * al256((size_t)l * l * 8));
  double* err = scal + 1;
  int rc;
  hipLaunchKernelGGL(fro_kernel, dim3(1), dim3(256), 0, st, A, l, ld, scal);
  hipLaunchKernelGGL(ns_init_kernel, dim3((l * l + 255) / 256), dim3(256), 0, st, A, l, ld, Y, Z, scal);
  DFTB_CHECK_LAUNCH();
  double herr = 1.0;
  int it = 0;
  const int maxit = 100;
  for (it = 0; it < maxit; ++it) {
    if ((rc = gemm_launch(small_gemm(Z, l, false, Y, l, P, l, l, l, l), st))) return rc;
    hipLaunchKernelGGL(ns_t_kernel, dim3(1), dim3(1024), 0, st, P, l, T, err);
    DFTB_CHECK_LAUNCH();
    if ((rc = gemm_launch(small_gemm(Y, l, false, T, l, Y2, l, l, l, l), st))) return rc;
    if ((rc = gemm_launch(small_gemm(T, l, false, Z, l, Z2, l, l, l, l), st))) return rc;
    if (hipMemcpyAsync(Y, Y2, (size_t)l * l * 8, hipMemcpyDeviceToDevice, st) != hipSuccess) return DFTB_ERR_LAUNCH;
    if (hipMemcpyAsync(Z, Z2, (size_t)l * l * 8, hipMemcpyDeviceToDevice, st) != hipSuccess) return DFTB_ERR_LAUNCH;
    if (it % 4 == 3) {
      if (hipMemcpyAsync(&herr, err, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) return DFTB_ERR_LAUNCH;
      if (hipStreamSynchronize(st) != hipSuccess) return DFTB_ERR_LAUNCH;
      if (!(herr == herr) || herr > 1e6) break;  // diverged: not SPD
      if (herr < 1e-14 * (l > 16 ? l : 16)) break;
    }
  }
  if (iters_out) *iters_out = it + 1;
  if (err_out) *err_out = herr;
  hipLaunchKernelGGL(ns_finish_kernel, dim3((l * l + 255) / 256), dim3(256), 0, st, Y, (long)l * l, scal, 0);
  hipLaunchKernelGGL(ns_finish_kernel, dim3((l * l + 255) / 256), dim3(256), 0, st, Z, (long)l * l, scal, 1);
  DFTB_CHECK_LAUNCH();
  if (!(herr < 1e-10)) return DFTB_ERR_NOT_SPD;
  return DFTB_OK;
}

extern "C" size_t dftb_sqrtm_workspace(int l) { return ns_workspace(l); }

extern "C" int dftb_sqrtm(const double* A, int l, int ld, double* Asq, double* Ainvsq, int* iters, double* err,
                          void* workspace, size_t ws_bytes, void* stream) {
  if (l <= 0 || ld < l || !A || !Asq || !Ainvsq || !workspace) return DFTB_ERR_ARG;
  if (ws_bytes < ns_workspace(l)) return DFTB_ERR_WORKSPACE;
  return sqrtm_ns(A, l, ld, Asq, Ainvsq, (char*)workspace, (hipStream_t)stream, iters, err);
}

// err = sum |Mh - Mh^T|  (Solver.py:766-770)
__global__ __launch_bounds__(256) void herm_err_kernel(const double* M, int l, double* diag) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    int r = (int)(i / l), c = (int)(i % l);
    s += fabs(M[(long)r * l + c] - M[(long)c * l + r]);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) diag[0] = sh[0] + sh[1] + sh[2] + sh[3];
}

// w = sqrt(w2) (flag non-positive w2); scales: sR[n] = 1/sqrt(w), sL[n] = 1/w
__global__ void roots_kernel(const double* w2, int nev, double* w, double* sR, double* sL, int* info) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nev) return;
  double x = w2[n];
  if (!(x > 0.0)) {
    info[0] = DFTB_ERR_NOT_SPD;
    info[1] = -1 - n;
    x = 1e-300;
  }
  double wn = sqrt(x);
  w[n] = wn;
  sR[n] = 1.0 / sqrt(wn);
  sL[n] = 1.0 / wn;
}

__global__ void scale_cols_kernel(double* X, int rows, int cols, const double* s) {
  long total = (long)rows * cols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    X[i] *= s[i % cols];
}

__global__ void biortho_kernel(const double* G, int nev, double* diag) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nev * nev; i += blockDim.x) {
    int r = i / nev, c = i % nev;
    s += fabs(G[i] - (r == c ? 1.0 : 0.0));
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) diag[1] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void zero_info_kernel(int* info) {
  info[0] = 0;
  info[1] = 0;
}

struct SubLayout {
  size_t Sq, Sinv, X, H, w2, sR, sL, G, eig, total;
};
static SubLayout sub_layout(int l, int nev) {
  SubLayout S;
  size_t off = 0;
  S.Sq = off; off += al256((size_t)l * l * 8);
  S.Sinv = off; off += al256((size_t)l * l * 8);
  S.X = off; off += al256((size_t)l * l * 8);
  S.H = off; off += al256((size_t)l * l * 8);
  S.w2 = off; off += al256((size_t)nev * 8);
  S.sR = off; off += al256((size_t)nev * 8);
  S.sL = off; off += al256((size_t)nev * 8);
  S.G = off; off += al256((size_t)nev * nev * 8);
  size_t e1 = eig_layout(l, nev).total, e2 = ns_workspace(l);
  S.eig = off; off += al256(e1 > e2 ? e1 : e2);
  S.total = off;
  return S;
}

extern "C" size_t dftb_subspace_workspace(int l, int nev) { return sub_layout(l, nev).total; }

extern "C" int dftb_subspace_nonhermitian(const double* Mp, const double* Mm, int l, int ld, int nev, double* w,
                                          double* Tb, double* Rb, double* Lb, int* info, double* diag,
                                          void* workspace, size_t ws_bytes, void* stream) {
  if (l <= 0 || nev <= 0 || nev > l || ld < l || l > EIG_MAXL) return DFTB_ERR_ARG;
  if (!Mp || !Mm || !w || !Tb || !Rb || !Lb || !info || !diag || !workspace) return DFTB_ERR_ARG;
  SubLayout S = sub_layout(l, nev);
  if (ws_bytes < S.total) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* Sq = (double*)(ws + S.Sq);
  double* Sinv = (double*)(ws + S.Sinv);
  double* X = (double*)(ws + S.X);
  double* H = (double*)(ws + S.H);
  double* w2 = (double*)(ws + S.w2);
  double* sR = (double*)(ws + S.sR);
  double* sL = (double*)(ws + S.sL);
  double* G = (double*)(ws + S.G);
  int rc;
  hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, st, info);
  DFTB_CHECK_LAUNCH();
  // Mmsq = sqrtm(Mm)
  int nsit = 0;
  double nserr = 0.0;
  rc = sqrtm_ns(Mm, l, ld, Sq, Sinv, ws + S.eig, st, &nsit, &nserr);
  if (rc == DFTB_ERR_NOT_SPD) {
    int h[2] = {DFTB_ERR_NOT_SPD, nsit};
    hipMemcpyAsync(info, h, sizeof(h), hipMemcpyHostToDevice, st);
    hipStreamSynchronize(st);
    return DFTB_OK;
  }
  if (rc) return rc;
  // Mh = Mmsq Mp Mmsq, hermiticity check
  if ((rc = gemm_launch(small_gemm(Mp, ld, false, Sq, l, X, l, l, l, l), st))) return rc;
  if ((rc = gemm_launch(small_gemm(Sq, l, false, X, l, H, l, l, l, l), st))) return rc;
  hipLaunchKernelGGL(herm_err_kernel, dim3(1), dim3(256), 0, st, H, l, diag);
  DFTB_CHECK_LAUNCH();
  if ((rc = dftb_syevx(H, l, l, nev, w2, Tb, ws + S.eig, eig_layout(l, nev).total, stream))) return rc;
  hipLaunchKernelGGL(roots_kernel, dim3((nev + 63) / 64), dim3(64), 0, st, w2, nev, w, sR, sL, info);
  DFTB_CHECK_LAUNCH();
  // Rb = Mmsq Tb / sqrt(w);  Lb = Mp Rb / w
  if ((rc = gemm_launch(small_gemm(Sq, l, false, Tb, nev, Rb, nev, l, nev, l), st))) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((l * nev + 255) / 256), dim3(256), 0, st, Rb, l, nev, sR);
  DFTB_CHECK_LAUNCH();
  if ((rc = gemm_launch(small_gemm(Mp, ld, false, Rb, nev, Lb, nev, l, nev, l), st))) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((l * nev + 255) / 256), dim3(256), 0, st, Lb, l, nev, sL);
  DFTB_CHECK_LAUNCH();
  // biorthonormality check Lb^T Rb = 1 (Solver.py:791-798)
  if ((rc = gemm_launch(small_gemm(Lb, nev, true, Rb, nev, G, nev, nev, nev, l), st))) return rc;
  hipLaunchKernelGGL(biortho_kernel, dim3(1), dim3(256), 0, st, G, nev, diag);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_subspace_hermitian(const double* H, int l, int ld, int nev, double* w2, double* T, int* info,
                                       void* workspace, size_t ws_bytes, void* stream) {
  if (l <= 0 || nev <= 0 || nev > l || ld < l || l > EIG_MAXL || !H || !w2 || !T || !info || !workspace)
    return DFTB_ERR_ARG;
  SubLayout S = sub_layout(l, nev);
  if (ws_bytes < S.total) return DFTB_ERR_WORKSPACE;
  hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, info);
  DFTB_CHECK_LAUNCH();
  return dftb_syevx(H, l, ld, nev, w2, T, (char*)workspace + S.eig, eig_layout(l, nev).total, stream);
}

// Replace the leading l x l block of M by its symmetric part; asym[0] = max |M - M^T| / max |M|.
__global__ __launch_bounds__(1024) void symmetrize_kernel(double* M, int l, int ld, double* asym) {
  __shared__ double sh[16], sm[16];
  double a = 0.0, mx = 0.0;
  for (long i = threadIdx.x; i < (long)l * l; i += blockDim.x) {
    int r = (int)(i / l), c = (int)(i % l);
    if (c > r) {
      double x = M[(long)r * ld + c], y = M[(long)c * ld + r];
      a = fmax(a, fabs(x - y));
      mx = fmax(mx, fmax(fabs(x), fabs(y)));
      double s = 0.5 * (x + y);
      M[(long)r * ld + c] = s;
      M[(long)c * ld + r] = s;
    } else if (c == r) {
      mx = fmax(mx, fabs(M[(long)r * ld + c]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o, 64));
    mx = fmax(mx, __shfl_xor(mx, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = a;
    sm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0.0, X = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      A = fmax(A, sh[i]);
      X = fmax(X, sm[i]);
    }
    asym[0] = X > 0.0 ? A / X : 0.0;
  }
}

extern "C" int dftb_symmetrize(double* M, int l, int ld, double* asym, void* stream) {
  if (l <= 0 || ld < l || !M || !asym) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(symmetrize_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, M, l, ld, asym);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
