// (c) Davidson support kernels: residuals, norms, preconditioner, guess
// vectors and block orthonormalisation (FP64 throughout).
//
// Reference: DFTB/Solver.py:809-897 (nonHermitianDavidson) and 480-530
// (HermitianDavidson).  These passes stream over l * n_ov doubles and are
// HBM-bound; norms use one workgroup per vector so reductions are
// deterministic.
#include "common.h"
#include "gemm.h"

static dim3 sgrid(long n) {
  long b = (n + 255) / 256;
  if (b > 32768) b = 32768;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    sh[0] = s;
  }
  __syncthreads();
  s = sh[0];
  __syncthreads();
  return s;
}

// Residuals and row norms stream k vectors of n doubles.  Each vector is split
// over NORM_SLICES workgroups (enough to keep every CU streaming for the k = 20
// to 40 vectors of a Davidson step); each slice writes its partial sum of
// squares, and a second kernel adds the slices of each vector in a fixed order
// (deterministic).  The caller's workspace holds k * NORM_SLICES partials.
#define NORM_SLICES 32

// W[j] = P[j] - w[j] X[j]; partial[j][s] = sum over slice s of W[j]^2
__global__ __launch_bounds__(256) void residual_kernel(const double* __restrict__ P, const double* __restrict__ X,
                                                       const double* __restrict__ w, long n, double* __restrict__ W,
                                                       double* __restrict__ partial) {
  __shared__ double sh[4];
  const long j = blockIdx.y;
  const long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long b = blockIdx.x * chunk;
  const long e = b + chunk < n ? b + chunk : n;
  const double wj = w[j];
  const double* p = P + j * n;
  const double* x = X + j * n;
  double* o = W + j * n;
  double s = 0.0;
  for (long i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double r = p[i] - wj * x[i];
    o[i] = r;
    s += r * r;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[j * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void row_norms_kernel(const double* __restrict__ X, long n,
                                                        double* __restrict__ partial) {
  __shared__ double sh[4];
  const long j = blockIdx.y;
  const long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long b = blockIdx.x * chunk;
  const long e = b + chunk < n ? b + chunk : n;
  const double* x = X + j * n;
  double s = 0.0;
  for (long i = b + threadIdx.x; i < e; i += blockDim.x) s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[j * gridDim.x + blockIdx.x] = s;
}

// norms[j] = sqrt(sum_s partial[j][s]), slices added in order
__global__ void slice_norms_kernel(const double* __restrict__ partial, int k, int slices, double* __restrict__ norms) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double s = 0.0;
  for (int t = 0; t < slices; ++t) s += partial[(long)j * slices + t];
  norms[j] = sqrt(s);
}

__global__ void precondition_kernel(const double* __restrict__ W, const double* __restrict__ w,
                                    const int* __restrict__ sel, int nsel, const double* __restrict__ omega, long n,
                                    double* __restrict__ Q) {
  const long total = (long)nsel * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long t = idx / n, e = idx - t * n;
    const int src = sel[t];
    const double om = omega[e];
    double d = w[src] - om;
    if (fabs(d) < 1.0e-6) d = 1.0e-6 * om;
    Q[idx] = W[(long)src * n + e] / d;
  }
}

__global__ void scale_rows_kernel(double* X, const double* s, int k, long n) {
  const long total = (long)k * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x)
    X[idx] *= s[idx / n];
}

__global__ void unit_vectors_kernel(const long* idx, int k, long n, double* X) {
  const long total = (long)k * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long t = i / n, e = i - t * n;
    X[i] = (e == idx[t]) ? 1.0 : 0.0;
  }
}

extern "C" size_t dftb_norms_workspace(int k) { return (size_t)(k > 0 ? k : 0) * NORM_SLICES * sizeof(double); }

extern "C" int dftb_residual(const double* P, const double* X, const double* w, int k, long n, double* W, double* norms,
                             void* workspace, size_t ws_bytes, void* stream) {
  if (k <= 0) return DFTB_OK;
  if (!P || !X || !w || !W || !norms || !workspace || n <= 0) return DFTB_ERR_ARG;
  if (ws_bytes < dftb_norms_workspace(k)) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(residual_kernel, dim3(NORM_SLICES, k), dim3(256), 0, st, P, X, w, n, W, part);
  hipLaunchKernelGGL(slice_norms_kernel, dim3((k + 63) / 64), dim3(64), 0, st, (const double*)part, k, NORM_SLICES,
                     norms);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_row_norms(const double* X, int k, long n, double* norms, void* workspace, size_t ws_bytes,
                              void* stream) {
  if (k <= 0) return DFTB_OK;
  if (!X || !norms || !workspace || n <= 0) return DFTB_ERR_ARG;
  if (ws_bytes < dftb_norms_workspace(k)) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)workspace;
  hipLaunchKernelGGL(row_norms_kernel, dim3(NORM_SLICES, k), dim3(256), 0, st, X, n, part);
  hipLaunchKernelGGL(slice_norms_kernel, dim3((k + 63) / 64), dim3(64), 0, st, (const double*)part, k, NORM_SLICES,
                     norms);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_precondition(const double* W, const double* w, const int* sel, int nsel, const double* omega,
                                 long n, double* Q, void* stream) {
  if (nsel <= 0) return DFTB_OK;
  if (!W || !w || !sel || !omega || !Q) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(precondition_kernel, sgrid((long)nsel * n), dim3(256), 0, (hipStream_t)stream, W, w, sel, nsel,
                     omega, n, Q);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_scale_rows(double* X, const double* s, int k, long n, void* stream) {
  if (k <= 0) return DFTB_OK;
  hipLaunchKernelGGL(scale_rows_kernel, sgrid((long)k * n), dim3(256), 0, (hipStream_t)stream, X, s, k, n);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_unit_vectors(const long* idx, int k, long n, double* X, void* stream) {
  if (k <= 0) return DFTB_OK;
  if (!idx || !X) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(unit_vectors_kernel, sgrid((long)k * n), dim3(256), 0, (hipStream_t)stream, idx, k, n, X);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// ---------------------------------------------------------------------------
// Block orthonormalisation
// ---------------------------------------------------------------------------
// Cholesky-QR factor of the Gram matrix G = Q Q^T (m x m) with rank dropping:
// a row whose pivot falls below tol2 * ref[j] is flagged and zeroed.  Writes
// Linv = L^{-1} (rows of dropped vectors zero).  One workgroup; G, L and Linv in
// LDS when 3 m^2 doubles fit (CHOLQR_LDS_M), else in global memory.
#define CHOLQR_LDS_M 52
__device__ void cholqr_core(const double* G, int m, const double* __restrict__ ref, double tol2, double* L,
                            double* Linv, int* __restrict__ keep, int* s_keep) {
  for (int i = threadIdx.x; i < m * m; i += blockDim.x) {
    L[i] = 0.0;
    Linv[i] = 0.0;
  }
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    if (threadIdx.x == 0) {
      double d = G[j * m + j];
      for (int k = 0; k < j; ++k) d -= L[j * m + k] * L[j * m + k];
      double r = ref ? ref[j] * ref[j] : 1.0;
      // the pivot of a dependent row is Cholesky cancellation noise ~ eps * G_jj:
      // require it to clear both the caller's tolerance and 1e-12 * G_jj
      int ok = keep[j] && (d > tol2 * r) && (d > 1.0e-12 * G[j * m + j]) && (d > 0.0);
      *s_keep = ok;
      keep[j] = ok;
      L[j * m + j] = ok ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double ljj = L[j * m + j];
    const int kj = *s_keep;
    for (int i = j + 1 + threadIdx.x; i < m; i += blockDim.x) {
      double s = 0.0;
      if (kj) {
        s = G[i * m + j];
        for (int k = 0; k < j; ++k) s -= L[i * m + k] * L[j * m + k];
        s /= ljj;
      }
      L[i * m + j] = s;
    }
    __syncthreads();
  }
  // Linv by forward substitution, one column per thread
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    for (int i = c; i < m; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= L[i * m + k] * Linv[k * m + c];
      Linv[i * m + c] = s / L[i * m + i];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m * m; i += blockDim.x) {
    int r = i / m;
    if (!keep[r]) Linv[i] = 0.0;
  }
}

__global__ __launch_bounds__(256) void cholqr_kernel(const double* __restrict__ G, int m, const double* __restrict__ ref,
                                                     double tol2, double* __restrict__ L, double* __restrict__ Linv,
                                                     int* __restrict__ keep) {
  __shared__ int s_keep;
  cholqr_core(G, m, ref, tol2, L, Linv, keep, &s_keep);
}

__global__ __launch_bounds__(256) void cholqr_lds_kernel(const double* __restrict__ G, int m,
                                                         const double* __restrict__ ref, double tol2,
                                                         double* __restrict__ L, double* __restrict__ Linv,
                                                         int* __restrict__ keep) {
  __shared__ double sm[3 * CHOLQR_LDS_M * CHOLQR_LDS_M];
  __shared__ int s_keep;
  double* sG = sm;
  double* sL = sm + m * m;
  double* sLi = sL + m * m;
  for (int i = threadIdx.x; i < m * m; i += blockDim.x) sG[i] = G[i];
  __syncthreads();
  cholqr_core(sG, m, ref, tol2, sL, sLi, keep, &s_keep);
  __syncthreads();
  for (int i = threadIdx.x; i < m * m; i += blockDim.x) {
    L[i] = sL[i];
    Linv[i] = sLi[i];
  }
}

static void cholqr_launch(const double* G, int m, const double* ref, double tol2, double* L, double* Linv, int* keep,
                          hipStream_t st) {
  if (m <= CHOLQR_LDS_M)
    hipLaunchKernelGGL(cholqr_lds_kernel, dim3(1), dim3(256), 0, st, G, m, ref, tol2, L, Linv, keep);
  else
    hipLaunchKernelGGL(cholqr_kernel, dim3(1), dim3(256), 0, st, G, m, ref, tol2, L, Linv, keep);
}

// dst[i][e] = sum_{k <= i} Linv[i][k] src[k][e] for a lower-triangular m x m Linv
// (m <= MP): one column e per thread with src[0..m)[e] in registers; the entries
// of Linv have workgroup-uniform addresses (scalar loads, no LDS traffic).  Streams
// src once and writes dst once.
template <int MP>
__global__ __launch_bounds__(256) void tri_apply_kernel(const double* __restrict__ Linv, int m,
                                                        const double* __restrict__ src, double* __restrict__ dst,
                                                        long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double x[MP];
#pragma unroll
    for (int k = 0; k < MP; ++k) x[k] = (k < m) ? src[(long)k * n + e] : 0.0;
    for (int i = 0; i < m; ++i) {
      const double* li = Linv + i * m;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < MP; ++k)
        if (k <= i) s += li[k] * x[k];       // uniform branch: the triangle only
      dst[(long)i * n + e] = s;
    }
  }
}

// dst = Linv src (Linv m x m lower triangular, src / dst m x n, distinct)
static int tri_apply(const double* Linv, int m, const double* src, double* dst, long n, hipStream_t st) {
  long g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  const dim3 grid((unsigned)g), blk(256);
  if (m <= 8) hipLaunchKernelGGL(tri_apply_kernel<8>, grid, blk, 0, st, Linv, m, src, dst, n);
  else if (m <= 16) hipLaunchKernelGGL(tri_apply_kernel<16>, grid, blk, 0, st, Linv, m, src, dst, n);
  else if (m <= 24) hipLaunchKernelGGL(tri_apply_kernel<24>, grid, blk, 0, st, Linv, m, src, dst, n);
  else if (m <= 32) hipLaunchKernelGGL(tri_apply_kernel<32>, grid, blk, 0, st, Linv, m, src, dst, n);
  else if (m <= 40) hipLaunchKernelGGL(tri_apply_kernel<40>, grid, blk, 0, st, Linv, m, src, dst, n);
  else if (m <= 48) hipLaunchKernelGGL(tri_apply_kernel<48>, grid, blk, 0, st, Linv, m, src, dst, n);
  else {
    GemmArgs u = gemm_args_default();
    u.A = Linv; u.lda = m;
    u.B = src; u.ldb = n;
    u.C = dst; u.ldc = n;
    u.M = m; u.N = (int)n; u.K = m; u.tile = 64;
    return gemm_launch(u, st);
  }
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

__global__ void init_keep_kernel(int* keep, int m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) keep[i] = 1;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct OrthoLayout {
  int ks;
  long kchunk;
  size_t part, C, G, L, Linv, ref, flag, nrm, tmp, total;
};

static OrthoLayout ortho_layout(int l, int m, long n) {
  OrthoLayout o;
  // enough K slices to fill the GPU (>= ~2048 workgroups of 64 x 64 tiles of the l x m
  // projection) with chunks of >= 2048 elements
  const long tiles = (long)((l > m ? l : m) + 63) / 64 * ((m + 63) / 64);
  long ks = (n + 2047) / 2048;
  const long want = (2048 + tiles - 1) / tiles;
  if (ks > want) ks = want;
  if (ks > 512) ks = 512;
  if (ks < 1) ks = 1;
  o.ks = (int)ks;
  o.kchunk = ((n + o.ks - 1) / o.ks + 31) / 32 * 32;
  o.ks = (int)((n + o.kchunk - 1) / o.kchunk);
  long lm = (long)(l > m ? l : m) * m;
  size_t off = 0;
  o.part = off; off += al256((size_t)o.ks * lm * 8);
  o.C = off; off += al256((size_t)lm * 8);
  o.G = off; off += al256((size_t)m * m * 8);
  o.L = off; off += al256((size_t)m * m * 8);
  o.Linv = off; off += al256((size_t)m * m * 8);
  o.ref = off; off += al256((size_t)m * 8);
  o.flag = off; off += al256(sizeof(int));
  o.nrm = off; off += al256((size_t)m * NORM_SLICES * 8);
  o.tmp = off; off += al256((size_t)m * n * 8);
  o.total = off;
  return o;
}

extern "C" size_t dftb_orthonormalize_workspace(int l, int m, long n) { return ortho_layout(l, m, n).total; }

// G[p][q] = sum_x X[p][x] Y[q][x]   (P x Q, split-K + deterministic reduction)
static int gram(const double* X, int P, const double* Y, int Q, long n, const OrthoLayout& o, char* ws, double* out,
                hipStream_t st) {
  GemmArgs g = gemm_args_default();
  g.A = X; g.lda = n;
  g.B = Y; g.ldb = n; g.b_trans = 1;
  g.C = ws + o.part; g.ldc = Q; g.sCk = (long)P * Q;
  g.M = P; g.N = Q; g.K = (int)n;
  g.ksplit = o.ks; g.kchunk = (int)o.kchunk; g.tile = 64;
  int rc = gemm_launch(g, st);
  if (rc) return rc;
  return gemm_reduce_launch((double*)(ws + o.part), (long)P * Q, o.ks, out, (long)P * Q, 1.0, 0.0, st);
}

// out[p * ldo + q] = sum_s part[s][p][q]  (slices added in order)
__global__ void gram_reduce_kernel(const double* __restrict__ part, int ks, int P, int Q, double* __restrict__ out,
                                   long ldo) {
  const long PQ = (long)P * Q;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < PQ; i += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int t = 0; t < ks; ++t) s += part[t * PQ + i];
    const long p = i / Q, q = i - p * Q;
    out[p * ldo + q] = s;
  }
}

static OrthoLayout gram_layout(int P, int Q, long n) {
  // enough K slices to fill the GPU (>= ~2048 workgroups of 64 x 64 tiles) with
  // chunks of >= 2048 elements, so the small operand's chunk stays in L2 across
  // the row tiles
  OrthoLayout o = {};
  const long tiles = (long)((P + 63) / 64) * ((Q + 63) / 64);
  long ks = (n + 2047) / 2048;
  const long want = (2048 + tiles - 1) / tiles;
  if (ks > want) ks = want;
  if (ks > 512) ks = 512;
  if (ks < 1) ks = 1;
  o.kchunk = ((n + ks - 1) / ks + 31) / 32 * 32;
  o.ks = (int)((n + o.kchunk - 1) / o.kchunk);
  o.part = 0;
  o.total = al256((size_t)o.ks * P * Q * 8);
  return o;
}

extern "C" size_t dftb_gram_workspace(int P, int Q, long n) {
  if (P <= 0 || Q <= 0 || n <= 0) return 0;
  return gram_layout(P, Q, n).total;
}

extern "C" int dftb_gram(const double* X, long ldx, int P, const double* Y, long ldy, int Q, long n, double* out,
                         long ldo, void* workspace, size_t ws_bytes, void* stream) {
  if (P <= 0 || Q <= 0) return DFTB_OK;
  if (!X || !Y || !out || n <= 0 || ldx < n || ldy < n || ldo < Q || !workspace) return DFTB_ERR_ARG;
  OrthoLayout o = gram_layout(P, Q, n);
  if (ws_bytes < o.total) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  GemmArgs g = gemm_args_default();
  g.A = X; g.lda = ldx;
  g.B = Y; g.ldb = ldy; g.b_trans = 1;
  g.C = workspace; g.ldc = Q; g.sCk = (long)P * Q;
  g.M = P; g.N = Q; g.K = (int)n;
  g.ksplit = o.ks; g.kchunk = (int)o.kchunk; g.tile = 64;
  int rc = gemm_launch(g, st);
  if (rc) return rc;
  const long PQ = (long)P * Q;
  long blocks = (PQ + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const double*)workspace, o.ks, P,
                     Q, out, ldo);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// flag = 1 when some vector kept less than half of its squared norm through the
// projection (G_jj < ref_j^2 / 2): the Daniel-Gragg-Kaufman-Stewart test for a
// second classical Gram-Schmidt pass
__global__ void dgks_kernel(const double* __restrict__ G, int m, const double* __restrict__ ref, int* __restrict__ flag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < m; j += blockDim.x)
    if (G[(long)j * m + j] < 0.5 * ref[j] * ref[j]) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) flag[0] = any;
}

extern "C" int dftb_orthonormalize(const double* basis, int l, double* Q, int m, long n, double tol, int* keep,
                                   void* workspace, size_t ws_bytes, void* stream) {
  if (m <= 0) return DFTB_OK;
  if (!Q || !keep || !workspace || n <= 0 || l < 0 || (l > 0 && !basis)) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  OrthoLayout o = ortho_layout(l, m, n);
  if (ws_bytes < o.total) return DFTB_ERR_WORKSPACE;
  char* ws = (char*)workspace;
  double* C = (double*)(ws + o.C);
  double* G = (double*)(ws + o.G);
  double* L = (double*)(ws + o.L);
  double* Linv = (double*)(ws + o.Linv);
  double* ref = (double*)(ws + o.ref);
  int* flag = (int*)(ws + o.flag);
  double* tmp = (double*)(ws + o.tmp);
  int rc;
  hipLaunchKernelGGL(init_keep_kernel, dim3((m + 255) / 256), dim3(256), 0, st, keep, m);
  DFTB_CHECK_LAUNCH();
  // reference norms before the projections
  hipLaunchKernelGGL(row_norms_kernel, dim3(NORM_SLICES, m), dim3(256), 0, st, (const double*)Q, n,
                     (double*)(ws + o.nrm));
  hipLaunchKernelGGL(slice_norms_kernel, dim3((m + 63) / 64), dim3(64), 0, st, (const double*)(ws + o.nrm), m,
                     NORM_SLICES, ref);
  DFTB_CHECK_LAUNCH();
  // classical Gram-Schmidt against the basis; a second pass when the DGKS test
  // asks for it (G = Q Q^T of the projected block serves the test and the first
  // Cholesky-QR)
  for (int pass = 0; pass < 2 && l > 0; ++pass) {
    if ((rc = gram(basis, l, Q, m, n, o, ws, C, st))) return rc;  // C[p][j] = <basis_p, Q_j>
    GemmArgs u = gemm_args_default();
    u.A = C; u.lda = m; u.a_trans = 1;  // (j, p) -> C[p][j]
    u.B = basis; u.ldb = n;
    u.C = Q; u.ldc = n;
    u.M = m; u.N = (int)n; u.K = l; u.alpha = -1.0; u.beta = 1.0; u.tile = 64;
    if ((rc = gemm_launch(u, st))) return rc;
    if ((rc = gram(Q, m, Q, m, n, o, ws, G, st))) return rc;
    if (pass == 0) {
      hipLaunchKernelGGL(dgks_kernel, dim3(1), dim3(64), 0, st, (const double*)G, m, (const double*)ref, flag);
      DFTB_CHECK_LAUNCH();
      int hflag = 1;
      if (hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return DFTB_ERR_LAUNCH;
      if (hipStreamSynchronize(st) != hipSuccess) return DFTB_ERR_LAUNCH;
      if (!hflag) break;
    }
  }
  // without a basis the first Cholesky-QR needs the block's own Gram matrix
  if (l == 0 && (rc = gram(Q, m, Q, m, n, o, ws, G, st))) return rc;
  // two passes of Cholesky-QR inside the block: Q -> tmp -> Q
  for (int pass = 0; pass < 2; ++pass) {
    const double* src = pass == 0 ? Q : tmp;
    double* dst = pass == 0 ? tmp : Q;
    if (pass == 1 && (rc = gram(src, m, src, m, n, o, ws, G, st))) return rc;
    cholqr_launch(G, m, pass == 0 ? ref : (const double*)nullptr, pass == 0 ? tol * tol : 1.0e-24, L, Linv, keep, st);
    DFTB_CHECK_LAUNCH();
    if ((rc = tri_apply(Linv, m, src, dst, n, st))) return rc;
  }
  return DFTB_OK;
}

// ---------------------------------------------------------------------------
// Small elementwise helpers for the host driver
// ---------------------------------------------------------------------------
// out[j][e] = X[j][e] * (rs ? rs[j] : 1) * (cs ? cs[e] : 1)
__global__ void scale2_kernel(const double* __restrict__ X, const double* __restrict__ rs,
                              const double* __restrict__ cs, int k, long n, double* __restrict__ out) {
  const long total = (long)k * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long j = idx / n, e = idx - j * n;
    double v = X[idx];
    if (rs) v *= rs[j];
    if (cs) v *= cs[e];
    out[idx] = v;
  }
}

// s[j] = sum_e X[j][e]^2 * wgt[e]   (Lambda2 = sum_ia C_ia^2 O_ia, LR_TDDFTB.py:561-565)
__global__ __launch_bounds__(1024) void wsumsq_kernel(const double* __restrict__ X, const double* __restrict__ wgt,
                                                      long n, double* __restrict__ s) {
  __shared__ double sh[16];
  const long j = blockIdx.x;
  const double* x = X + j * n;
  double acc = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) acc += x[i] * x[i] * wgt[i];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) s[j] = acc;
}

// O[i][a] /= sqrt(Oii[i] * Oaa[a])  on the padded [opad][vpad] grid (padding -> 0)
__global__ void oia_norm_kernel(double* O, const double* Oii, const double* Oaa, int nocc, int nvirt, int vpad,
                                long total) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / vpad), a = (int)(idx % vpad);
    O[idx] = (i < nocc && a < nvirt) ? O[idx] / sqrt(Oii[i] * Oaa[a]) : 0.0;
  }
}

extern "C" int dftb_scale2(const double* X, const double* rs, const double* cs, int k, long n, double* out,
                           void* stream) {
  if (k <= 0 || n <= 0) return DFTB_OK;
  if (!X || !out) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(scale2_kernel, sgrid((long)k * n), dim3(256), 0, (hipStream_t)stream, X, rs, cs, k, n, out);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_weighted_sumsq(const double* X, const double* wgt, int k, long n, double* s, void* stream) {
  if (k <= 0) return DFTB_OK;
  if (!X || !wgt || !s) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(wsumsq_kernel, dim3(k), dim3(1024), 0, (hipStream_t)stream, X, wgt, n, s);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_oia_normalize(double* O, const double* Oii, const double* Oaa, int nocc, int nvirt, int opad,
                                  int vpad, void* stream) {
  long total = (long)opad * vpad;
  hipLaunchKernelGGL(oia_norm_kernel, sgrid(total), dim3(256), 0, (hipStream_t)stream, O, Oii, Oaa, nocc, nvirt, vpad,
                     total);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
