// (c) FP64 subspace eigensolver and the Davidson subspace steps.
//
// The reference solves the l x l subspace problems with LAPACK
// (scipy.linalg.sqrtm + numpy.linalg.eigh, DFTB/Solver.py:762-772, and eigh at
// Solver.py:484).  Only the lowest roots are used, so the engine computes the
// lowest nev eigenpairs on the GPU:
//   1. Householder reduction to tridiagonal form (one workgroup; the matrix
//      stays L2-resident, the Householder vector and the product vector in LDS),
//   2. Sturm-sequence multisection of the wanted eigenvalues (several points per
//      eigenvalue and round),
//   3. inverse iteration with partial-pivot tridiagonal LU, vectors of close
//      eigenvalues (gap < 1e-3 ||T||) re-orthogonalised within their cluster,
//   4. back-transformation with the stored reflectors.
// The non-Hermitian step follows Solver.py:759-798: Mmsq = sqrtm(Mm) (coupled
// Newton-Schulz iteration on MFMA GEMMs), Mh = Mmsq Mp Mmsq, the lowest nev
// eigenpairs of Mh, Rb = Mmsq Tb / sqrt(w), Lb = Mp Rb / w.  The Davidson
// iterations take the same roots, Rb and Lb from a blocked Cholesky factor of Mm
// (dftb_subspace_nonhermitian_rl) and Tb once at the end (dftb_subspace_tb).
// For l >= EIG_MULTI_L the tridiagonalisation runs over many workgroups, one
// launch per column; the multisection and inverse iteration stay in one.
#include <math.h>
#include "common.h"
#include "gemm.h"
#include "blockops.h"

#define EIG_THREADS 512
#define EIG_MAXL 4096
#define EIG_MULTI_L 64   // from this size on: multi-workgroup reduction, then the one-workgroup tail

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
// One launch per column j (hh_step) finishes step j and starts step j + 1:
//   - every workgroup rebuilds v_j from the updated column j (col, written by the
//     previous launch) and w = p - (t/2)(p.v) v from p = t A22 v (previous launch);
//   - every workgroup forms the updated column j + 1 itself (from row j + 1, by
//     symmetry) and v_{j+1} from it; workgroup 0 stores it for the next launch;
//   - each wave updates one row of the trailing block, A -= v w^T + w v^T, for the
//     columns beyond j + 1, and with the values it has just written forms that
//     row's entry of p' = t' A22' v_{j+1} for the next launch.
// Nothing reads what the launch writes: column / row j + 1 are never written
// (their updated values travel in col), p and col alternate between two buffers,
// and column j below the subdiagonal receives v_j (rows >= 1) for the
// back-transformation.  The update term (v_r w_c) + (w_r v_c) is evaluated
// without contraction, so entries (r, c) and (c, r) stay bitwise equal and row
// j + 1 is column j + 1 exactly.
#define HH_THREADS 512

__global__ __launch_bounds__(256) void sym_copy_kernel(const double* __restrict__ A, int l, int ld, double* __restrict__ W) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    const long r = i / l, c = i % l;
    W[i] = 0.5 * (A[r * ld + c] + A[c * ld + r]);
  }
}

// Householder vector of the n2 entries x (global or LDS) into sv: returns t (0 when
// the entries are all zero), beta via *beta_out.  Ends with the block synchronised.
__device__ double hh_vec(const double* x, int n2, double* sv, double* sh, double* beta_out) {
  const int tid = threadIdx.x;
  double part = 0.0;
  for (int r = tid; r < n2; r += blockDim.x) {
    const double a = x[r];
    sv[r] = a;
    part += a * a;
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
  if (tid == 0) sv[0] = v0;
  __syncthreads();
  *beta_out = beta;
  return 2.0 / vtv;
}

// (a b) + (c d) rounded without contraction: symmetric in swapping the pairs
__device__ __forceinline__ double sym_upd(double a, double b, double c, double d) {
  return __dadd_rn(__dmul_rn(a, b), __dmul_rn(c, d));
}

// Step 0's inputs: col = row 0 of W (= column 0), P[r] = t0 (A22 v0)_r.
__global__ __launch_bounds__(HH_THREADS) void hh_start_kernel(const double* __restrict__ W, int l,
                                                              double* __restrict__ col, double* __restrict__ P) {
  __shared__ double sv[EIG_MAXL];
  __shared__ double sh[HH_THREADS / 64];
  const int n2 = l - 1, tid = threadIdx.x;
  if (blockIdx.x == 0)
    for (int r = tid; r < l; r += blockDim.x) col[r] = W[r];
  double beta;
  const double t = hh_vec(W + 1, n2, sv, sh, &beta);
  const int lane = tid & 63, nw = blockDim.x >> 6;
  for (int r = blockIdx.x * nw + (tid >> 6); r < n2; r += gridDim.x * nw) {
    const double* row = W + (long)(1 + r) * l + 1;
    double s = 0.0;
    for (int c = lane; c < n2; c += 64) s += row[c] * sv[c];
    s = wave_sum(s);
    if (lane == 0) P[r] = t * s;
  }
}

// col[0] = updated (j, j), col[1 + r] = updated (j + 1 + r, j); P[r] = t_j (A22 v_j)_r
// (r < n2).  Writes coln (updated column j + 1 from its diagonal down) and Pn (n2 - 1).
// Workgroup bid of nblk takes rows 1 + (bid * waves + wave) + k * nblk * waves.
__device__ void hh_step_body(double* __restrict__ W, int l, int j, const double* __restrict__ col,
                             double* __restrict__ coln, const double* __restrict__ P, double* __restrict__ Pn,
                             double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau,
                             double* __restrict__ hv0, int bid, int nblk, double* sv, double* sw, double* sn,
                             double* sh) {
  const int n2 = l - j - 1, tid = threadIdx.x;
  double beta;
  const double t = hh_vec(col + 1, n2, sv, sh, &beta);
  if (bid == 0 && tid == 0) {
    tau[j] = t;
    e[j] = beta;
    d[j] = col[0];
    hv0[j] = (t == 0.0) ? 0.0 : sv[0];
  }
  __syncthreads();
  if (t == 0.0) {
    // nothing to annihilate: no update, and column j below the subdiagonal is zero
    for (int r = tid; r < n2; r += blockDim.x) {
      sv[r] = 0.0;
      sw[r] = 0.0;
    }
  } else {
    double part = 0.0;
    for (int r = tid; r < n2; r += blockDim.x) {
      const double pr = P[r];
      sw[r] = pr;
      part += pr * sv[r];
    }
    const double K = 0.5 * t * blk_sum(part, sh);
    for (int r = tid; r < n2; r += blockDim.x) sw[r] -= K * sv[r];
  }
  __syncthreads();
  // updated column j + 1 = updated row j + 1 (rows j + 1 ..)
  const double* rowp = W + (long)(j + 1) * l + (j + 1);
  const double v0 = sv[0], w0 = sw[0];
  double part = 0.0;
  for (int r = tid; r < n2; r += blockDim.x) {
    const double a = rowp[r] - sym_upd(v0, sw[r], w0, sv[r]);
    if (bid == 0) coln[r] = a;
    if (r >= 1) {
      sn[r - 1] = a;
      part += a * a;
    }
  }
  const double alpha2 = blk_sum(part, sh);
  const double x0 = sn[0];
  __syncthreads();
  double tn = 0.0;
  if (alpha2 > 1e-300) {
    const double alpha = sqrt(alpha2);
    const double bn = (x0 >= 0.0) ? -alpha : alpha;
    const double vn0 = x0 - bn;
    tn = 2.0 / (alpha2 - x0 * x0 + vn0 * vn0);
    if (tid == 0) sn[0] = vn0;
  }
  __syncthreads();
  const int lane = tid & 63, nw = blockDim.x >> 6;
  for (int r = 1 + bid * nw + (tid >> 6); r < n2; r += nblk * nw) {
    double* row = W + (long)(j + 1 + r) * l + (j + 1);
    const double vr = sv[r], wr = sw[r];
    double s = 0.0;
    for (int c = 1 + lane; c < n2; c += 64) {
      const double a = row[c] - sym_upd(vr, sw[c], wr, sv[c]);
      row[c] = a;
      s += a * sn[c - 1];
    }
    s = wave_sum(s);
    if (lane == 0) {
      Pn[r - 1] = tn * s;
      W[(long)(j + 1 + r) * l + j] = vr;   // v_j below its first entry (0 when t = 0)
    }
  }
}

__global__ __launch_bounds__(HH_THREADS) void hh_step_kernel(double* __restrict__ W, int l, int j,
                                                             const double* __restrict__ col, double* __restrict__ coln,
                                                             const double* __restrict__ P, double* __restrict__ Pn,
                                                             double* __restrict__ d, double* __restrict__ e,
                                                             double* __restrict__ tau, double* __restrict__ hv0) {
  __shared__ double sv[EIG_MAXL];   // v_j
  __shared__ double sw[EIG_MAXL];   // w_j
  __shared__ double sn[EIG_MAXL];   // v_{j+1}
  __shared__ double sh[HH_THREADS / 64];
  hh_step_body(W, l, j, col, coln, P, Pn, d, e, tau, hv0, blockIdx.x, gridDim.x, sv, sw, sn, sh);
}

// The last 2 x 2 block: the updated (l-2, l-2) and (l-1, l-2) from col.
__global__ void hh_tail_kernel(double* __restrict__ W, int l, const double* __restrict__ col) {
  if (threadIdx.x == 0) {
    W[(long)(l - 2) * l + (l - 2)] = col[0];
    W[(long)(l - 1) * l + (l - 2)] = col[1];
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
  if (l < EIG_MULTI_L) {
    hipLaunchKernelGGL(syevx_kernel, dim3(1), dim3(EIG_THREADS), 0, st, A, l, ld, nev, evals, evecs, ws,
                       (int*)nullptr);
    DFTB_CHECK_LAUNCH();
    return DFTB_OK;
  }
  double* W = (double*)(ws + Lo.W);
  double* d = (double*)(ws + Lo.d);
  double* e = (double*)(ws + Lo.e);
  double* tau = (double*)(ws + Lo.tau);
  // product vectors and updated columns (two of each, alternating) and the
  // reflectors' first entries: 5 l doubles of the tail's scratch (not yet in use)
  double* P2[2] = {(double*)(ws + Lo.scratch), (double*)(ws + Lo.scratch) + l};
  double* C2[2] = {P2[1] + l, P2[1] + 2 * l};
  double* hv0 = C2[1] + l;
  hipLaunchKernelGGL(sym_copy_kernel, dim3(ceil_div((long)l * l, 256) < 4096 ? ceil_div((long)l * l, 256) : 4096),
                     dim3(256), 0, st, A, l, ld, W);
  const int waves = HH_THREADS / 64;
  {
    int g = (l - 1 + waves - 1) / waves;    // one wave per row
    if (g > 256) g = 256;
    hipLaunchKernelGGL(hh_start_kernel, dim3(g), dim3(HH_THREADS), 0, st, (const double*)W, l, C2[0], P2[0]);
  }
  for (int j = 0; j + 2 < l; ++j) {
    const int rows = l - j - 2;             // rows updated (the first row of the block is not)
    int g = (rows + waves - 1) / waves;
    if (g > 256) g = 256;
    hipLaunchKernelGGL(hh_step_kernel, dim3(g), dim3(HH_THREADS), 0, st, W, l, j, (const double*)C2[j & 1],
                       C2[(j + 1) & 1], (const double*)P2[j & 1], P2[(j + 1) & 1], d, e, tau, hv0);
  }
  hipLaunchKernelGGL(hh_tail_kernel, dim3(1), dim3(64), 0, st, W, l, (const double*)C2[(l - 2) & 1]);
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
#define NS_MAXIT 100
#define RED_BLOCKS 256

// atomic maximum of non-negative doubles through their bit patterns (which order
// like the values); the target starts at +0.0
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

__device__ double blk_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double m = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, sh[i]);
  return m;
}

static int red_blocks(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > RED_BLOCKS ? RED_BLOCKS : g));
}

// Y0 = sym(A) / c, Z0 = I; errs[0 .. NS_MAXIT) = 0
__global__ void ns_init_kernel(const double* A, int l, int ld, double* Y, double* Z, const double* scale, double* errs) {
  const double c = scale[0];
  const long g0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g0 < NS_MAXIT) errs[g0] = 0.0;
  for (long i = g0; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    Y[i] = 0.5 * (A[(long)r * ld + q] + A[(long)q * ld + r]) / c;
    Z[i] = (r == q) ? 1.0 : 0.0;
  }
}

// part[b] = largest absolute row sum of the symmetric part over the workgroup's rows
// (one wave per row); the maximum over the parts bounds the largest eigenvalue
// (Gershgorin) and scales the spectrum into (0, 1].
__global__ __launch_bounds__(256) void gersh_part_kernel(const double* A, int l, int ld, double* part) {
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double m = 0.0;
  for (int i = blockIdx.x * 4 + wv; i < l; i += gridDim.x * 4) {
    double s = 0.0;
    for (int j = lane; j < l; j += 64) s += fabs(0.5 * (A[(long)i * ld + j] + A[(long)j * ld + i]));
    m = fmax(m, wave_sum(s));
  }
  m = blk_max(m, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(64) void max_parts_kernel(const double* part, int n, double* out) {
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = fmax(m, part[i]);
    out[0] = fmax(m, 1e-300);
  }
}

__global__ __launch_bounds__(64) void sum_parts_kernel(const double* part, int n, double* out) {
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s;
  }
}

// T = (3I - P)/2 ; err[0] = max |P - I| (atomic, err zeroed by ns_init)
__global__ __launch_bounds__(256) void ns_t_kernel(const double* P, int l, double* T, double* err) {
  __shared__ double sh[4];
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    int r = (int)(i / l), q = (int)(i % l);
    double id = (r == q) ? 1.0 : 0.0;
    double p = P[i];
    m = fmax(m, fabs(p - id));
    T[i] = 0.5 * (3.0 * id - p);
  }
  m = blk_max(m, sh);
  if (threadIdx.x == 0) atomic_max_nonneg(err, m);
}

// out = X * sqrt(c) (sqrt) or X / sqrt(c) (inverse sqrt)
__global__ void ns_finish_kernel(const double* X, double* out, long n, const double* scale, int inverse) {
  const double f = inverse ? 1.0 / sqrt(scale[0]) : sqrt(scale[0]);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) out[i] = X[i] * f;
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

static size_t ns_workspace(int l) { return al256((size_t)l * l * 8) * 5 + al256((size_t)(1 + NS_MAXIT) * 8); }

// Y <- A^{1/2}, Z <- A^{-1/2}; iters[0] = iterations used; returns DFTB_ERR_NOT_SPD if not converged.
static int sqrtm_ns(const double* A, int l, int ld, double* Y, double* Z, char* ws, hipStream_t st, int* iters_out,
                    double* err_out) {
  const size_t blk = al256((size_t)l * l * 8);
  double* T = (double*)ws;
  double* P = (double*)(ws + blk);
  double* Y2 = (double*)(ws + 2 * blk);
  double* Z2 = (double*)(ws + 3 * blk);
  double* scal = (double*)(ws + 5 * blk);
  double* errs = scal + 1;                     // max |Z Y - I| of every iteration
  const long n = (long)l * l;
  const int g = red_blocks(n);
  int rc;
  const int gr = (l + 3) / 4 < RED_BLOCKS ? (l + 3) / 4 : RED_BLOCKS;
  hipLaunchKernelGGL(gersh_part_kernel, dim3(gr), dim3(256), 0, st, A, l, ld, T);
  hipLaunchKernelGGL(max_parts_kernel, dim3(1), dim3(64), 0, st, (const double*)T, gr, scal);
  hipLaunchKernelGGL(ns_init_kernel, dim3((l * l + 255) / 256 > 1 ? (l * l + 255) / 256 : 1), dim3(256), 0, st, A, l,
                     ld, Y, Z, (const double*)scal, errs);
  DFTB_CHECK_LAUNCH();
  // the iterates alternate between the outputs and Y2 / Z2
  double *y = Y, *z = Z, *y2 = Y2, *z2 = Z2;
  double herr = 1.0;
  int it = 0;
  for (it = 0; it < NS_MAXIT; ++it) {
    if ((rc = gemm_launch(small_gemm(z, l, false, y, l, P, l, l, l, l), st))) return rc;
    hipLaunchKernelGGL(ns_t_kernel, dim3(g), dim3(256), 0, st, (const double*)P, l, T, errs + it);
    DFTB_CHECK_LAUNCH();
    if ((rc = gemm_launch(small_gemm(y, l, false, T, l, y2, l, l, l, l), st))) return rc;
    if ((rc = gemm_launch(small_gemm(T, l, false, z, l, z2, l, l, l, l), st))) return rc;
    double* t;
    t = y; y = y2; y2 = t;
    t = z; z = z2; z2 = t;
    if (it % 4 == 3) {
      if (hipMemcpyAsync(&herr, errs + it, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
        return DFTB_ERR_LAUNCH;
      if (hipStreamSynchronize(st) != hipSuccess) return DFTB_ERR_LAUNCH;
      if (!(herr == herr) || herr > 1e6) break;  // diverged: not SPD
      if (herr < 1e-14 * (l > 16 ? l : 16)) break;
    }
  }
  if (iters_out) *iters_out = it + 1;
  if (err_out) *err_out = herr;
  const int gf = (l * l + 255) / 256 > 1 ? (l * l + 255) / 256 : 1;
  hipLaunchKernelGGL(ns_finish_kernel, dim3(gf), dim3(256), 0, st, (const double*)y, Y, n, (const double*)scal, 0);
  hipLaunchKernelGGL(ns_finish_kernel, dim3(gf), dim3(256), 0, st, (const double*)z, Z, n, (const double*)scal, 1);
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

// err = sum |Mh - Mh^T|  (Solver.py:766-770): per-workgroup parts, added in order
__global__ __launch_bounds__(256) void herm_part_kernel(const double* M, int l, double* part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
    int r = (int)(i / l), c = (int)(i % l);
    s += fabs(M[(long)r * l + c] - M[(long)c * l + r]);
  }
  s = blk_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
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
  {
    const int g = red_blocks((long)l * l);
    double* part = (double*)(ws + S.eig);     // free between the square root and the eigensolver
    hipLaunchKernelGGL(herm_part_kernel, dim3(g), dim3(256), 0, st, (const double*)H, l, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64), 0, st, (const double*)part, g, diag);
    DFTB_CHECK_LAUNCH();
  }
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

// ---- Cholesky route of the non-Hermitian subspace step ------------------------
// With Mm = L L^T, H' = L^T Mp L is orthogonally similar to Mh = Mm^{1/2} Mp Mm^{1/2}
// (Mh = Q^T H' Q with Q = L^{-1} Mm^{1/2}): the same roots w^2, and for
// H' T' = T' diag(w^2) the eigenvectors of Mh are Tb = Q^T T', so that
// Rb = Mm^{1/2} Tb / sqrt(w) = L T' / sqrt(w) and Lb = Mp Rb / w.  The iterations
// need only the Cholesky factor (blocked: a one-workgroup factor of each 64 x 64
// diagonal block with its inverse, the panel below as a GEMM with that inverse,
// the trailing update as a GEMM); Tb itself, which only C = bs Tb after the last
// iteration uses, is Mm^{-1/2} Rb diag(sqrt(w)) (dftb_subspace_tb).
#define CHOL_NB 64

__global__ void fill_zero_kernel(double* X, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) X[i] = 0.0;
}

// Factor the b x b diagonal block at (p, p) of W (lower triangle, trailing updates
// applied) into L[p.., p..] (zeros above the diagonal).  One wave: lane i holds row
// i in registers and column k travels by lane shuffles.  A non-positive pivot sets
// info (DFTB_ERR_NOT_SPD, pivot index) and continues with a unit pivot, so that
// later launches see finite values.
__global__ __launch_bounds__(64) void chol_diag_kernel(const double* __restrict__ W, double* __restrict__ L, int l,
                                                       int p, int b, int* __restrict__ info) {
  const int i = threadIdx.x;
  double row[CHOL_NB];
#pragma unroll
  for (int j = 0; j < CHOL_NB; ++j) row[j] = (i < b && j <= i) ? W[(long)(p + i) * l + (p + j)] : 0.0;
  int bad = -1;
#pragma unroll
  for (int k = 0; k < CHOL_NB; ++k) {
    if (k < b) {
      double d = __shfl(row[k], k, 64);
      if (!(d > 0.0)) {
        if (bad < 0) bad = k;
        d = 1.0;
      }
      const double piv = sqrt(d);
      const double lik = (i == k) ? piv : (i > k ? row[k] / piv : 0.0);
      row[k] = lik;
#pragma unroll
      for (int j = k + 1; j < CHOL_NB; ++j) {
        const double ljk = __shfl(lik, j, 64);
        if (i >= j) row[j] -= lik * ljk;
      }
    }
  }
  if (i == 0 && bad >= 0 && info[0] == 0) {
    info[0] = DFTB_ERR_NOT_SPD;
    info[1] = p + bad;
  }
  if (i < b)
#pragma unroll
    for (int j = 0; j < CHOL_NB; ++j)
      if (j < b) L[(long)(p + i) * l + (p + j)] = (j <= i) ? row[j] : 0.0;
}

// Linv (b x b) = inverse of the factor block L[p.., p..]: one wave, lane c forms
// column c by column-oriented forward substitution (rows of the factor read from
// LDS at uniform addresses); entries above the diagonal come out zero.
__global__ __launch_bounds__(64) void chol_inv_kernel(const double* __restrict__ L, int l, int p, int b,
                                                      double* __restrict__ Linv) {
  __shared__ double sL[CHOL_NB][CHOL_NB + 1];
  const int i = threadIdx.x;
  for (int j = 0; j < CHOL_NB; ++j) sL[i][j] = (i < b && j < b) ? L[(long)(p + i) * l + (p + j)] : 0.0;
  __syncthreads();
  double x[CHOL_NB];
#pragma unroll
  for (int r = 0; r < CHOL_NB; ++r) x[r] = (r == i) ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < CHOL_NB; ++k) {
    if (k < b) {
      const double xk = x[k] / sL[k][k];
      x[k] = xk;
#pragma unroll
      for (int r = k + 1; r < CHOL_NB; ++r) x[r] -= sL[r][k] * xk;
    }
  }
  if (i < b)
#pragma unroll
    for (int r = 0; r < CHOL_NB; ++r)
      if (r < b) Linv[r * b + i] = x[r];
}

// L (l x l, zero above the diagonal) with L L^T = sym(Mm); W: l x l scratch.
static int cholesky_blocked(const double* Mm, int l, int ld, double* W, double* L, double* Linv, int* info,
                            hipStream_t st) {
  const int gcopy = ceil_div((long)l * l, 256) < 4096 ? ceil_div((long)l * l, 256) : 4096;
  hipLaunchKernelGGL(sym_copy_kernel, dim3(gcopy), dim3(256), 0, st, Mm, l, ld, W);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(gcopy), dim3(256), 0, st, L, (long)l * l);
  DFTB_CHECK_LAUNCH();
  int rc;
  for (int p = 0; p < l; p += CHOL_NB) {
    const int b = l - p < CHOL_NB ? l - p : CHOL_NB;
    hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(64), 0, st, (const double*)W, L, l, p, b, info);
    DFTB_CHECK_LAUNCH();
    const int r = l - p - b;
    if (r <= 0) break;
    hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(64), 0, st, (const double*)L, l, p, b, Linv);
    DFTB_CHECK_LAUNCH();
    // panel: L21 = W21 L11^{-T}
    GemmArgs g = gemm_args_default();
    g.A = W + (long)(p + b) * l + p; g.lda = l;
    g.B = Linv; g.ldb = b; g.b_trans = 1;
    g.C = L + (long)(p + b) * l + p; g.ldc = l;
    g.M = r; g.N = b; g.K = b; g.tile = 64;
    if ((rc = gemm_launch(g, st))) return rc;
    // trailing: W22 -= L21 L21^T
    GemmArgs u = gemm_args_default();
    u.A = L + (long)(p + b) * l + p; u.lda = l;
    u.B = L + (long)(p + b) * l + p; u.ldb = l; u.b_trans = 1;
    u.C = W + (long)(p + b) * l + (p + b); u.ldc = l;
    u.M = r; u.N = r; u.K = b; u.alpha = -1.0; u.beta = 1.0; u.tile = 64;
    if ((rc = gemm_launch(u, st))) return rc;
  }
  return DFTB_OK;
}

extern "C" int dftb_subspace_nonhermitian_rl(const double* Mp, const double* Mm, int l, int ld, int nev, double* w,
                                             double* Rb, double* Lb, int* info, double* diag, void* workspace,
                                             size_t ws_bytes, void* stream) {
  if (l <= 0 || nev <= 0 || nev > l || ld < l || l > EIG_MAXL) return DFTB_ERR_ARG;
  if (!Mp || !Mm || !w || !Rb || !Lb || !info || !diag || !workspace) return DFTB_ERR_ARG;
  SubLayout S = sub_layout(l, nev);
  if (ws_bytes < S.total) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* W = (double*)(ws + S.Sq);      // Cholesky scratch, then T'
  double* L = (double*)(ws + S.Sinv);
  double* X = (double*)(ws + S.X);       // the diagonal blocks' inverses, then Mp L
  double* H = (double*)(ws + S.H);
  double* w2 = (double*)(ws + S.w2);
  double* sR = (double*)(ws + S.sR);
  double* sL = (double*)(ws + S.sL);
  double* G = (double*)(ws + S.G);
  int rc;
  hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, st, info);
  DFTB_CHECK_LAUNCH();
  if ((rc = cholesky_blocked(Mm, l, ld, W, L, X, info, st))) return rc;
  // H' = L^T Mp L, Hermiticity check
  if ((rc = gemm_launch(small_gemm(Mp, ld, false, L, l, X, l, l, l, l), st))) return rc;
  if ((rc = gemm_launch(small_gemm(L, l, true, X, l, H, l, l, l, l), st))) return rc;
  {
    const int g = red_blocks((long)l * l);
    double* part = (double*)(ws + S.eig);
    hipLaunchKernelGGL(herm_part_kernel, dim3(g), dim3(256), 0, st, (const double*)H, l, part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64), 0, st, (const double*)part, g, diag);
    DFTB_CHECK_LAUNCH();
  }
  double* Tp = W;
  if ((rc = dftb_syevx(H, l, l, nev, w2, Tp, ws + S.eig, eig_layout(l, nev).total, stream))) return rc;
  hipLaunchKernelGGL(roots_kernel, dim3((nev + 63) / 64), dim3(64), 0, st, w2, nev, w, sR, sL, info);
  DFTB_CHECK_LAUNCH();
  // Rb = L T' / sqrt(w);  Lb = Mp Rb / w
  if ((rc = gemm_launch(small_gemm(L, l, false, Tp, nev, Rb, nev, l, nev, l), st))) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((l * nev + 255) / 256), dim3(256), 0, st, Rb, l, nev, sR);
  DFTB_CHECK_LAUNCH();
  if ((rc = gemm_launch(small_gemm(Mp, ld, false, Rb, nev, Lb, nev, l, nev, l), st))) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((l * nev + 255) / 256), dim3(256), 0, st, Lb, l, nev, sL);
  DFTB_CHECK_LAUNCH();
  if ((rc = gemm_launch(small_gemm(Lb, nev, true, Rb, nev, G, nev, nev, nev, l), st))) return rc;
  hipLaunchKernelGGL(biortho_kernel, dim3(1), dim3(256), 0, st, G, nev, diag);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

__global__ void sqrt_vec_kernel(const double* w, int n, double* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = sqrt(fmax(w[i], 0.0));
}

extern "C" int dftb_subspace_tb(const double* Mm, int l, int ld, int nev, const double* w, const double* Rb,
                                double* Tb, void* workspace, size_t ws_bytes, void* stream) {
  if (l <= 0 || nev <= 0 || nev > l || ld < l || l > EIG_MAXL || !Mm || !w || !Rb || !Tb || !workspace)
    return DFTB_ERR_ARG;
  SubLayout S = sub_layout(l, nev);
  if (ws_bytes < S.total) return DFTB_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* Sq = (double*)(ws + S.Sq);
  double* Sinv = (double*)(ws + S.Sinv);
  double* sw = (double*)(ws + S.sR);
  int nsit = 0;
  double nserr = 0.0;
  int rc = sqrtm_ns(Mm, l, ld, Sq, Sinv, ws + S.eig, st, &nsit, &nserr);
  if (rc) return rc;
  // Tb = Mm^{-1/2} Rb diag(sqrt(w))
  if ((rc = gemm_launch(small_gemm(Sinv, l, false, Rb, nev, Tb, nev, l, nev, l), st))) return rc;
  hipLaunchKernelGGL(sqrt_vec_kernel, dim3((nev + 63) / 64), dim3(64), 0, st, w, nev, sw);
  hipLaunchKernelGGL(scale_cols_kernel, dim3((l * nev + 255) / 256), dim3(256), 0, st, Tb, l, nev, (const double*)sw);
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

// Replace the leading l x l block of M by its symmetric part; asym[0] = max |M - M^T| / max |M|
// (asym[1], asym[2]: the two maxima, accumulated atomically from zero).
__global__ void symmetrize_zero_kernel(double* asym) {
  if (threadIdx.x < 3) asym[threadIdx.x] = 0.0;
}

__global__ __launch_bounds__(256) void symmetrize_kernel(double* M, int l, int ld, double* asym) {
  __shared__ double sh[4];
  double a = 0.0, mx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)l * l; i += (long)gridDim.x * blockDim.x) {
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
  a = blk_max(a, sh);
  mx = blk_max(mx, sh);
  if (threadIdx.x == 0) {
    atomic_max_nonneg(asym + 1, a);
    atomic_max_nonneg(asym + 2, mx);
  }
}

__global__ void symmetrize_ratio_kernel(double* asym) {
  if (threadIdx.x == 0) asym[0] = asym[2] > 0.0 ? asym[1] / asym[2] : 0.0;
}

extern "C" int dftb_symmetrize(double* M, int l, int ld, double* asym, void* stream) {
  if (l <= 0 || ld < l || !M || !asym) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(symmetrize_zero_kernel, dim3(1), dim3(64), 0, st, asym);
  hipLaunchKernelGGL(symmetrize_kernel, dim3(red_blocks((long)l * l)), dim3(256), 0, st, M, l, ld, asym);
  hipLaunchKernelGGL(symmetrize_ratio_kernel, dim3(1), dim3(64), 0, st, asym);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
