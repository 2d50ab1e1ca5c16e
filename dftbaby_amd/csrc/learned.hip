// Learned Davidson guess / preconditioner: feature, restricted-subspace and
// MLP training/inference kernels (FP32 network, FP64 physics).
//
// Per-pair features for pair ia (all cheap, O(N) after two GEMMs):
//   f0 = omega_ia, f1 = ||q^{ia}||_2 over atoms, f2 = A_{ia,ia}, f3 = B_{ia,ia},
//   with A_{ia,ia} = omega_ia + 2 (ia|ia) - (ii|aa)_lr, B_{ia,ia} = 2 (ia|ia) - (ia|ia)_lr
//   (the diagonals of Solver.buildAB, Solver.py:35-96).
// The MLP (width 128, 3 hidden layers, SiLU) runs its matmuls on the MFMA GEMM
// (FP32 operands); the kernels here are its bias/activation/loss/Adam steps.
#include "common.h"
#include "gemm.h"

static dim3 lgrid(long n) {
  long b = (n + 255) / 256;
  if (b > 32768) b = 32768;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// cd[x] = sum_A P[A][x] * Q[A][x]   (column-wise dot over atoms; P: [nat][ldp], Q: [nat][ldq])
__global__ void coldot_kernel(const double* __restrict__ P, long ldp, const double* __restrict__ Q, long ldq,
                              int nat, long N, double* __restrict__ out) {
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int A = 0; A < nat; ++A) s += P[(long)A * ldp + x] * Q[(long)A * ldq + x];
    out[x] = s;
  }
}

extern "C" int dftb_coldot(const double* P, long ldp, const double* Q, long ldq, int nat, long N, double* out,
                           void* stream) {
  if (!P || !Q || !out || nat <= 0 || N <= 0 || ldp < N || ldq < N) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(coldot_kernel, lgrid(N), dim3(256), 0, (hipStream_t)stream, P, ldp, Q, ldq, nat, N, out);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// Assemble features [N][nf] (float) from FP64 pieces on the padded grid.
//  omega, qn2 = sum_A q^2, cc = (ia|ia), cl = (ia|ia)_lr, xl = (ii|aa)_lr  (all [opad*vpad])
__global__ void features_kernel(const double* omega, const double* qn2, const double* cc, const double* cl,
                                const double* xl, long N, int nf, float* F) {
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < N; x += (long)gridDim.x * blockDim.x) {
    double w = omega[x];
    double A = w + 2.0 * cc[x] - xl[x];
    double B = 2.0 * cc[x] - cl[x];
    float* f = F + x * nf;
    f[0] = (float)w;
    f[1] = (float)sqrt(qn2[x]);
    f[2] = (float)A;
    f[3] = (float)B;
    if (nf > 4) f[4] = (float)(A - B);
    if (nf > 5) f[5] = (float)sqrt(fmax((A + B) * (A - B), 0.0));
  }
}

extern "C" int dftb_pair_features(const double* omega, const double* qn2, const double* cc, const double* cl,
                                  const double* xl, long N, int nf, float* F, void* stream) {
  if (nf < 4 || nf > 6 || N <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(features_kernel, lgrid(N), dim3(256), 0, (hipStream_t)stream, omega, qn2, cc, cl, xl, N, nf, F);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// Restricted (A+B), (A-B) on M selected pairs p = (ip, ap) (engine layouts):
//   (A+-B)_pq = omega_p d_pq + 4 (p|q) - (i_p j_q|a_p b_q)_lr -+ (i_p b_q|j_q a_p)_lr
// with (p|q) = Kc[p][q] precomputed (Coulomb, via GEMMs), and the exchange
// integrals from q_oo[i][A][j] * qf_vv[b][A][a] and q_ov[i][A][b] * qf_ov[j][A][a].
__global__ void restricted_ab_kernel(const int* __restrict__ ip, const int* __restrict__ ap, int M, int nat, int opad,
                                     int vpad, const double* __restrict__ omega, const double* __restrict__ Kc,
                                     const double* __restrict__ q_oo, const double* __restrict__ q_ov,
                                     const double* __restrict__ qf_ov, const double* __restrict__ qf_vv, int lc,
                                     double* __restrict__ ApB, double* __restrict__ AmB) {
  const long total = (long)M * M;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int p = (int)(t / M), q = (int)(t % M);
    const int i = ip[p], a = ap[p], j = ip[q], b = ap[q];
    double kA = 0.0, kB = 0.0;
    if (lc) {
      const double* qij = q_oo + ((long)i * nat) * opad + j;   // q_oo[i][A][j], stride opad over A
      const double* qab = qf_vv + ((long)b * nat) * vpad + a;  // qf_vv[b][A][a]
      const double* qib = q_ov + ((long)i * nat) * vpad + b;   // q_ov[i][A][b]
      const double* qja = qf_ov + ((long)j * nat) * vpad + a;  // qf_ov[j][A][a]
      for (int A = 0; A < nat; ++A) {
        kA += qij[(long)A * opad] * qab[(long)A * vpad];
        kB += qib[(long)A * vpad] * qja[(long)A * vpad];
      }
    }
    const double d = (p == q) ? omega[(long)i * vpad + a] : 0.0;
    const double c = 4.0 * Kc[t];
    ApB[t] = d + c - kA - kB;
    AmB[t] = d - kA + kB;
  }
}

extern "C" int dftb_restricted_ab(const int* ip, const int* ap, int M, int nat, int opad, int vpad, const double* omega,
                                  const double* Kc, const double* q_oo, const double* q_ov, const double* qf_ov,
                                  const double* qf_vv, int lc, double* ApB, double* AmB, void* stream) {
  if (M <= 0 || !ip || !ap || !omega || !Kc || !ApB || !AmB) return DFTB_ERR_ARG;
  if (lc && (!q_oo || !q_ov || !qf_ov || !qf_vv)) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(restricted_ab_kernel, lgrid((long)M * M), dim3(256), 0, (hipStream_t)stream, ip, ap, M, nat,
                     opad, vpad, omega, Kc, q_oo, q_ov, qf_ov, qf_vv, lc, ApB, AmB);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

// ---------------------------------------------------------------------------
// MLP pieces (row-major activations [rows][width], FP32)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Z[r][c] += b[c]; H = act ? silu(Z) : Z     (Z kept for the backward pass)
__global__ void bias_act_kernel(float* Z, const float* b, long rows, int cols, int act, float* H) {
  const long total = rows * cols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    float z = Z[i] + b[i % cols];
    Z[i] = z;
    H[i] = act ? silu(z) : z;
  }
}

// G[r][c] *= dsilu(Z[r][c])
__global__ void act_bwd_kernel(float* G, const float* Z, long total) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    G[i] *= dsilu(Z[i]);
}

// db[c] = sum_r G[r][c]   (one thread per column, deterministic)
__global__ void colsum_kernel(const float* G, long rows, int cols, float* db) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (long r = 0; r < rows; ++r) s += G[r * cols + c];
  db[c] = (float)s;
}

// weighted MSE over nout outputs: loss = sum_r sum_o w_o (Y - T)^2 / rows;  G = dloss/dY
__global__ void mse_kernel(const float* Y, const float* T, const float* w, long rows, int nout, float* G,
                           float* partial) {
  __shared__ float sh[256];
  float acc = 0.0f;
  const long total = rows * nout;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int o = (int)(i % nout);
    float d = Y[i] - T[i];
    acc += w[o] * d * d;
    G[i] = 2.0f * w[o] * d / (float)rows;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] / (float)rows;
}

// Adam step on a flat parameter vector
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float wd) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float gi = g[i] + wd * p[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// (x - mean) / std per feature column (FP32 in place)
__global__ void normalize_kernel(float* X, long rows, int cols, const float* mean, const float* inv_std) {
  const long total = rows * cols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int c = (int)(i % cols);
    X[i] = (X[i] - mean[c]) * inv_std[c];
  }
}

extern "C" int dftb_mlp_bias_act(float* Z, const float* b, long rows, int cols, int act, float* H, void* stream) {
  if (!Z || !b || !H || rows <= 0 || cols <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(bias_act_kernel, lgrid(rows * cols), dim3(256), 0, (hipStream_t)stream, Z, b, rows, cols, act, H);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_mlp_act_bwd(float* G, const float* Z, long n, void* stream) {
  if (!G || !Z || n <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(act_bwd_kernel, lgrid(n), dim3(256), 0, (hipStream_t)stream, G, Z, n);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_mlp_colsum(const float* G, long rows, int cols, float* db, void* stream) {
  if (!G || !db || rows <= 0 || cols <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(colsum_kernel, dim3((cols + 127) / 128), dim3(128), 0, (hipStream_t)stream, G, rows, cols, db);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_mlp_mse(const float* Y, const float* T, const float* w, long rows, int nout, float* G,
                            float* partial, int nblocks, void* stream) {
  if (!Y || !T || !w || !G || !partial || rows <= 0 || nblocks <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, Y, T, w, rows, nout, G, partial);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                         int step, float wd, void* stream) {
  if (!p || !g || !m || !v || n <= 0 || step <= 0) return DFTB_ERR_ARG;
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, lgrid(n), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2,
                     wd);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_mlp_normalize(float* X, long rows, int cols, const float* mean, const float* inv_std,
                                  void* stream) {
  if (!X || !mean || !inv_std || rows <= 0 || cols <= 0) return DFTB_ERR_ARG;
  hipLaunchKernelGGL(normalize_kernel, lgrid(rows * cols), dim3(256), 0, (hipStream_t)stream, X, rows, cols, mean,
                     inv_std);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
