// (b) Casida matvecs (A+B)V and (A-B)V for a block of vectors.
//
// Reference: DFTB/Solver.py:583-635 (_ApBv/_AmBv) and the Fortran twins
// DFTB/extensions/tddftb.f90:114-287.  The reference evaluates, per vector,
// the Coulomb term and two long-range exchange terms with np.tensordot, and
// recomputes the gamma_lr application inside every product.  Here gamma_lr is
// folded into q_vv and q_ov once per geometry, the block of vectors is
// processed together, and each exchange term is two MFMA GEMMs over an atom
// slice:
//   K1:  T[n][A][j][a] = sum_b V[n][j][b] qf_vv[b][A][a]        (expand)
//        X1[n][i][a]  += sum_{A,j} q_oo[i][A][j] T[n][A][j][a]   (contract)
//   K2:  U[n][A][b][a] = sum_j V[n][j][b] qf_ov[j][A][a]        (expand)
//        X2[n][i][a]  += sum_{A,b} q_ov[i][A][b] U[n][A][b][a]   (contract)
// and the results combine as (A+B)V = omega V + 4C - X1 - X2,
// (A-B)V = omega V - X1 + X2, where C = q^T gamma q V is the Coulomb term.
#include "common.h"
#include "gemm.h"

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
typedef unsigned short u16_t;

struct MvLayout {
  long N;        // vector length opad*vpad
  int ks;        // split-K slices for the Coulomb projection
  int ac;        // atoms per exchange slice
  int aoc;       // vectors per AO-exchange chunk
  size_t vin, tpart, t, g, g32, x1, x2, tu, ao, total;
};

// AO exchange intermediates per vector (elements): B1 [2 npad][max(opad, vpad)],
// BH [2 npad][2 npad], B3 [vpad][2 npad]; BF16x3 keeps two BH (BH and its
// transpose) and two B3 per vector
static void ao_sizes(const CasidaOperator* op, long* e1, long* eh, long* e3) {
  const long np2 = 2L * op->npad, w = op->opad > op->vpad ? op->opad : op->vpad;
  *e1 = np2 * w;
  *eh = np2 * np2;
  *e3 = (long)op->vpad * np2;
}
// bytes per element summed over the arrays of one operand (BF16x3: hi + lo)
static int ao_esz(int precision) { return precision == 0 ? 8 : 4; }

static int in_bytes(int precision) { return precision == 0 ? 8 : 4; }  // f64, f32, bf16 hi+lo

static MvLayout mv_layout(const CasidaOperator* op, int nvec, bool hermitian) {
  MvLayout L;
  L.N = (long)op->opad * op->vpad;
  L.ks = (int)((L.N + 32767) / 32768);
  if (L.ks > 64) L.ks = 64;
  if (L.ks < 1) L.ks = 1;
  const int s = in_bytes(op->precision);
  long rows = op->opad > op->vpad ? op->opad : op->vpad;
  size_t per_atom = (size_t)nvec * rows * op->vpad * s;
  int ac = op->atom_chunk;
  if (ac <= 0) {
    const size_t budget = (size_t)16 << 30;
    ac = (int)(budget / (per_atom ? per_atom : 1));
    if (ac < 1) ac = 1;
  }
  if (ac > op->nat) ac = op->nat;
  L.ac = ac;
  size_t off = 0;
  // hermitian: sqrt(omega)-scaled FP64 copy; mixed: FP32 copy (+ bf16 hi/lo pair for BF16x3)
  size_t vin_bytes = hermitian ? (size_t)nvec * L.N * 8 : 0;
  if (op->precision > 0) vin_bytes += (size_t)nvec * L.N * 4;
  if (op->precision == 2 && !hermitian) vin_bytes += (size_t)nvec * L.N * 8;  // hi/lo and transposed hi/lo
  L.vin = off; off += al256(vin_bytes);
  L.tpart = off; off += al256((size_t)L.ks * nvec * op->nat * 8);
  L.t = off; off += al256((size_t)nvec * op->nat * 8);
  L.g = off; off += al256((size_t)nvec * op->nat * 8);
  L.g32 = off; off += al256((size_t)nvec * op->nat * 4);
  bool exch = !hermitian && op->lc;
  const bool ao = exch && op->exchange == 1;
  L.x1 = off; off += exch ? al256((size_t)nvec * L.N * 8) : 0;
  L.x2 = off; off += exch ? al256((size_t)nvec * L.N * 8) : 0;
  L.tu = off; off += (exch && !ao) ? al256(per_atom * ac) : 0;
  L.aoc = 0;
  L.ao = off;
  if (ao) {
    long e1, eh, e3;
    ao_sizes(op, &e1, &eh, &e3);
    const int dup = op->precision == 2 ? 2 : 1;
    const size_t per_vec = (size_t)(e1 + dup * (eh + e3)) * ao_esz(op->precision);
    const size_t budget = (size_t)16 << 30;
    long c = (long)(budget / (per_vec ? per_vec : 1));
    if (c < 1) c = 1;
    if (c > nvec) c = nvec;
    L.aoc = (int)c;
    off += al256(per_vec * c + 3 * 256);
  }
  L.total = off;
  return L;
}

extern "C" size_t dftb_casida_matvec_workspace(const CasidaOperator* op, int nvec) {
  if (!op || nvec <= 0) return 0;
  return mv_layout(op, nvec, false).total;
}
extern "C" size_t dftb_hermitian_matvec_workspace(const CasidaOperator* op, int nvec) {
  if (!op || nvec <= 0) return 0;
  return mv_layout(op, nvec, true).total;
}

__global__ void combine_kernel(const double* __restrict__ omega, const double* __restrict__ V, long N, long total,
                               double* __restrict__ P, double* __restrict__ M, const double* __restrict__ X1,
                               const double* __restrict__ X2, int lc) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    double ov = omega[idx % N] * V[idx];
    double c = P[idx];
    if (lc) {
      double k1 = X1[idx], k2 = X2[idx];
      P[idx] = ov + c - k1 - k2;
      M[idx] = ov - k1 + k2;
    } else {
      P[idx] = ov + c;
      M[idx] = ov;
    }
  }
}

__global__ void zero_kernel(double* x, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = 0.0;
}

__global__ void scale_vec_kernel(const double* __restrict__ s, const double* __restrict__ V, long N, long total,
                                 double* __restrict__ out) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x)
    out[idx] = s[idx % N] * V[idx];
}

__global__ void herm_finish_kernel(const double* __restrict__ om, const double* __restrict__ sq,
                                   const double* __restrict__ V, long N, long total, double* __restrict__ R) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long e = idx % N;
    R[idx] = om[e] * V[idx] + sq[e] * R[idx];
  }
}

__global__ void cvt32_kernel(const double* x, long n, float* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void cvtpair_kernel(const double* x, long n, unsigned short* hi, unsigned short* lo) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float f = (float)x[i];
    unsigned short h = bf16_bits_rne(f);
    hi[i] = h;
    lo[i] = bf16_bits_rne(f - bf16_bits_to_float(h));
  }
}

// out[b][c][r] = split(x[b][r][c]) for r < R, zeros for R <= r < Rpad, c < Cc (tiled through LDS)
__global__ void transpose_pair_kernel(const double* __restrict__ x, long R, long Rpad, long Cc, long sIn, long ldIn,
                                      unsigned short* __restrict__ hi, unsigned short* __restrict__ lo, long sOut,
                                      long ldOut) {
  __shared__ float tile[32][33];
  const long bt = blockIdx.z;
  const long r0 = (long)blockIdx.y * 32, c0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int k = ty; k < 32; k += 8) {
    long r = r0 + k, c = c0 + tx;
    float v = 0.0f;
    if (r < R && c < Cc) v = (float)x[bt * sIn + r * ldIn + c];
    tile[k][tx] = v;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    long c = c0 + k, r = r0 + tx;
    if (c < Cc && r < Rpad) {
      float f = tile[tx][k];
      unsigned short h = bf16_bits_rne(f);
      long o = bt * sOut + c * ldOut + r;
      hi[o] = h;
      lo[o] = bf16_bits_rne(f - bf16_bits_to_float(h));
    }
  }
}

static int transpose_pair(const double* x, long batch, long R, long Rpad, long Cc, long sIn, long ldIn,
                          unsigned short* hi, unsigned short* lo, long sOut, long ldOut, hipStream_t st) {
  if (batch <= 0 || Rpad <= 0 || Cc <= 0) return DFTB_OK;
  dim3 grid(ceil_div(Cc, 32), ceil_div(Rpad, 32), (unsigned)batch);
  hipLaunchKernelGGL(transpose_pair_kernel, grid, dim3(256), 0, st, x, R, Rpad, Cc, sIn, ldIn, hi, lo, sOut, ldOut);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_convert_bf16pair_atom_major(const double* q, long np, long nppad, int nat, long qpad,
                                                unsigned short* hi, unsigned short* lo, void* stream) {
  // q[p][A][x] (p < np) -> out[A][x][p] with p padded to nppad
  if (!q || !hi || !lo || nat <= 0) return DFTB_ERR_ARG;
  return transpose_pair(q, nat, np, nppad, qpad, qpad, (long)nat * qpad, hi, lo, qpad * nppad, nppad,
                        (hipStream_t)stream);
}

static dim3 sgrid(long n) {
  long b = (n + 255) / 256;
  if (b > 32768) b = 32768;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// Coulomb chain: out[n][x] = alpha * sum_A (sum_B gamma[A][B] sum_y q_c[B][y] Vx[n][y]) q_c[A][x]
static int coulomb(const CasidaOperator* op, const MvLayout& L, char* ws, const void* Vx, int nvec, double alpha,
                   double* out, hipStream_t st) {
  const bool f32 = op->precision > 0;
  int rc;
  double* tpart = (double*)(ws + L.tpart);
  double* t = (double*)(ws + L.t);
  double* g = (double*)(ws + L.g);
  float* g32 = (float*)(ws + L.g32);
  long kchunk = ((L.N + L.ks - 1) / L.ks + 31) / 32 * 32;
  int ks = (int)((L.N + kchunk - 1) / kchunk);
  GemmArgs a = gemm_args_default();
  a.in_type = f32 ? GEMM_IN_F32 : GEMM_IN_F64;
  a.A = Vx; a.lda = L.N;
  a.B = op->q_c; a.ldb = L.N; a.b_trans = 1;
  a.C = tpart; a.ldc = op->nat; a.sCk = (long)nvec * op->nat;
  a.M = nvec; a.N = op->nat; a.K = (int)L.N;
  a.ksplit = ks; a.kchunk = (int)kchunk; a.tile = 64;
  if ((rc = gemm_launch(a, st))) return rc;
  if ((rc = gemm_reduce_launch(tpart, (long)nvec * op->nat, ks, t, (long)nvec * op->nat, 1.0, 0.0, st))) return rc;
  GemmArgs b = gemm_args_default();
  b.A = t; b.lda = op->nat;
  b.B = op->gamma; b.ldb = op->nat;
  b.C = g; b.ldc = op->nat;
  b.M = nvec; b.N = op->nat; b.K = op->nat; b.tile = 64;
  if ((rc = gemm_launch(b, st))) return rc;
  const void* gx = g;
  if (f32) {
    hipLaunchKernelGGL(cvt32_kernel, sgrid((long)nvec * op->nat), dim3(256), 0, st, g, (long)nvec * op->nat, g32);
    DFTB_CHECK_LAUNCH();
    gx = g32;
  }
  GemmArgs c = gemm_args_default();
  c.in_type = f32 ? GEMM_IN_F32 : GEMM_IN_F64;
  c.A = gx; c.lda = op->nat;
  c.B = op->q_c; c.ldb = L.N;
  c.C = out; c.ldc = L.N;
  c.M = nvec; c.N = (int)L.N; c.K = op->nat; c.alpha = alpha; c.tile = 64;
  return gemm_launch(c, st);
}

// One GEMM of the AO exchange in the operand type of the precision: C = alpha A B^T
// (btrans = 1) or alpha A B (btrans = 0, FP64/FP32 only); final = FP64 output.
static int ao_gemm(int prec, const void* A, const void* A2, long lda, const void* B, const void* B2, long ldb,
                   long sB, int btrans, void* C, void* C2, long ldc, long sC, int M, int N, int K, int batch,
                   double alpha, bool final, const CasidaOperator* op, bool had, hipStream_t st, int tile = 0) {
  GemmArgs g = gemm_args_default();
  // a tuned pipelined tile (256: 256 x 128, 512: 256 x 256) applies where the
  // automatic choice is the pipelined kernel too
  g.tile = (prec == 2 && (tile == 256 || tile == 512) && M >= 256 && N >= 128 && K >= 128) ? tile : 0;
  g.in_type = prec == 0 ? GEMM_IN_F64 : (prec == 1 ? GEMM_IN_F32 : GEMM_IN_BF16X3);
  g.out_type = final ? GEMM_OUT_F64 : (prec == 0 ? GEMM_OUT_F64 : (prec == 1 ? GEMM_OUT_F32 : GEMM_OUT_BF16PAIR));
  g.A = A; g.A2 = A2; g.lda = lda; g.sA = 0;
  g.B = B; g.B2 = B2; g.ldb = ldb; g.sB = sB; g.b_trans = btrans;
  g.C = C; g.C2 = C2; g.ldc = ldc; g.sC = sC;
  g.M = M; g.N = N; g.K = K; g.batch = batch; g.alpha = alpha;
  if (had) { g.had = op->gam_ao; g.hmod = op->npad; g.ldh = op->nat + 1; g.hmap = op->amap; }
  return gemm_launch(g, st);
}

// Long-range exchange in the AO basis.  With the Mulliken transition charges
// q^{pq}_A = 1/2 sum_{mu in A} (C_mu,p Y_mu,q + Y_mu,p C_mu,q), Y = S C, the two
// exchange products of a vector V (o x v) are
//   K1 V = 1/4 [C_o; Y_o]^T { G o ([Y_o; C_o] V [Y_v; C_v]^T) } [C_v; Y_v]
//   K2 V = 1/4 [C_o; Y_o]^T { G o ([Y_v; C_v] V^T [Y_o; C_o]^T) } [C_v; Y_v]
// (G: gamma_lr expanded to orbital pairs, o: Hadamard product), i.e. four GEMMs
// each over 2 npad-sized AO dimensions instead of the Nat x o x v x (o + v)
// transition-charge contractions.  Every GEMM is in the K-contiguous (NT) form.
// G is symmetric, so the bracket of K2 is the transpose of that of K1 and K2
// starts from BH (six GEMMs per vector instead of eight): in BF16x3 mode (NT
// GEMMs only) the Hadamard GEMM also writes BH^T; FP64 / FP32 read BH in the NN
// form.
static int ao_exchange(const CasidaOperator* op, const MvLayout& L, char* ws, int nvec, const void* V,
                       const void* Vlo, double* X1, double* X2, hipStream_t st) {
  const int prec = op->precision;
  const long np2 = 2L * op->npad, opad = op->opad, vpad = op->vpad, N = L.N;
  long e1, eh, e3;
  ao_sizes(op, &e1, &eh, &e3);
  const long vs = prec == 0 ? 8 : (prec == 1 ? 4 : 2);   // bytes per vector element (per array)
  const long es = vs;                                     // intermediates use the same element type
  const int dup = prec == 2 ? 2 : 1;                      // BH, B3 per vector
  int rc;
  for (int n0 = 0; n0 < nvec; n0 += L.aoc) {
    const int c = (nvec - n0 < L.aoc) ? (nvec - n0) : L.aoc;
    char* b1 = (char*)(((uintptr_t)(ws + L.ao) + 255) & ~(uintptr_t)255);
    char* bh = b1 + (size_t)c * e1 * es;
    char* b3 = bh + (size_t)dup * c * eh * es;
    char *b1l = nullptr, *bhl = nullptr, *b3l = nullptr;
    if (prec == 2) {
      b1l = b3 + (size_t)dup * c * e3 * es;
      bhl = b1l + (size_t)c * e1 * es;
      b3l = bhl + (size_t)dup * c * eh * es;
    }
    const char* v0 = (const char*)V + (size_t)n0 * N * vs;
    const char* v0l = Vlo ? (const char*)Vlo + (size_t)n0 * N * vs : nullptr;
    double* x1 = X1 + (size_t)n0 * N;
    double* x2 = X2 + (size_t)n0 * N;
    // ---- K1: B1 = [Y_v; C_v] V^T, BH = G o ([Y_o; C_o] B1^T), B3 = [C_v; Y_v]^T BH^T, X1 = 1/4 W B3^T
    if ((rc = ao_gemm(prec, op->ao_v[0], op->ao_v[1], vpad, v0, v0l, vpad, N, 1, b1, b1l, opad, e1, (int)np2,
                      (int)opad, (int)vpad, c, 1.0, false, op, false, st, op->tile_expand))) return rc;
    if (prec == 2) {
      // BH and BH^T (= the K2 bracket) for the c vectors, then B3 of both in one launch
      GemmArgs g = gemm_args_default();
      g.in_type = GEMM_IN_BF16X3;
      g.out_type = GEMM_OUT_BF16PAIR;
      g.A = op->ao_o[0]; g.A2 = op->ao_o[1]; g.lda = opad; g.sA = 0;
      g.B = b1; g.B2 = b1l; g.ldb = opad; g.sB = e1; g.b_trans = 1;
      g.C = bh; g.C2 = bhl; g.ldc = np2; g.sC = eh;
      g.Ct = bh + (size_t)c * eh * es; g.Ct2 = bhl + (size_t)c * eh * es; g.ldct = np2; g.sCt = eh;
      g.M = (int)np2; g.N = (int)np2; g.K = (int)opad; g.batch = c; g.alpha = 1.0;
      g.had = op->gam_ao; g.hmod = op->npad; g.ldh = op->nat + 1; g.hmap = op->amap;
      g.tile = (op->tile_expand == 256 || op->tile_expand == 512) ? op->tile_expand : 0;
      if ((rc = gemm_launch(g, st))) return rc;
      if ((rc = ao_gemm(prec, op->ao_cyvt[0], op->ao_cyvt[1], np2, bh, bhl, np2, eh, 1, b3, b3l, np2, e3,
                        (int)vpad, (int)np2, (int)np2, 2 * c, 1.0, false, op, false, st, op->tile_contract))) return rc;
      if ((rc = ao_gemm(prec, op->ao_w[0], op->ao_w[1], np2, b3, b3l, np2, e3, 1, x1, nullptr, vpad, N, (int)opad,
                        (int)vpad, (int)np2, c, 0.25, true, op, false, st, op->tile_contract))) return rc;
      const size_t o3 = (size_t)c * e3 * es;
      if ((rc = ao_gemm(prec, op->ao_w[0], op->ao_w[1], np2, b3 + o3, b3l + o3, np2, e3, 1, x2, nullptr, vpad, N,
                        (int)opad, (int)vpad, (int)np2, c, 0.25, true, op, false, st, op->tile_contract))) return rc;
      continue;
    }
    if ((rc = ao_gemm(prec, op->ao_o[0], op->ao_o[1], opad, b1, b1l, opad, e1, 1, bh, bhl, np2, eh, (int)np2,
                      (int)np2, (int)opad, c, 1.0, false, op, true, st))) return rc;
    if ((rc = ao_gemm(prec, op->ao_cyvt[0], op->ao_cyvt[1], np2, bh, bhl, np2, eh, 1, b3, b3l, np2, e3,
                      (int)vpad, (int)np2, (int)np2, c, 1.0, false, op, false, st))) return rc;
    if ((rc = ao_gemm(prec, op->ao_w[0], op->ao_w[1], np2, b3, b3l, np2, e3, 1, x1, nullptr, vpad, N, (int)opad,
                      (int)vpad, (int)np2, c, 0.25, true, op, false, st))) return rc;
    // ---- K2 (FP64 / FP32): B3 = [C_v; Y_v]^T BH (NN form: the K2 bracket is BH^T), X2 as above
    if ((rc = ao_gemm(prec, op->ao_cyvt[0], op->ao_cyvt[1], np2, bh, bhl, np2, eh, 0, b3, b3l, np2, e3,
                      (int)vpad, (int)np2, (int)np2, c, 1.0, false, op, false, st))) return rc;
    if ((rc = ao_gemm(prec, op->ao_w[0], op->ao_w[1], np2, b3, b3l, np2, e3, 1, x2, nullptr, vpad, N, (int)opad,
                      (int)vpad, (int)np2, c, 0.25, true, op, false, st))) return rc;
  }
  return DFTB_OK;
}

static int check_op(const CasidaOperator* op) {
  if (!op) return DFTB_ERR_ARG;
  if (op->nat <= 0 || op->nocc <= 0 || op->nvirt <= 0) return DFTB_ERR_ARG;
  if (op->opad < op->nocc || op->vpad < op->nvirt || (op->opad & 7) || (op->vpad & 7)) return DFTB_ERR_ARG;
  if (op->precision < 0 || op->precision > 2) return DFTB_ERR_ARG;
  if (!op->omega || !op->gamma) return DFTB_ERR_ARG;
  if (op->lc && op->exchange == 1) {
    if (op->npad <= 0 || (op->npad & 7) || !op->gam_ao || !op->amap) return DFTB_ERR_ARG;
    if (!op->ao_v[0] || !op->ao_o[0] || !op->ao_cyvt[0] || !op->ao_w[0]) return DFTB_ERR_ARG;
    if (op->precision == 2 && (!op->ao_v[1] || !op->ao_o[1] || !op->ao_cyvt[1] || !op->ao_w[1])) return DFTB_ERR_ARG;
  } else if (op->lc) {
    if (op->exchange != 0) return DFTB_ERR_ARG;
    if (!op->q_oo[0] || !op->q_ov[0] || !op->qf_ov[0] || !op->qf_vv[0]) return DFTB_ERR_ARG;
    if (op->precision == 2 && (!op->q_oo[1] || !op->q_ov[1] || !op->qf_ov[1] || !op->qf_vv[1])) return DFTB_ERR_ARG;
  }
  return DFTB_OK;
}

extern "C" int dftb_casida_matvec(const CasidaOperator* op, const double* V, int nvec, double* ApBV, double* AmBV,
                                  void* workspace, size_t ws_bytes, void* stream) {
  int rc = check_op(op);
  if (rc) return rc;
  if (nvec <= 0) return DFTB_OK;
  if (!V || !ApBV || !AmBV || !workspace) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  MvLayout L = mv_layout(op, nvec, false);
  if (ws_bytes < L.total) return DFTB_ERR_WORKSPACE;
  char* ws = (char*)workspace;
  const long total = (long)nvec * L.N;
  const int prec = op->precision;

  // operand copies of the vectors: FP32 (Coulomb in both mixed modes, exchange in
  // FP32 mode) and bf16 hi/lo pairs (exchange in BF16x3 mode)
  const void* Vx = V;
  const void* Vlo = nullptr;
  const void* Vc = V;
  if (prec > 0) {
    float* v32 = (float*)(ws + L.vin);
    hipLaunchKernelGGL(cvt32_kernel, sgrid(total), dim3(256), 0, st, V, total, v32);
    DFTB_CHECK_LAUNCH();
    Vc = v32;
    Vx = v32;
    if (prec == 2) {
      unsigned short* hi = (unsigned short*)(v32 + total);
      unsigned short* lo = hi + total;
      hipLaunchKernelGGL(cvtpair_kernel, sgrid(total), dim3(256), 0, st, V, total, hi, lo);
      DFTB_CHECK_LAUNCH();
      Vx = hi;
      Vlo = lo;
      // V^T per vector: [n][b][j] (K-contiguous operand of the transition-charge K2 expand)
      if (op->lc && op->exchange != 1) {
        unsigned short* thi = lo + total;
        unsigned short* tlo = thi + total;
        if ((rc = transpose_pair(V, nvec, op->nocc, op->opad, op->vpad, L.N, op->vpad, thi, tlo, L.N, op->opad, st)))
          return rc;
      }
    }
  }
  // q_c = NULL: the Coulomb term is left to the caller ((A+B)V starts from zero)
  if (op->q_c) {
    if ((rc = coulomb(op, L, ws, Vc, nvec, 4.0, ApBV, st))) return rc;
  } else {
    hipLaunchKernelGGL(zero_kernel, sgrid(total), dim3(256), 0, st, ApBV, total);
    DFTB_CHECK_LAUNCH();
  }

  if (op->lc) {
    double* X1 = (double*)(ws + L.x1);
    double* X2 = (double*)(ws + L.x2);
    hipLaunchKernelGGL(zero_kernel, sgrid(total), dim3(256), 0, st, X1, total);
    hipLaunchKernelGGL(zero_kernel, sgrid(total), dim3(256), 0, st, X2, total);
    DFTB_CHECK_LAUNCH();
    const int it = prec == 0 ? GEMM_IN_F64 : (prec == 1 ? GEMM_IN_F32 : GEMM_IN_BF16X3);
    const int ot = prec == 0 ? GEMM_OUT_F64 : (prec == 1 ? GEMM_OUT_F32 : GEMM_OUT_BF16PAIR);
    const long opad = op->opad, vpad = op->vpad, nat = op->nat;
    char* tu = ws + L.tu;
    const long tu_elems = (long)nvec * L.ac * (opad > vpad ? opad : vpad) * vpad;
    void* T = tu;
    void* Tlo = prec == 2 ? (void*)(tu + tu_elems * 2) : nullptr;
    const long esz = prec == 0 ? 8 : (prec == 1 ? 4 : 2);
    const int nat_eff = (op->exchange == 1) ? 0
                        : ((op->atom_limit > 0 && op->atom_limit < op->nat) ? op->atom_limit : op->nat);
    if (op->exchange == 1) {
      if ((rc = ao_exchange(op, L, ws, nvec, Vx, Vlo, X1, X2, st))) return rc;
    }
    for (int A0 = 0; prec < 2 && A0 < nat_eff; A0 += L.ac) {
      const int ac = (nat_eff - A0 < L.ac) ? (nat_eff - A0) : L.ac;
      // ---- K1 expand: T[n][a'][j][a] = sum_b V[n][j][b] qf_vv[b][A0+a'][a]
      GemmArgs e = gemm_args_default();
      e.in_type = it; e.out_type = ot;
      e.A = Vx; e.A2 = Vlo; e.lda = vpad; e.sA = L.N;
      e.B = (const char*)op->qf_vv[0] + A0 * vpad * esz;
      e.B2 = prec == 2 ? (const char*)op->qf_vv[1] + A0 * vpad * esz : nullptr;
      e.ldb = nat * vpad; e.sB = 0;
      e.C = T; e.C2 = Tlo; e.ldc = vpad; e.cdiv = vpad; e.cstr = opad * vpad; e.sC = (long)ac * opad * vpad;
      e.M = (int)opad; e.N = (int)(ac * vpad); e.K = (int)vpad; e.batch = nvec; e.tile = op->tile_expand;
      if ((rc = gemm_launch(e, st))) return rc;
      // ---- K1 contract: X1[n][i][a] += sum_{a',j} q_oo[i][A0+a'][j] T[n][a'][j][a]
      GemmArgs c = gemm_args_default();
      c.in_type = it; c.out_type = GEMM_OUT_F64;
      c.A = (const char*)op->q_oo[0] + A0 * opad * esz;
      c.A2 = prec == 2 ? (const char*)op->q_oo[1] + A0 * opad * esz : nullptr;
      c.lda = nat * opad; c.sA = 0;
      c.B = T; c.B2 = Tlo; c.ldb = vpad; c.sB = (long)ac * opad * vpad;
      c.C = X1; c.ldc = vpad; c.sC = L.N; c.beta = 1.0;
      c.M = op->nocc; c.N = (int)vpad; c.K = (int)(ac * opad); c.batch = nvec; c.tile = op->tile_contract;
      if ((rc = gemm_launch(c, st))) return rc;
      // ---- K2 expand: U[n][a'][b][a] = sum_j V[n][j][b] qf_ov[j][A0+a'][a]
      GemmArgs e2 = gemm_args_default();
      e2.in_type = it; e2.out_type = ot;
      e2.A = Vx; e2.A2 = Vlo; e2.lda = vpad; e2.sA = L.N; e2.a_trans = 1;
      e2.B = (const char*)op->qf_ov[0] + A0 * vpad * esz;
      e2.B2 = prec == 2 ? (const char*)op->qf_ov[1] + A0 * vpad * esz : nullptr;
      e2.ldb = nat * vpad; e2.sB = 0;
      e2.C = T; e2.C2 = Tlo; e2.ldc = vpad; e2.cdiv = vpad; e2.cstr = vpad * vpad; e2.sC = (long)ac * vpad * vpad;
      e2.M = (int)vpad; e2.N = (int)(ac * vpad); e2.K = op->nocc; e2.batch = nvec; e2.tile = op->tile_expand;
      if ((rc = gemm_launch(e2, st))) return rc;
      // ---- K2 contract: X2[n][i][a] += sum_{a',b} q_ov[i][A0+a'][b] U[n][a'][b][a]
      GemmArgs c2 = gemm_args_default();
      c2.in_type = it; c2.out_type = GEMM_OUT_F64;
      c2.A = (const char*)op->q_ov[0] + A0 * vpad * esz;
      c2.A2 = prec == 2 ? (const char*)op->q_ov[1] + A0 * vpad * esz : nullptr;
      c2.lda = nat * vpad; c2.sA = 0;
      c2.B = T; c2.B2 = Tlo; c2.ldb = vpad; c2.sB = (long)ac * vpad * vpad;
      c2.C = X2; c2.ldc = vpad; c2.sC = L.N; c2.beta = 1.0;
      c2.M = op->nocc; c2.N = (int)vpad; c2.K = (int)(ac * vpad); c2.batch = nvec; c2.tile = op->tile_contract;
      if ((rc = gemm_launch(c2, st))) return rc;
    }
    // BF16x3: NT GEMMs on K-contiguous layouts.  qf_vv is [nat][vpad][vpad] and
    // qf_ov [nat][vpad][opad] (atom-major, see dftb_convert_bf16pair_atom_major);
    // the intermediates are stored transposed, T[n][a][a'][j] and U[n][a][a'][b].
    const unsigned short* VTh = (const unsigned short*)Vlo + total;
    const unsigned short* VTl = VTh + total;
    for (int A0 = 0; prec == 2 && A0 < nat_eff; A0 += L.ac) {
      const int ac = (nat_eff - A0 < L.ac) ? (nat_eff - A0) : L.ac;
      // K1 expand with the slice as the row operand: rows (a', a) of qf_vv, columns j
      // of V; row (a', a) is stored at a * ac * opad + a' * opad, so the 16 lanes of
      // one MFMA row group write 16 consecutive j
      GemmArgs e = gemm_args_default();
      e.in_type = GEMM_IN_BF16X3; e.out_type = GEMM_OUT_BF16PAIR;
      e.A = (const u16_t*)op->qf_vv[0] + A0 * vpad * vpad;
      e.A2 = (const u16_t*)op->qf_vv[1] + A0 * vpad * vpad;
      e.lda = vpad; e.sA = 0;
      e.B = Vx; e.B2 = Vlo; e.b_trans = 1; e.ldb = vpad; e.sB = L.N;
      e.C = T; e.C2 = Tlo; e.rdiv = vpad; e.rstr = opad; e.rs2 = (long)ac * opad;
      e.sC = (long)ac * opad * vpad;
      e.M = (int)(ac * vpad); e.N = (int)opad; e.K = (int)vpad; e.batch = nvec; e.tile = op->tile_expand;
      if ((rc = gemm_launch(e, st))) return rc;
      GemmArgs c = gemm_args_default();
      c.in_type = GEMM_IN_BF16X3; c.out_type = GEMM_OUT_F64;
      c.A = (const u16_t*)op->q_oo[0] + A0 * opad;
      c.A2 = (const u16_t*)op->q_oo[1] + A0 * opad;
      c.lda = nat * opad; c.sA = 0;
      c.B = T; c.B2 = Tlo; c.b_trans = 1; c.ldb = (long)ac * opad; c.sB = (long)ac * opad * vpad;
      c.C = X1; c.ldc = vpad; c.sC = L.N; c.beta = 1.0;
      c.M = op->nocc; c.N = (int)vpad; c.K = (int)(ac * opad); c.batch = nvec; c.tile = op->tile_contract;
      if ((rc = gemm_launch(c, st))) return rc;
      // K2 expand, same orientation: rows (a', a) of qf_ov, columns b of V^T
      GemmArgs e2 = gemm_args_default();
      e2.in_type = GEMM_IN_BF16X3; e2.out_type = GEMM_OUT_BF16PAIR;
      e2.A = (const u16_t*)op->qf_ov[0] + A0 * vpad * opad;
      e2.A2 = (const u16_t*)op->qf_ov[1] + A0 * vpad * opad;
      e2.lda = opad; e2.sA = 0;
      e2.B = VTh; e2.B2 = VTl; e2.b_trans = 1; e2.ldb = opad; e2.sB = L.N;
      e2.C = T; e2.C2 = Tlo; e2.rdiv = vpad; e2.rstr = vpad; e2.rs2 = (long)ac * vpad;
      e2.sC = (long)ac * vpad * vpad;
      e2.M = (int)(ac * vpad); e2.N = (int)vpad; e2.K = (int)opad; e2.batch = nvec; e2.tile = op->tile_expand;
      if ((rc = gemm_launch(e2, st))) return rc;
      GemmArgs c2 = gemm_args_default();
      c2.in_type = GEMM_IN_BF16X3; c2.out_type = GEMM_OUT_F64;
      c2.A = (const u16_t*)op->q_ov[0] + A0 * vpad;
      c2.A2 = (const u16_t*)op->q_ov[1] + A0 * vpad;
      c2.lda = nat * vpad; c2.sA = 0;
      c2.B = T; c2.B2 = Tlo; c2.b_trans = 1; c2.ldb = (long)ac * vpad; c2.sB = (long)ac * vpad * vpad;
      c2.C = X2; c2.ldc = vpad; c2.sC = L.N; c2.beta = 1.0;
      c2.M = op->nocc; c2.N = (int)vpad; c2.K = (int)(ac * vpad); c2.batch = nvec; c2.tile = op->tile_contract;
      if ((rc = gemm_launch(c2, st))) return rc;
    }
    hipLaunchKernelGGL(combine_kernel, sgrid(total), dim3(256), 0, st, op->omega, V, L.N, total, ApBV, AmBV, X1, X2, 1);
  } else {
    hipLaunchKernelGGL(combine_kernel, sgrid(total), dim3(256), 0, st, op->omega, V, L.N, total, ApBV, AmBV,
                       (const double*)nullptr, (const double*)nullptr, 0);
  }
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_hermitian_matvec(const CasidaOperator* op, const double* sqrt_omega, const double* om,
                                     const double* V, int nvec, double* RV, void* workspace, size_t ws_bytes,
                                     void* stream) {
  if (!op || op->nat <= 0 || !op->q_c || !op->gamma || (op->opad & 7) || (op->vpad & 7)) return DFTB_ERR_ARG;
  if (op->precision != 0 && op->precision != 1) return DFTB_ERR_ARG;
  if (nvec <= 0) return DFTB_OK;
  if (!sqrt_omega || !om || !V || !RV || !workspace) return DFTB_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  MvLayout L = mv_layout(op, nvec, true);
  if (ws_bytes < L.total) return DFTB_ERR_WORKSPACE;
  char* ws = (char*)workspace;
  const long total = (long)nvec * L.N;
  double* Vs = (double*)(ws + L.vin);
  hipLaunchKernelGGL(scale_vec_kernel, sgrid(total), dim3(256), 0, st, sqrt_omega, V, L.N, total, Vs);
  DFTB_CHECK_LAUNCH();
  const void* Vx = Vs;
  if (op->precision == 1) {
    float* v32 = (float*)(Vs + total);
    hipLaunchKernelGGL(cvt32_kernel, sgrid(total), dim3(256), 0, st, Vs, total, v32);
    DFTB_CHECK_LAUNCH();
    Vx = v32;
  }
  int rc = coulomb(op, L, ws, Vx, nvec, 4.0, RV, st);
  if (rc) return rc;
  hipLaunchKernelGGL(herm_finish_kernel, sgrid(total), dim3(256), 0, st, om, sqrt_omega, V, L.N, total, RV);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
