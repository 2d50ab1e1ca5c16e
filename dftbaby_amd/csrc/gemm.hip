// Batched, strided MFMA GEMM for gfx950 (FP64 and FP32 operands).
//
//   C[b] (+)= alpha * op(A[b]) * op(B[b]) + beta * C[b]
//
// It carries every dense contraction of the Casida matvec and of the
// Davidson loop: the gamma-fold of the transition charges, the "expand"
// and "contract" halves of the two long-range exchange terms, the Coulomb
// projections, Gram matrices and subspace linear combinations.
//
// Tiling: 256 threads = 4 wave64s in a 2x2 arrangement; each wave owns a
// (BM/2)x(BN/2) block of 16x16 MFMA tiles (v_mfma_f64_16x16x4f64 /
// v_mfma_f32_16x16x4f32).  Operand tiles are staged K-major in LDS
// (double buffered, one barrier per K step) so that the 16 lanes that feed
// one MFMA row/column read 16 consecutive LDS words.
//
// Output modes: plain row-major, or a column and/or row remap where output
// column n is split as (n / cdiv, n % cdiv) and written at (n / cdiv) * cstr +
// (n % cdiv) * cs2 (rows likewise with rdiv, rstr, rs2).  The remap writes the
// exchange intermediates directly in the order the following contraction
// consumes.  Split-K writes
// per-slice partial results at C + slice * sCk (reduced by gemm_reduce).
#include "common.h"
#include "gemm.h"

template <typename T> struct IsF64 { static constexpr bool value = false; };
template <> struct IsF64<double> { static constexpr bool value = true; };
template <typename T> struct AccT;
template <> struct AccT<double> { typedef d4_t type; };
template <> struct AccT<float> { typedef f4_t type; };

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void store_out(const GemmArgs& p, long off, double val) {
  if (p.out_type == GEMM_OUT_F64) {
    double* C = (double*)p.C + off;
    if (p.beta != 0.0) val += p.beta * (*C);
    *C = val;
  } else if (p.out_type == GEMM_OUT_F32) {
    float* C = (float*)p.C + off;
    if (p.beta != 0.0) val += p.beta * (double)(*C);
    *C = (float)val;
  } else {  // bf16 hi/lo pair
    float x = (float)val;
    unsigned short hi = bf16_bits_rne(x);
    unsigned short lo = bf16_bits_rne(x - bf16_bits_to_float(hi));
    ((unsigned short*)p.C)[off] = hi;
    ((unsigned short*)p.C2)[off] = lo;
  }
}

// WGM x (4 / WGM) waves over the BM x BN tile (2 x 2 for the square tiles; 1 x 4
// and 4 x 1 for the 32- and 48-wide tiles of the tall-skinny basis products).
template <typename T, int BM, int BN, int BK, bool AT, bool BT, int WGM = 2>
__global__ __launch_bounds__(256) void gemm_fp_kernel(GemmArgs p) {
  typedef typename AccT<T>::type V4;
  __shared__ T As[2][BK][BM + 1];
  __shared__ T Bs[2][BK][BN + 1];
  constexpr int WGN = 4 / WGM;
  constexpr int WBM = BM / WGM, WBN = BN / WGN;  // wave tile
  constexpr int TM = WBM / 16, TN = WBN / 16;
  static_assert(TM * 16 * WGM == BM && TN * 16 * WGN == BN, "tile / wave layout");
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "tile / thread layout");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  unsigned tx, ty, tz;
  xcd_tile_grouped<4>(tx, ty, tz);
  const long n0 = (long)tx * BN, m0 = (long)ty * BM;
  const int ks = tz % p.ksplit;
  const long b = tz / p.ksplit;
  const T* A = (const T*)p.A + b * p.sA;
  const T* B = (const T*)p.B + b * p.sB;
  const long kbeg = (long)ks * p.kchunk;
  const long kend = min((long)p.K, kbeg + p.kchunk);
  const long M = p.M, N = p.N;

  T ra[NA], rb[NB];
  auto gload = [&](long k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      int idx = tid + 256 * e;
      int kk, r;
      if (!AT) { kk = idx % BK; r = idx / BK; } else { r = idx % BM; kk = idx / BM; }
      long m = m0 + r, k = k0 + kk;
      T v = (T)0;
      if (m < M && k < kend) v = AT ? A[k * p.lda + m] : A[m * p.lda + k];
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      int idx = tid + 256 * e;
      int kk, r;
      if (!BT) { r = idx % BN; kk = idx / BN; } else { kk = idx % BK; r = idx / BK; }
      long n = n0 + r, k = k0 + kk;
      T v = (T)0;
      if (n < N && k < kend) v = BT ? B[n * p.ldb + k] : B[k * p.ldb + n];
      rb[e] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      int idx = tid + 256 * e;
      int kk, r;
      if (!AT) { kk = idx % BK; r = idx / BK; } else { r = idx % BM; kk = idx / BM; }
      As[buf][kk][r] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      int idx = tid + 256 * e;
      int kk, r;
      if (!BT) { r = idx % BN; kk = idx / BN; } else { kk = idx % BK; r = idx / BK; }
      Bs[buf][kk][r] = rb[e];
    }
  };

  V4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (V4)(0);

  const long nk = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    for (long kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) gload(kbeg + (kt + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T af[TM], bf[TN];
        const int kr = kk + (lane >> 4);
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[cur][kr][wm * WBM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][kr][wn * WBN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = mfma16(af[i], bf[j], acc[i][j]);
      }
      if (kt + 1 < nk) sstore(cur ^ 1);
      __syncthreads();
    }
  }

  // output offsets and Hadamard indices split into row and column parts, each
  // formed once per row / column of this lane
  const long cbase = b * p.sC + (long)ks * p.sCk;
  const bool had_on = p.had != nullptr;
  long coff[TN], hcol[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const long col = n0 + wn * WBN + j * 16 + (lane & 15);
    coff[j] = gemm_col_part(p, col);
    hcol[j] = had_on ? gemm_had_col(p, col) : 0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // accumulator row: v_mfma_f64_16x16x4f64 holds rows (lane/16) + 4r,
      // v_mfma_f32_16x16x4f32 rows 4(lane/16) + r (both: column lane%16)
      const int rr = IsF64<T>::value ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
      const long row = m0 + wm * WBM + i * 16 + rr;
      if (row >= M) continue;
      const long ro = cbase + gemm_row_part(p, row);
      const long hr = had_on ? gemm_had_row(p, row) : 0;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const long col = n0 + wn * WBN + j * 16 + (lane & 15);
        if (col < N)
          store_out(p, ro + coff[j], p.alpha * (had_on ? p.had[hr + hcol[j]] : 1.0) * (double)acc[i][j][r]);
      }
    }
}

// Sum split-K partials: C[b][i] = alpha * sum_s P[s][b][i] + beta * C[b][i]  (dense, row-major M x N)
__global__ void gemm_reduce_kernel(const double* P, long sP, int nsplit, double* C, long n, double alpha, double beta) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += P[k * sP + i];
  s *= alpha;
  if (beta != 0.0) s += beta * C[i];
  C[i] = s;
}

template <typename T, int BM, int BN, int BK, int WGM = 2>
static void launch_fp(const GemmArgs& p, hipStream_t st) {
  dim3 grid(ceil_div(p.N, BN), ceil_div(p.M, BM), p.batch * p.ksplit);
  if (!p.a_trans && !p.b_trans) hipLaunchKernelGGL((gemm_fp_kernel<T, BM, BN, BK, false, false, WGM>), grid, dim3(256), 0, st, p);
  else if (p.a_trans && !p.b_trans) hipLaunchKernelGGL((gemm_fp_kernel<T, BM, BN, BK, true, false, WGM>), grid, dim3(256), 0, st, p);
  else if (!p.a_trans && p.b_trans) hipLaunchKernelGGL((gemm_fp_kernel<T, BM, BN, BK, false, true, WGM>), grid, dim3(256), 0, st, p);
  else hipLaunchKernelGGL((gemm_fp_kernel<T, BM, BN, BK, true, true, WGM>), grid, dim3(256), 0, st, p);
}

// ---------------------------------------------------------------------------
// Launch timing: when enabled, every GEMM launch is bracketed by two HIP events
// on its own stream; dftb_gemm_timing_read synchronises them and returns the
// summed device time, launch count and algorithmic flops (2 M N K per batch
// entry; BF16x3 counts its three bf16 products) per kernel class.
#include <vector>
namespace {
struct TimedLaunch {
  hipEvent_t a, b;
  int cls;
  double flops;
};
bool g_timing = false;
std::vector<TimedLaunch> g_launches;
}  // namespace

void GemmTimer::begin(hipStream_t st) {
  slot = -1;
  if (!g_timing) return;
  TimedLaunch t{};
  if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess) return;
  hipEventRecord(t.a, st);
  g_launches.push_back(t);
  slot = (int)g_launches.size() - 1;
}

void GemmTimer::end(hipStream_t st, int cls, double flops) {
  if (slot < 0) return;
  g_launches[slot].cls = cls;
  g_launches[slot].flops = flops;
  hipEventRecord(g_launches[slot].b, st);
}

static void timing_clear() {
  for (auto& t : g_launches) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  g_launches.clear();
}

extern "C" int dftb_gemm_timing_enable(int on) {
  timing_clear();
  g_timing = on != 0;
  return DFTB_OK;
}

extern "C" int dftb_gemm_timing_read(int cls, double* ms_total, long* launches, double* flops) {
  if (cls < 0 || cls >= GT_NCLASS || !ms_total || !launches || !flops) return DFTB_ERR_ARG;
  double ms = 0.0, fl = 0.0;
  long n = 0;
  for (auto& t : g_launches) {
    if (t.cls != cls) continue;
    if (hipEventSynchronize(t.b) != hipSuccess) return DFTB_ERR_LAUNCH;
    float e = 0.0f;
    if (hipEventElapsedTime(&e, t.a, t.b) != hipSuccess) return DFTB_ERR_LAUNCH;
    ms += e;
    fl += t.flops;
    ++n;
  }
  *ms_total = ms;
  *launches = n;
  *flops = fl;
  return DFTB_OK;
}

int gemm_launch(const GemmArgs& p0, hipStream_t st, bool timed) {
  GemmArgs p = p0;
  if (p.M <= 0 || p.N <= 0 || p.batch <= 0) return DFTB_OK;
  if (p.K < 0 || !p.A || !p.B || !p.C) return DFTB_ERR_ARG;
  if (p.ksplit <= 0) p.ksplit = 1;
  if (p.kchunk <= 0) p.kchunk = p.K;
  if (p.out_type == GEMM_OUT_BF16PAIR && !p.C2) return DFTB_ERR_ARG;
  if (p.had && (p.hmod <= 0 || p.ldh <= 0)) return DFTB_ERR_ARG;  // Hadamard keys are taken modulo hmod
  const bool big = (p.M >= 128 && p.N >= 128) || p.tile == 128;
  const bool small = p.tile == 64 || !big;
  if (p.in_type == GEMM_IN_BF16X3) return gemm_bf16x3_launch(p, st);
  if (p.Ct) return DFTB_ERR_ARG;  // transposed copy: pipelined BF16x3 kernel only
  if (p.in_type != GEMM_IN_F64 && p.in_type != GEMM_IN_F32) return DFTB_ERR_ARG;
  GemmTimer tm;
  tm.slot = -1;
  if (timed) tm.begin(st);
  if (p.in_type == GEMM_IN_F64) {
    // tall-skinny products of the Davidson basis (one side 20-40 vectors, the other
    // hundreds of basis vectors or the pair index): narrow tiles cut the padded
    // MFMA work, which bounds these calls at the FP64 MFMA rate
    if (small && p.M <= 16 && p.N > 48) launch_fp<double, 16, 64, 16, 1>(p, st);
    else if (small && p.M <= 32 && p.N > 48) launch_fp<double, 32, 64, 16, 1>(p, st);
    else if (small && p.M <= 48 && p.N > 48) launch_fp<double, 48, 64, 16, 1>(p, st);
    else if (small && p.N <= 32 && p.M > 48) launch_fp<double, 64, 32, 16, 4>(p, st);
    else if (small && p.N <= 48 && p.M > 48) launch_fp<double, 64, 48, 16, 4>(p, st);
    else if (small && p.N <= 80 && p.M > 80) launch_fp<double, 64, 80, 16, 4>(p, st);   // both Gram halves of a block
    else if (small && p.M <= 48 && p.N <= 48 && p.M > 32) launch_fp<double, 48, 64, 16, 1>(p, st);  // a block's own Gram
    else if (small) launch_fp<double, 64, 64, 16>(p, st);
    else launch_fp<double, 128, 128, 16>(p, st);
  } else {
    // the Coulomb projection / expansion of a vector block: M = block size (<= 48 rows
    // per tile instead of 64 cuts the padded MFMA work by a quarter)
    if (small && p.M <= 48 && p.M > 32 && p.N >= 4096) launch_fp<float, 48, 128, 32, 1>(p, st);
    else if (small && p.M <= 48 && p.M > 32) launch_fp<float, 48, 64, 32, 1>(p, st);
    else if (small && p.M <= 16) launch_fp<float, 16, 64, 32, 1>(p, st);
    else if (small && p.M <= 32) launch_fp<float, 32, 64, 32, 1>(p, st);
    else if (small) launch_fp<float, 64, 64, 32>(p, st);
    else launch_fp<float, 128, 128, 32>(p, st);
  }
  const int cls = p.in_type == GEMM_IN_F64
                      ? (small ? GT_F64_64 : ((p.a_trans || p.b_trans) ? GT_F64_128_T : GT_F64_128))
                      : (small ? GT_F32_64 : GT_F32_128);
  tm.end(st, cls, 2.0 * p.M * p.N * (double)p.K * p.batch);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

int gemm_reduce_launch(const double* P, long sP, int nsplit, double* C, long n, double alpha, double beta, hipStream_t st) {
  if (n <= 0) return DFTB_OK;
  hipLaunchKernelGGL(gemm_reduce_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, P, sP, nsplit, C, n, alpha, beta);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_gemm(const GemmArgs* p, void* stream) {
  if (!p) return DFTB_ERR_ARG;
  return gemm_launch(*p, (hipStream_t)stream);
}
