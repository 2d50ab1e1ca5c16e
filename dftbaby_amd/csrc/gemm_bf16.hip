// BF16x3 split-precision MFMA GEMM for gfx950.
//
// Each operand x is carried as a pair (hi, lo) of bf16 with hi = bf16(x) and
// lo = bf16(x - hi), which keeps ~16 significant bits.  The product is
//   A B ~= A_hi B_hi + A_hi B_lo + A_lo B_hi
// with FP32 accumulation in v_mfma_f32_16x16x32_bf16 (the dropped lo*lo term
// is below 2^-16 relative).  Three bf16 MFMAs cost 3/16 of one FP32 MFMA at the
// MI355X dense rates (2.5 PFLOP/s bf16 vs 157 TFLOP/s fp32).
//
// Form: C = A B^T ("NT"): A[m][k] and B[n][k] both K-contiguous, so every lane
// loads its 8 consecutive k values for one row with a single 16-byte LDS read,
// exactly the operand shape of the 16x16x32 MFMA (lane l: row l%16,
// k = 8*(l/16) .. +7).  Tiles: 256 threads, BM x BN = 128 x 128, BK = 32, each
// wave a 64 x 64 block of 4x4 MFMA tiles; global->LDS staging with 16-byte
// loads, double buffered.
#include "common.h"
#include "gemm.h"

#define BF_BK 32
#define BF_LDS_ROW 40  // BK + 8 bf16 padding (80-byte rows keep 16-byte alignment)

typedef unsigned short u16;
typedef uint4 v16b;  // 16 bytes = 8 bf16

__device__ __forceinline__ void store_out_bf(const GemmArgs& p, long off, float val) {
  if (p.out_type == GEMM_OUT_F64) {
    double* C = (double*)p.C + off;
    double v = (double)val;
    if (p.beta != 0.0) v += p.beta * (*C);
    *C = v;
  } else if (p.out_type == GEMM_OUT_F32) {
    float* C = (float*)p.C + off;
    float v = val;
    if (p.beta != 0.0) v += (float)p.beta * (*C);
    *C = v;
  } else {
    u16 hi = bf16_bits_rne(val);
    u16 lo = bf16_bits_rne(val - bf16_bits_to_float(hi));
    ((u16*)p.C)[off] = hi;
    ((u16*)p.C2)[off] = lo;
  }
}

template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_bf3_kernel(GemmArgs p) {
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int VA = BM * BF_BK / 8 / 256;  // 16-byte vectors per thread per operand array
  constexpr int VB = BN * BF_BK / 8 / 256;
  __shared__ __attribute__((aligned(16))) u16 Ah[2][BM][BF_LDS_ROW];
  __shared__ __attribute__((aligned(16))) u16 Al[2][BM][BF_LDS_ROW];
  __shared__ __attribute__((aligned(16))) u16 Bh[2][BN][BF_LDS_ROW];
  __shared__ __attribute__((aligned(16))) u16 Bl[2][BN][BF_LDS_ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  unsigned tx, ty, tz;
  xcd_tile(tx, ty, tz);
  const long n0 = (long)tx * BN, m0 = (long)ty * BM;
  const int ks = tz % p.ksplit;
  const long b = tz / p.ksplit;
  const u16* Ahi = (const u16*)p.A + b * p.sA;
  const u16* Alo = (const u16*)p.A2 + b * p.sA;
  const u16* Bhi = (const u16*)p.B + b * p.sB;
  const u16* Blo = (const u16*)p.B2 + b * p.sB;
  const long kbeg = (long)ks * p.kchunk;
  const long kend = min((long)p.K, kbeg + p.kchunk);
  const long M = p.M, N = p.N;

  v16b rah[VA], ral[VA], rbh[VB], rbl[VB];
  const v16b zero = make_uint4(0u, 0u, 0u, 0u);
  auto gload = [&](long k0) {
#pragma unroll
    for (int e = 0; e < VA; ++e) {
      int idx = tid + 256 * e;
      int r = idx >> 2, kv = (idx & 3) * 8;
      long m = m0 + r, k = k0 + kv;
      bool ok = (m < M) && (k < kend);
      rah[e] = ok ? *(const v16b*)(Ahi + m * p.lda + k) : zero;
      ral[e] = ok ? *(const v16b*)(Alo + m * p.lda + k) : zero;
    }
#pragma unroll
    for (int e = 0; e < VB; ++e) {
      int idx = tid + 256 * e;
      int r = idx >> 2, kv = (idx & 3) * 8;
      long n = n0 + r, k = k0 + kv;
      bool ok = (n < N) && (k < kend);
      rbh[e] = ok ? *(const v16b*)(Bhi + n * p.ldb + k) : zero;
      rbl[e] = ok ? *(const v16b*)(Blo + n * p.ldb + k) : zero;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int e = 0; e < VA; ++e) {
      int idx = tid + 256 * e;
      int r = idx >> 2, kv = (idx & 3) * 8;
      *(v16b*)&Ah[buf][r][kv] = rah[e];
      *(v16b*)&Al[buf][r][kv] = ral[e];
    }
#pragma unroll
    for (int e = 0; e < VB; ++e) {
      int idx = tid + 256 * e;
      int r = idx >> 2, kv = (idx & 3) * 8;
      *(v16b*)&Bh[buf][r][kv] = rbh[e];
      *(v16b*)&Bl[buf][r][kv] = rbl[e];
    }
  };

  f4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = (f4_t)(0.0f);

  const long nk = (kend > kbeg) ? (kend - kbeg + BF_BK - 1) / BF_BK : 0;
  if (nk > 0) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    for (long kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) gload(kbeg + (kt + 1) * BF_BK);
      bf8_t ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm * (BM / 2) + i * 16 + (lane & 15);
        ah[i] = *(const bf8_t*)&Ah[cur][r][kq];
        al[i] = *(const bf8_t*)&Al[cur][r][kq];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = wn * (BN / 2) + j * 16 + (lane & 15);
        bh[j] = *(const bf8_t*)&Bh[cur][r][kq];
        bl[j] = *(const bf8_t*)&Bl[cur][r][kq];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      if (kt + 1 < nk) sstore(cur ^ 1);
      __syncthreads();
    }
  }

  const long cbase = b * p.sC + (long)ks * p.sCk;
  const float alpha = (float)p.alpha;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        long row = m0 + wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
        long col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
        if (row < M && col < N) {
          long off = (p.cdiv > 0) ? (col / p.cdiv) * p.cstr + (col % p.cdiv) * p.cs2 + row * p.ldc
                                  : row * p.ldc + col;
          store_out_bf(p, cbase + off, alpha * acc[i][j][r]);
        }
      }
}

int gemm_bf16x3_launch(const GemmArgs& p, hipStream_t st) {
  if (p.a_trans || !p.b_trans) return DFTB_ERR_ARG;
  if (!p.A2 || !p.B2) return DFTB_ERR_ARG;
  if ((p.lda & 7) || (p.ldb & 7) || (p.K & 7) || (p.sA & 7) || (p.sB & 7)) return DFTB_ERR_ARG;
  if (p.ksplit > 1 && (p.kchunk & 7)) return DFTB_ERR_ARG;
  if ((((uintptr_t)p.A) | ((uintptr_t)p.A2) | ((uintptr_t)p.B) | ((uintptr_t)p.B2)) & 15) return DFTB_ERR_ARG;
  const bool small = p.tile == 64 || (p.tile == 0 && (p.M < 128 || p.N < 128));
  if (small) {
    dim3 grid(ceil_div(p.N, 64), ceil_div(p.M, 64), p.batch * p.ksplit);
    hipLaunchKernelGGL((gemm_bf3_kernel<64, 64>), grid, dim3(256), 0, st, p);
  } else {
    dim3 grid(ceil_div(p.N, 128), ceil_div(p.M, 128), p.batch * p.ksplit);
    hipLaunchKernelGGL((gemm_bf3_kernel<128, 128>), grid, dim3(256), 0, st, p);
  }
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
