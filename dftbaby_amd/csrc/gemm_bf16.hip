// BF16x3 split-precision MFMA GEMM for gfx950.
//
// Each operand x is carried as a pair (hi, lo) of bf16 with hi = bf16(x) and
// lo = bf16(x - hi), which keeps ~16 significant bits.  The product is
//   A B ~= A_hi B_hi + A_hi B_lo + A_lo B_hi
// with FP32 accumulation in bf16 MFMAs (the dropped lo*lo term is below 2^-16
// relative).  Three bf16 MFMAs cost 3/16 of one FP32 MFMA at the MI355X dense
// rates (2.5 PFLOP/s bf16 vs 157 TFLOP/s fp32).
//
// Measured limits on gfx950 (register-only loops): v_mfma_f32_16x16x32_bf16
// sustains about 1.46 PFLOP/s, v_mfma_f32_32x32x16_bf16 about 2.46 PFLOP/s.  The
// pipelined kernel below is bound by its global->LDS feed (with the DMA removed
// the same loop reaches about 74% of 2.5 PFLOP/s, with the MFMAs removed it takes
// as long as the full kernel), which is why its tile order is grouped per XCD.
//
// Form: C = A B^T ("NT"): A[m][k] and B[n][k] both K-contiguous, so every lane
// loads its 8 consecutive k values for one row with a single 16-byte LDS read,
// exactly the operand shape of the 16x16x32 MFMA (lane l: row l%16,
// k = 8*(l/16) .. +7).  Tiles: 256 threads, BM x BN = 128 x 128, BK = 32, each
// wave a 64 x 64 block of 4x4 MFMA tiles; global->LDS staging with 16-byte
// loads, double buffered.
#include <stdlib.h>
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
  const bool had_on = p.had != nullptr;
  long coff[TN], hcol[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const long col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
    coff[j] = gemm_col_part(p, col);
    hcol[j] = had_on ? gemm_had_col(p, col) : 0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long row = m0 + wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
      if (row >= M) continue;
      const long ro = cbase + gemm_row_part(p, row);
      const long hr = had_on ? gemm_had_row(p, row) : 0;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const long col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
        if (col < N)
          store_out_bf(p, ro + coff[j], alpha * (had_on ? (float)p.had[hr + hcol[j]] : 1.0f) * acc[i][j][r]);
      }
    }
}

// ---------------------------------------------------------------------------
// Large-tile pipelined variant: 256 x 256 (or 256 x 128) output tile per
// workgroup of eight wave64s (4 along M x 2 along N), BK = 32, a two-stage (or
// three-stage) LDS ring filled by direct buffer->LDS loads (buffer_load_dwordx4
// ... lds).  Each wave instruction deposits 64 x 16 bytes contiguously in LDS;
// the lanes choose their global 16-byte chunk so that LDS row r holds k-chunk c
// in slot c ^ ((r >> 2) & 3), which makes the 16-byte fragment reads of 16 rows
// at one k-chunk hit disjoint banks.  Out-of-range rows and k-chunks get an
// offset past the buffer's num_records and load zeros.  The DMA is issued as
// inline asm so the compiler does not wait for it; the cross-wave hand-off is
// an explicit vmcnt wait followed by s_barrier.
#define P_BM 256
#define P_BK 32

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u16 lds_u16_t;

// raw buffer descriptor (stride 0, 32-bit data format) over [base, base + bytes)
__device__ __forceinline__ i4_t make_rsrc(const void* base, unsigned bytes) {
  const unsigned long long a = (unsigned long long)base;
  i4_t r;  // workgroup-uniform: pinned to scalar registers for the "s" operand of dma16
  r.x = __builtin_amdgcn_readfirstlane((int)(a & 0xffffffffu));
  r.y = __builtin_amdgcn_readfirstlane((int)((a >> 32) & 0xffffu));
  r.z = __builtin_amdgcn_readfirstlane((int)bytes);
  r.w = 0x00020000;
  return r;
}

__device__ __forceinline__ unsigned lds_addr(const u16* p) { return (unsigned)(unsigned long long)(lds_u16_t*)p; }

// 16 bytes per lane from buffer offset `off` into LDS at m0 + 16 * lane.  Issued
// as inline asm so that the compiler's waitcnt insertion does not serialise the
// prefetch behind later LDS reads; completion is ordered by P_WAIT_BARRIER.
__device__ __forceinline__ void dma16(i4_t rs, unsigned lds, unsigned off) {
  asm volatile("s_mov_b32 m0, %0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds" ::"s"(lds), "v"(off), "s"(rs)
               : "memory");
}

// wait until at most N buffer loads of this wave are outstanding and its LDS
// reads have landed, then barrier
#define P_WAIT_BARRIER(N) asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory")

// Hadamard factor of one output tile, read from LDS.  The factor of element
// (row, col) is had[ra(row) ldh + cb(col)] with ra, cb the row / column keys
// (hmap of row % hmod, or row % hmod itself).  After the main loop the ring is
// free: the key ranges [a0, a1] x [b0, b1] of the tile's rows and columns are
// reduced in LDS, and when that block of the factor fits, it is copied there
// as float once per tile (row stride nc) instead of being gathered from global
// memory per element.  Otherwise (scattered keys) the epilogue reads global
// memory as before.  Block-uniform result; call with every thread.
struct HadTile {
  const float* s;  // LDS block, or nullptr: read p.had
  int a0, b0;      // key offsets subtracted before indexing
  long ld;         // row stride of the block (nc, or p.ldh)
};
__device__ __forceinline__ int had_rkey(const GemmArgs& p, long row) {
  const unsigned rm = (unsigned)row % (unsigned)p.hmod;
  return p.hmap ? p.hmap[rm] : (int)rm;
}
template <int BM, int BN, int NT>
__device__ HadTile had_stage(const GemmArgs& p, long m0, long n0, u16* ring, int cap_floats) {
  const int tid = threadIdx.x;
  int* hb = (int*)ring;
  float* hs = (float*)ring + 64;
  __syncthreads();  // every wave is done reading the ring
  if (tid == 0) { hb[0] = 0x7fffffff; hb[1] = -1; hb[2] = 0x7fffffff; hb[3] = -1; }
  __syncthreads();
  int k[4] = {0x7fffffff, -1, 0x7fffffff, -1};  // min / max row key, min / max column key
  for (int r = tid; r < BM; r += NT)
    if (m0 + r < p.M) { const int a = had_rkey(p, m0 + r); k[0] = min(k[0], a); k[1] = max(k[1], a); }
  for (int c = tid; c < BN; c += NT)
    if (n0 + c < p.N) { const int b = had_rkey(p, n0 + c); k[2] = min(k[2], b); k[3] = max(k[3], b); }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    k[0] = min(k[0], __shfl_xor(k[0], o));
    k[1] = max(k[1], __shfl_xor(k[1], o));
    k[2] = min(k[2], __shfl_xor(k[2], o));
    k[3] = max(k[3], __shfl_xor(k[3], o));
  }
  if ((tid & 63) == 0) {
    atomicMin(&hb[0], k[0]); atomicMax(&hb[1], k[1]);
    atomicMin(&hb[2], k[2]); atomicMax(&hb[3], k[3]);
  }
  __syncthreads();
  const int a0 = hb[0], b0 = hb[2];
  const int nr = hb[1] - a0 + 1, nc = hb[3] - b0 + 1;
  HadTile h;
  if ((long)nr * nc > (long)(cap_floats - 64)) {
    h.s = nullptr; h.a0 = 0; h.b0 = 0; h.ld = p.ldh;
    return h;
  }
  const int wave = tid >> 6, lane = tid & 63;
  for (int r = wave; r < nr; r += NT / 64) {
    const double* src = p.had + (long)(a0 + r) * p.ldh + b0;
    for (int c = lane; c < nc; c += 64) hs[r * nc + c] = (float)src[c];
  }
  __syncthreads();
  h.s = hs; h.a0 = a0; h.b0 = b0; h.ld = nc;
  return h;
}

// NW wave64s per workgroup: NW / 2 along M, 2 along N; wave tile
// (256 / (NW / 2)) x (BN / 2) computed with v_mfma_f32_32x32x16_bf16 (M32) or
// with v_mfma_f32_16x16x32_bf16.  BN = 128: three-stage ring of 48 KiB stages;
// BN = 256: two-stage ring of 64 KiB stages (the ring must fit the 160 KiB of
// LDS), 1.5x the MFMA work per byte staged into LDS.
template <int NW, bool M32, int BN, bool TR = false>
__global__ __launch_bounds__(NW * 64) void gemm_bf3_pipe_kernel(GemmArgs p) {
  static_assert(BN == 128 || BN == 256, "pipelined tile width");
  static_assert(!TR || !M32, "transposed copy: 16x16 form");  // TR: bf16-pair output plus transposed copy
  constexpr int NST = BN == 128 ? 3 : 2;     // ring stages
  constexpr int WMW = NW / 2;                // waves along M
  constexpr int WROWS = P_BM / WMW;          // rows per wave (64 or 128)
  constexpr int WCOLS = BN / 2;              // columns per wave (64 or 128)
  constexpr int TM = WROWS / 16;             // 16x16 tiles per wave along M
  constexpr int TN = WCOLS / 16;             // 16x16 tiles per wave along N
  constexpr int TM32 = WROWS / 32;           // 32x32 tiles per wave along M
  constexpr int TN32 = WCOLS / 32;           // 32x32 tiles per wave along N
  constexpr int CA = 16 / NW, CB = (BN / 16) / NW;  // 1 KB DMA chunks per wave per A / B array
  constexpr int LOADS = 2 * CA + 2 * CB;     // DMA instructions per wave per stage
  // ring of NST stages, each [A hi | A lo | B hi | B lo]
  constexpr int SA = P_BM * P_BK, SB = BN * P_BK, STAGE = 2 * SA + 2 * SB;
  __shared__ __attribute__((aligned(1024))) u16 ring[NST * STAGE];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the LDS destination goes to M0
  const int wm = wave >> 1, wn = wave & 1;
  unsigned tx, ty, tz;
  xcd_tile_grouped<4>(tx, ty, tz);   // 4 x 8 tiles (1024 x 8 BN) per XCD at a time
  const long n0 = (long)tx * BN, m0 = (long)ty * P_BM;
  const int ks = tz % p.ksplit;
  const long b = tz / p.ksplit;
  const long kbeg = (long)ks * p.kchunk;
  const long kend = min((long)p.K, kbeg + p.kchunk);
  const long M = p.M, N = p.N;
  const long mrows = min((long)P_BM, M - m0), nrows = min((long)BN, N - n0);

  // buffer descriptors over this workgroup's rows; an offset equal to num_records loads zeros
  const unsigned nbA = (unsigned)(((mrows - 1) * p.lda + kend) * 2);
  const unsigned nbB = (unsigned)(((nrows - 1) * p.ldb + kend) * 2);
  const i4_t rah = make_rsrc((const u16*)p.A + b * p.sA + m0 * p.lda, nbA);
  const i4_t ral = make_rsrc((const u16*)p.A2 + b * p.sA + m0 * p.lda, nbA);
  const i4_t rbh = make_rsrc((const u16*)p.B + b * p.sB + n0 * p.ldb, nbB);
  const i4_t rbl = make_rsrc((const u16*)p.B2 + b * p.sB + n0 * p.ldb, nbB);
  // this lane's chunk: row 16 * ch + lane / 4 of the tile, k-chunk (lane & 3) ^ ((lane >> 4) & 3)
  const int cl = (lane & 3) ^ ((lane >> 4) & 3);
  unsigned oA[CA], oB[CB];
  {
    const long r = lane >> 2;
#pragma unroll
    for (int i = 0; i < CA; ++i) {
      const long rr = (wave * CA + i) * 16 + r;
      oA[i] = rr < mrows ? (unsigned)((rr * p.lda + kbeg + 8 * cl) * 2) : nbA;
    }
#pragma unroll
    for (int i = 0; i < CB; ++i) {
      const long rr = (wave * CB + i) * 16 + r;
      oB[i] = rr < nrows ? (unsigned)((rr * p.ldb + kbeg + 8 * cl) * 2) : nbB;
    }
  }
  const long kl = kbeg + 8 * cl;  // k of this lane's chunk at stage 0

#define P_LD(rs, base, ch, off) dma16(rs, lds_addr((base) + (ch) * 512), off)
#define P_ISSUE_A(kt, Ah, Al)                                                        \
  do {                                                                               \
    const bool kok_ = kl + (long)(kt) * P_BK < kend;                                 \
    const unsigned kb_ = (unsigned)((kt) * P_BK * 2);                                \
    _Pragma("unroll") for (int i = 0; i < CA; ++i) {                                 \
      const unsigned o_ = (kok_ && oA[i] != nbA) ? oA[i] + kb_ : nbA;                \
      P_LD(rah, Ah, wave * CA + i, o_);                                              \
      P_LD(ral, Al, wave * CA + i, o_);                                              \
    }                                                                                \
  } while (0)
#define P_ISSUE_B(kt, Bh, Bl)                                                        \
  do {                                                                               \
    const bool kok_ = kl + (long)(kt) * P_BK < kend;                                 \
    const unsigned kb_ = (unsigned)((kt) * P_BK * 2);                                \
    _Pragma("unroll") for (int i = 0; i < CB; ++i) {                                 \
      const unsigned o_ = (kok_ && oB[i] != nbB) ? oB[i] + kb_ : nbB;                \
      P_LD(rbh, Bh, wave * CB + i, o_);                                              \
      P_LD(rbl, Bl, wave * CB + i, o_);                                              \
    }                                                                                \
  } while (0)
#define P_ISSUE(kt, Ah, Al, Bh, Bl) \
  do {                              \
    P_ISSUE_A(kt, Ah, Al);          \
    P_ISSUE_B(kt, Bh, Bl);          \
  } while (0)


  const int nk = (kend > kbeg) ? (int)((kend - kbeg + P_BK - 1) / P_BK) : 0;
#define P_SLOT(st) (ring + (st) * STAGE)
  // Stages 0 .. NST - 2 are issued up front.  At the top of step kt the loads of
  // stages kt .. min(kt + NST - 2, nk - 1) are in flight; stage kt is complete once
  // at most the later ones remain (vmcnt counts this wave's loads in issue order),
  // and the barrier also tells every wave that all reads of step kt - 1 are done,
  // so the slot of stage kt + NST - 1 (= that of kt - 1) may be refilled.
#define P_MAINLOOP(COMPUTE)                                                           \
  if (nk > 0) {                                                                       \
    { u16* S0 = P_SLOT(0); P_ISSUE(0, S0, S0 + SA, S0 + 2 * SA, S0 + 2 * SA + SB); }  \
    if (NST == 3 && nk > 1) { u16* S1 = P_SLOT(1); P_ISSUE(1, S1, S1 + SA, S1 + 2 * SA, S1 + 2 * SA + SB); } \
    int cur = 0;                                                                      \
    for (int kt = 0; kt < nk; ++kt) {                                                 \
      if (NST == 3 && kt + 1 < nk) P_WAIT_BARRIER(LOADS); else P_WAIT_BARRIER(0);     \
      if (kt + NST - 1 < nk) {                                                        \
        const int nx = cur == 0 ? NST - 1 : cur - 1; /* slot of stage kt + NST - 1 */ \
        u16* S = P_SLOT(nx);                                                          \
        P_ISSUE(kt + NST - 1, S, S + SA, S + 2 * SA, S + 2 * SA + SB);                \
      }                                                                               \
      const u16* C = P_SLOT(cur);                                                     \
      COMPUTE(C, C + SA, C + 2 * SA, C + 2 * SA + SB);                                \
      cur = cur == NST - 1 ? 0 : cur + 1;                                             \
    }                                                                                 \
  }
  const long cbase = b * p.sC + (long)ks * p.sCk;
  const float alpha = (float)p.alpha;
  const bool had_on = p.had != nullptr;

  if constexpr (M32) {
    // fragment of a 32x32x16 MFMA: lane l reads row l % 32 at k-chunk 2 s + l / 32
    // (k-step s of the stage), stored in slot chunk ^ ((row >> 2) & 3)
    typedef float f16v_t __attribute__((ext_vector_type(16)));
    f16v_t acc[TM32][TN32];
#pragma unroll
    for (int i = 0; i < TM32; ++i)
#pragma unroll
      for (int j = 0; j < TN32; ++j) acc[i][j] = (f16v_t)(0.0f);
    const int sw = (lane >> 2) & 3;
    const int sl0 = ((lane >> 5) ^ sw) * 8, sl1 = ((2 + (lane >> 5)) ^ sw) * 8;
    const int arow = (wm * WROWS + (lane & 31)) * P_BK;
    const int brow = (wn * WCOLS + (lane & 31)) * P_BK;
#define P_COMPUTE32(Ah, Al, Bh, Bl)                                                              \
  do {                                                                                           \
    bf8_t ah_[TM32][2], al_[TM32][2], bh_[TN32][2], bl_[TN32][2];                                      \
    _Pragma("unroll") for (int i = 0; i < TM32; ++i) {                                           \
      ah_[i][0] = *(const bf8_t*)((Ah) + arow + i * 32 * P_BK + sl0);                            \
      ah_[i][1] = *(const bf8_t*)((Ah) + arow + i * 32 * P_BK + sl1);                            \
      al_[i][0] = *(const bf8_t*)((Al) + arow + i * 32 * P_BK + sl0);                            \
      al_[i][1] = *(const bf8_t*)((Al) + arow + i * 32 * P_BK + sl1);                            \
    }                                                                                            \
    _Pragma("unroll") for (int j = 0; j < TN32; ++j) {                                              \
      bh_[j][0] = *(const bf8_t*)((Bh) + brow + j * 32 * P_BK + sl0);                            \
      bh_[j][1] = *(const bf8_t*)((Bh) + brow + j * 32 * P_BK + sl1);                            \
      bl_[j][0] = *(const bf8_t*)((Bl) + brow + j * 32 * P_BK + sl0);                            \
      bl_[j][1] = *(const bf8_t*)((Bl) + brow + j * 32 * P_BK + sl1);                            \
    }                                                                                            \
    /* product-major order: independent MFMAs separate two updates of one accumulator */        \
    _Pragma("unroll") for (int s_ = 0; s_ < 2; ++s_) {                                           \
      _Pragma("unroll") for (int i = 0; i < TM32; ++i) _Pragma("unroll") for (int j = 0; j < TN32; ++j) \
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al_[i][s_], bh_[j][s_], acc[i][j], 0, 0, 0); \
      _Pragma("unroll") for (int i = 0; i < TM32; ++i) _Pragma("unroll") for (int j = 0; j < TN32; ++j) \
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah_[i][s_], bl_[j][s_], acc[i][j], 0, 0, 0); \
      _Pragma("unroll") for (int i = 0; i < TM32; ++i) _Pragma("unroll") for (int j = 0; j < TN32; ++j) \
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah_[i][s_], bh_[j][s_], acc[i][j], 0, 0, 0); \
    }                                                                                            \
  } while (0)
    P_MAINLOOP(P_COMPUTE32)
#undef P_COMPUTE32
    const HadTile ht = had_on ? had_stage<P_BM, BN, NW * 64>(p, m0, n0, ring, NST * STAGE / 2)
                              : HadTile{nullptr, 0, 0, 0};
    // epilogue: accumulator entry q of tile (i, j) is row 8 (q / 4) + 4 (l / 32) + q % 4,
    // column l % 32.  Column parts of the offsets and of the Hadamard index are formed
    // once; each tile row block loads its Hadamard factors before its stores.
    long coff[TN32], hcol[TN32];
#pragma unroll
    for (int j = 0; j < TN32; ++j) {
      const long col = n0 + wn * WCOLS + j * 32 + (lane & 31);
      coff[j] = gemm_col_part(p, col);
      hcol[j] = had_on ? (long)(had_rkey(p, col) - ht.b0) : 0;
    }
#define P_EPI32(STMT)                                                                  \
  _Pragma("unroll") for (int i = 0; i < TM32; ++i) {                                   \
    float hv_[16][TN32];                                                                  \
    _Pragma("unroll") for (int q = 0; q < 16; ++q) {                                   \
      const long row = m0 + wm * WROWS + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3); \
      const long hr = had_on ? (long)(had_rkey(p, row) - ht.a0) * ht.ld : 0;           \
      if (!had_on) {                                                                   \
        _Pragma("unroll") for (int j = 0; j < TN32; ++j) hv_[q][j] = 1.0f;             \
      } else if (ht.s) {                                                               \
        _Pragma("unroll") for (int j = 0; j < TN32; ++j) hv_[q][j] = ht.s[hr + hcol[j]]; \
      } else {                                                                         \
        _Pragma("unroll") for (int j = 0; j < TN32; ++j) hv_[q][j] = (float)p.had[hr + hcol[j]]; \
      }                                                                                \
    }                                                                                  \
    _Pragma("unroll") for (int q = 0; q < 16; ++q) {                                   \
      const long row = m0 + wm * WROWS + i * 32 + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3); \
      if (row < M) {                                                                   \
        const long ro = cbase + gemm_row_part(p, row);                                 \
        _Pragma("unroll") for (int j = 0; j < TN32; ++j) {                                \
          const long col = n0 + wn * WCOLS + j * 32 + (lane & 31);                        \
          if (col < N) {                                                               \
            const long off = ro + coff[j];                                             \
            const float val = alpha * hv_[q][j] * acc[i][j][q];                        \
            STMT;                                                                      \
          }                                                                            \
        }                                                                              \
      }                                                                                \
    }                                                                                  \
  }
    if (p.out_type == GEMM_OUT_F64) {
      double* Cd = (double*)p.C;
      if (p.beta != 0.0) {
        const double beta = p.beta;
        P_EPI32(Cd[off] = (double)val + beta * Cd[off]);
      } else {
        P_EPI32(Cd[off] = (double)val);
      }
    } else if (p.out_type == GEMM_OUT_F32) {
      float* Cf = (float*)p.C;
      const float beta = (float)p.beta;
      P_EPI32(Cf[off] = beta != 0.0f ? val + beta * Cf[off] : val);
    } else {
      u16* Ch = (u16*)p.C;
      u16* Cl = (u16*)p.C2;
      P_EPI32(const u16 h = bf16_bits_rne(val); Ch[off] = h; Cl[off] = bf16_bits_rne(val - bf16_bits_to_float(h)));
    }
#undef P_EPI32
  } else {
    f4_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = (f4_t)(0.0f);
    // fragment reads: row r = 16 q + (lane & 15) holds k-chunk c = lane >> 4 in slot c ^ ((lane >> 2) & 3)
    const int slot = ((lane >> 4) ^ ((lane >> 2) & 3)) * 8;
    const int arow = (wm * WROWS + (lane & 15)) * P_BK + slot;
    const int brow = (wn * WCOLS + (lane & 15)) * P_BK + slot;
    // B fragments in groups of four columns of 16 (one group per 64 columns) so the
    // 256-wide tile's fragments and accumulators fit the 256 registers of a wave
#define P_COMPUTE(Ah, Al, Bh, Bl)                                                                    \
  do {                                                                                               \
    bf8_t ah_[TM], al_[TM];                                                                          \
    _Pragma("unroll") for (int i = 0; i < TM; ++i) {                                                 \
      ah_[i] = *(const bf8_t*)((Ah) + arow + i * 16 * P_BK);                                         \
      al_[i] = *(const bf8_t*)((Al) + arow + i * 16 * P_BK);                                         \
    }                                                                                                \
    _Pragma("unroll") for (int g_ = 0; g_ < TN / 4; ++g_) {                                          \
      bf8_t bh_[4], bl_[4];                                                                          \
      _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                                \
        bh_[j] = *(const bf8_t*)((Bh) + brow + (4 * g_ + j) * 16 * P_BK);                            \
        bl_[j] = *(const bf8_t*)((Bl) + brow + (4 * g_ + j) * 16 * P_BK);                            \
      }                                                                                              \
      _Pragma("unroll") for (int i = 0; i < TM; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)   \
        acc[i][4 * g_ + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al_[i], bh_[j], acc[i][4 * g_ + j], 0, 0, 0); \
      _Pragma("unroll") for (int i = 0; i < TM; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)   \
        acc[i][4 * g_ + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah_[i], bl_[j], acc[i][4 * g_ + j], 0, 0, 0); \
      _Pragma("unroll") for (int i = 0; i < TM; ++i) _Pragma("unroll") for (int j = 0; j < 4; ++j)   \
        acc[i][4 * g_ + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah_[i], bh_[j], acc[i][4 * g_ + j], 0, 0, 0); \
    }                                                                                                \
  } while (0)
    P_MAINLOOP(P_COMPUTE)
#undef P_COMPUTE
    // transposed copy: each wave stages 16 columns x 64 rows (bf16 hi and lo) at
    // a time at the end of the ring, after the Hadamard block
    constexpr int TW_LD = 72;                  // u16 per staged column (WROWS <= 64 rows + pad, 16-byte rows)
    constexpr int TW = 2 * 16 * TW_LD;         // u16 per wave
    constexpr int TW_OFF = NST * STAGE - NW * TW;
    const int had_cap = (TR ? TW_OFF : NST * STAGE) / 2;
    const HadTile ht = had_on ? had_stage<P_BM, BN, NW * 64>(p, m0, n0, ring, had_cap)
                              : HadTile{nullptr, 0, 0, 0};
    // column parts once; row parts per tile row (hoisting all of them would hold
    // 4 TM 64-bit offsets beside the accumulators)
    long coff[TN];
    int hcol[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const long col = n0 + wn * WCOLS + j * 16 + (lane & 15);
      coff[j] = gemm_col_part(p, col);
      hcol[j] = had_on ? had_rkey(p, col) - ht.b0 : 0;
    }
#define P_EPI(STMT, ...)                                                       \
  _Pragma("unroll") for (int i = 0; i < TM; ++i) {                             \
    long roff_[4], hrow_[4];                                                   \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) {                            \
      const long row = m0 + wm * WROWS + i * 16 + 4 * (lane >> 4) + r;         \
      roff_[r] = cbase + gemm_row_part(p, row);                                \
      hrow_[r] = had_on ? (long)(had_rkey(p, row) - ht.a0) * ht.ld : 0;       \
    }                                                                          \
    _Pragma("unroll") for (int j = 0; j < TN; ++j) {                           \
      const long col = n0 + wn * WCOLS + j * 16 + (lane & 15);                 \
      const long row0 = m0 + wm * WROWS + i * 16 + 4 * (lane >> 4);           \
      float hv_[4];                                                            \
      if (!had_on) {                                                           \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) hv_[r] = 1.0f;           \
      } else if (ht.s) {                                                       \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) hv_[r] = ht.s[hrow_[r] + hcol[j]]; \
      } else {                                                                 \
        _Pragma("unroll") for (int r = 0; r < 4; ++r) hv_[r] = (float)p.had[hrow_[r] + hcol[j]]; \
      }                                                                        \
      float vv_[4];                                                            \
      _Pragma("unroll") for (int r = 0; r < 4; ++r) {                          \
        const long row = row0 + r;                                             \
        vv_[r] = alpha * hv_[r] * acc[i][j][r];                                \
        if (row < M && col < N) {                                              \
          const long off = roff_[r] + coff[j];                                 \
          const float val = vv_[r];                                            \
          STMT;                                                                \
        }                                                                      \
      }                                                                        \
      __VA_ARGS__;                                                             \
    }                                                                          \
  }
    if (TR) {
    } else if (p.out_type == GEMM_OUT_F64) {
      double* Cd = (double*)p.C;
      if (p.beta != 0.0) {
        const double beta = p.beta;
        P_EPI(Cd[off] = (double)val + beta * Cd[off], );
      } else {
        P_EPI(Cd[off] = (double)val, );
      }
    } else if (p.out_type == GEMM_OUT_F32) {
      float* Cf = (float*)p.C;
      const float beta = (float)p.beta;
      P_EPI(Cf[off] = beta != 0.0f ? val + beta * Cf[off] : val, );
    } else {
      u16* Ch = (u16*)p.C;
      u16* Cl = (u16*)p.C2;
      P_EPI(const u16 h = bf16_bits_rne(val); Ch[off] = h; Cl[off] = bf16_bits_rne(val - bf16_bits_to_float(h)), );
    }
    if constexpr (TR) {
      u16* Ch = (u16*)p.C;
      u16* Cl = (u16*)p.C2;
      u16* Th = (u16*)p.Ct;
      u16* Tl = (u16*)p.Ct2;
      const long tb = b * p.sCt;
      {
        // With the transposed copy the loops run column block outer, so that each
        // accumulator dies once stored.  The wave's WROWS rows of 16 columns go
        // column-major into its LDS buffer (a lane's four consecutive rows as one
        // 8-byte write) and are read back as 16-byte row segments of the copy: 8
        // lanes fill 128 bytes of one transposed row.  The buffer is private to
        // the wave (no barrier).
        u16* tw = ring + TW_OFF + wave * TW;
        const int q = lane >> 4, c = lane & 15;
        // row parts of the output offset and of the Hadamard index, once per tile
        long ro[TM][4], hr[TM][4];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const long row = m0 + wm * WROWS + i * 16 + 4 * q + r;
            ro[i][r] = cbase + gemm_row_part(p, row);
            hr[i][r] = had_on ? (long)(had_rkey(p, row) - ht.a0) * ht.ld : 0;
          }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const long col = n0 + wn * WCOLS + j * 16 + c;
#pragma unroll
          for (int i = 0; i < TM; ++i) {
            unsigned h[4], l[4];
            float hv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
              hv[r] = !had_on ? 1.0f : (ht.s ? ht.s[hr[i][r] + hcol[j]] : (float)p.had[hr[i][r] + hcol[j]]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const long row = m0 + wm * WROWS + i * 16 + 4 * q + r;
              const float val = alpha * hv[r] * acc[i][j][r];
              h[r] = bf16_bits_rne(val);
              l[r] = bf16_bits_rne(val - bf16_bits_to_float((u16)h[r]));
              if (row < M && col < N) {
                const long off = ro[i][r] + coff[j];
                Ch[off] = (u16)h[r];
                Cl[off] = (u16)l[r];
              }
            }
            const int o = c * TW_LD + i * 16 + 4 * q;
            *(uint2*)(tw + o) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
            *(uint2*)(tw + 16 * TW_LD + o) = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
          }
          asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
          // 16 columns x WROWS / 8 segments of 8 rows, 64 per pass
          constexpr int SEG = WROWS / 8;
#pragma unroll
          for (int t = 0; t < WROWS / 32; ++t) {
            const int e = lane + 64 * t, cc = e / SEG, rr = (e % SEG) * 8;
            const uint4 H = *(const uint4*)(tw + cc * TW_LD + rr);
            const uint4 L = *(const uint4*)(tw + 16 * TW_LD + cc * TW_LD + rr);
            const long tcol = n0 + wn * WCOLS + j * 16 + cc, trow = m0 + wm * WROWS + rr;
            if (tcol < N && trow < M) {
              *(uint4*)(Th + tb + tcol * p.ldct + trow) = H;
              *(uint4*)(Tl + tb + tcol * p.ldct + trow) = L;
            }
          }
          asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        }
      }
    }
#undef P_EPI
  }
#undef P_MAINLOOP
#undef P_SLOT
#undef P_LD
#undef P_ISSUE
#undef P_ISSUE_A
#undef P_ISSUE_B
}

// wave count of the pipelined kernel (8: two waves per SIMD hide each other's stalls)
#ifndef DFTB_BF3_WAVES
#define DFTB_BF3_WAVES 8
#endif

int gemm_bf16x3_launch(const GemmArgs& p, hipStream_t st) {
  if (p.a_trans || !p.b_trans) return DFTB_ERR_ARG;
  if (!p.A2 || !p.B2) return DFTB_ERR_ARG;
  if ((p.lda & 7) || (p.ldb & 7) || (p.K & 7) || (p.sA & 7) || (p.sB & 7)) return DFTB_ERR_ARG;
  if (p.ksplit > 1 && (p.kchunk & 7)) return DFTB_ERR_ARG;
  if ((((uintptr_t)p.A) | ((uintptr_t)p.A2) | ((uintptr_t)p.B) | ((uintptr_t)p.B2)) & 15) return DFTB_ERR_ARG;
  const bool small = p.tile == 64 || (p.tile == 0 && (p.M < 128 || p.N < 128));
  bool pipe = p.tile == 256 || p.tile == 512 || (p.tile == 0 && p.M >= 256 && p.N >= 128 && p.K >= 4 * P_BK);
  if (p.Ct) {
    // transposed copy: pipelined kernel, 16x16 form, bf16 pairs stored as 16-byte
    // row segments (8 rows), so M, ldct, sCt multiples of 8 and 16-byte aligned arrays
    if (p.tile == 64 || p.tile == 128 || p.out_type != GEMM_OUT_BF16PAIR || !p.Ct2 || p.ksplit != 1) return DFTB_ERR_ARG;
    if ((p.M & 7) || (p.ldct & 7) || (p.sCt & 7) || p.ldct < p.M) return DFTB_ERR_ARG;
    if ((((uintptr_t)p.Ct) | ((uintptr_t)p.Ct2)) & 15) return DFTB_ERR_ARG;
    pipe = true;
  }
  if (pipe) {
    // tile 256: 256 x 128 tiles; tile 512: 256 x 256 tiles; automatic: 256 x 256
    // when N >= 256 unless DFTB_BF3_BN=128 asks for the narrower tile
    static const int bn_env = getenv("DFTB_BF3_BN") ? atoi(getenv("DFTB_BF3_BN")) : 256;
    const int bn = p.tile == 512 ? 256 : p.tile == 256 ? 128 : (bn_env == 128 || p.N < 256) ? 128 : 256;
    // 32-bit buffer offsets: each workgroup's operand span must stay below 4 GiB
    const long span_a = ((long)P_BM * p.lda + p.K) * 2, span_b = ((long)bn * p.ldb + p.K) * 2;
    if (span_a >= 0xffff0000L || span_b >= 0xffff0000L) return DFTB_ERR_ARG;
    dim3 grid(ceil_div(p.N, bn), ceil_div(p.M, P_BM), p.batch * p.ksplit);
    GemmTimer tm;
    tm.begin(st);
    // 16x16x32 by default: the kernel is bound by its global->LDS feed, where the
    // 16x16 form measured 2-4% faster (256 x 128 tiles); DFTB_BF3_M32=1 selects the
    // 32x32x16 form
    static const int m32_env = getenv("DFTB_BF3_M32") ? atoi(getenv("DFTB_BF3_M32")) : 0;
    const int m32 = m32_env;
    const dim3 block(DFTB_BF3_WAVES * 64);
    if (p.Ct) {
      if (bn == 256)
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, false, 256, true>), grid, block, 0, st, p);
      else
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, false, 128, true>), grid, block, 0, st, p);
    } else if (bn == 256) {
      if (m32)
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, true, 256>), grid, block, 0, st, p);
      else
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, false, 256>), grid, block, 0, st, p);
    } else {
      if (m32)
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, true, 128>), grid, block, 0, st, p);
      else
        hipLaunchKernelGGL((gemm_bf3_pipe_kernel<DFTB_BF3_WAVES, false, 128>), grid, block, 0, st, p);
    }
    tm.end(st, GT_BF3_PIPE, 3.0 * 2.0 * p.M * p.N * (double)p.K * p.batch);
    DFTB_CHECK_LAUNCH();
    return DFTB_OK;
  }
  GemmTimer tm;
  tm.begin(st);
  if (small) {
    dim3 grid(ceil_div(p.N, 64), ceil_div(p.M, 64), p.batch * p.ksplit);
    hipLaunchKernelGGL((gemm_bf3_kernel<64, 64>), grid, dim3(256), 0, st, p);
  } else {
    dim3 grid(ceil_div(p.N, 128), ceil_div(p.M, 128), p.batch * p.ksplit);
    hipLaunchKernelGGL((gemm_bf3_kernel<128, 128>), grid, dim3(256), 0, st, p);
  }
  tm.end(st, GT_BF3_TILE, 3.0 * 2.0 * p.M * p.N * (double)p.K * p.batch);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
