#pragma once
#include "common.h"
#include "../../include/dftbaby_amd.h"

// timed = false keeps the launch out of the optional launch-timing record (which
// creates HIP events): the allocation-free entry points use it.
int gemm_launch(const GemmArgs& p, hipStream_t st, bool timed = true);

// Optional launch timing (see dftb_gemm_timing_*): kernel classes
enum { GT_F64_128 = 0, GT_F64_64 = 1, GT_F32_128 = 2, GT_F32_64 = 3, GT_BF3_PIPE = 4, GT_BF3_TILE = 5,
       GT_F64_128_T = 6, GT_NCLASS = 7 };
struct GemmTimer {
  int slot;
  void begin(hipStream_t st);
  void end(hipStream_t st, int cls, double flops);
};
int gemm_bf16x3_launch(const GemmArgs& p, hipStream_t st);
int gemm_reduce_launch(const double* P, long sP, int nsplit, double* C, long n, double alpha, double beta,
                       hipStream_t st);

// Output offset of element (row, col) within one batch entry = row part + column
// part (row remap: (row / rdiv) rstr + (row % rdiv) rs2, else row ldc; column
// remap likewise with cdiv, cstr, cs2).
__device__ __forceinline__ long gemm_row_part(const GemmArgs& p, long row) {
  return (p.rdiv > 0) ? (row / p.rdiv) * p.rstr + (row % p.rdiv) * p.rs2 : row * p.ldc;
}
__device__ __forceinline__ long gemm_col_part(const GemmArgs& p, long col) {
  return (p.cdiv > 0) ? (col / p.cdiv) * p.cstr + (col % p.cdiv) * p.cs2 : col;
}
// Row and column parts of the Hadamard index (had[row part + column part]); any
// row / column gives an index inside the factor matrix.  Only for p.had != null.
__device__ __forceinline__ long gemm_had_row(const GemmArgs& p, long row) {
  const unsigned rm = (unsigned)row % (unsigned)p.hmod;
  return (long)(p.hmap ? p.hmap[rm] : (int)rm) * p.ldh;
}
__device__ __forceinline__ long gemm_had_col(const GemmArgs& p, long col) {
  const unsigned cm = (unsigned)col % (unsigned)p.hmod;
  return p.hmap ? p.hmap[cm] : (long)cm;
}


static inline GemmArgs gemm_args_default() {
  GemmArgs p;
  p.A = p.A2 = p.B = p.B2 = nullptr;
  p.C = p.C2 = nullptr;
  p.lda = p.ldb = p.ldc = 0;
  p.sA = p.sB = p.sC = 0;
  p.cdiv = 0;
  p.cstr = 0;
  p.cs2 = 1;
  p.rdiv = 0;
  p.rstr = 0;
  p.rs2 = 0;
  p.had = nullptr;
  p.hmod = 1;
  p.ldh = 0;
  p.hmap = nullptr;
  p.M = p.N = p.K = 0;
  p.batch = 1;
  p.ksplit = 1;
  p.kchunk = 0;
  p.sCk = 0;
  p.alpha = 1.0;
  p.beta = 0.0;
  p.a_trans = p.b_trans = 0;
  p.in_type = GEMM_IN_F64;
  p.out_type = GEMM_OUT_F64;
  p.tile = 0;
  p.Ct = p.Ct2 = nullptr;
  p.ldct = p.sCt = 0;
  return p;
}
