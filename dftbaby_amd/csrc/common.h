// Shared definitions for the gfx950 LR-TD-DFTB kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

// Status codes returned by every C-ABI entry point.
enum {
  DFTB_OK = 0,
  DFTB_ERR_ARG = 1,        // invalid argument (shape, pointer, mode)
  DFTB_ERR_LAUNCH = 2,     // HIP launch / runtime error
  DFTB_ERR_NOT_SPD = 3,    // subspace matrix (A-B projection) not positive definite
  DFTB_ERR_WORKSPACE = 4,  // workspace too small
};

#define DFTB_CHECK_LAUNCH()                                  \
  do {                                                       \
    hipError_t e_ = hipGetLastError();                       \
    if (e_ != hipSuccess) {                                  \
      fprintf(stderr, "dftbaby_amd: %s\n", hipGetErrorString(e_)); \
      return DFTB_ERR_LAUNCH;                                \
    }                                                        \
  } while (0)

static inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ unsigned short bf16_bits_rne(float x) {
  unsigned int u = __float_as_uint(x);
  unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_float(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8 XCDs
// of the MI355X, each with its own L2.  Map the linear dispatch id to a tile id
// so that every XCD walks a contiguous chunk of the (x fastest, then y, then z)
// tile space and neighbouring tiles, which share operand rows, share an L2.
// The map is a bijection of [0, nx*ny*nz).
__device__ __forceinline__ void xcd_tile(unsigned& tx, unsigned& ty, unsigned& tz) {
  const unsigned nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  const unsigned total = nx * ny * nz;
  const unsigned bid = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned q = total / 8u, r = total % 8u;
  const unsigned xcd = bid % 8u, slot = bid / 8u;
  const unsigned t = (xcd < r) ? xcd * (q + 1u) + slot : r * (q + 1u) + (xcd - r) * q + slot;
  tx = t % nx;
  ty = (t / nx) % ny;
  tz = t / (nx * ny);
}

// As xcd_tile, with the tiles of each z slice visited in groups of G tile rows
// (column by column inside a group), so that the tiles one XCD runs at the same
// time form a G x (CUs / G) block: its operand panels are fetched once per block
// instead of once per tile row.  A bijection of [0, nx*ny*nz).
template <unsigned G>
__device__ __forceinline__ void xcd_tile_grouped(unsigned& tx, unsigned& ty, unsigned& tz) {
  const unsigned nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  const unsigned total = nx * ny * nz;
  const unsigned bid = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned q = total / 8u, r = total % 8u;
  const unsigned xcd = bid % 8u, slot = bid / 8u;
  const unsigned t = (xcd < r) ? xcd * (q + 1u) + slot : r * (q + 1u) + (xcd - r) * q + slot;
  const unsigned plane = nx * ny;
  tz = t / plane;
  const unsigned u = t - tz * plane;
  const unsigned g = u / (G * nx);
  const unsigned y0 = g * G;
  const unsigned gs = (ny - y0 < G) ? ny - y0 : G;
  const unsigned w = u - g * G * nx;
  ty = y0 + w % gs;
  tx = w / gs;
}
