// (a) Mulliken transition charges, gamma fold and operand conversions.
//
// q^{pq}_A = 1/2 sum_{mu in A} [X_{mu p} Y_{mu q} + X_{mu q} Y_{mu p}],  X = C, Y = S C
// (DFTB/extensions/tddftb.f90:50-110).  S C is one MFMA GEMM; the charge kernel
// is bound by its HBM writes (Nat * P * Q outputs, each 2*nmu FMAs); its reads
// (X, Y rows of one atom per workgroup) are small next to them.
#include <mutex>
#include <vector>
#include "common.h"
#include "gemm.h"

#define TC_MAXMU 16
#define TC_LDS_BYTES (48 * 1024)

// One workgroup owns one atom A, a chunk of p and S stripes of 256 q (one q per
// thread and stripe).  The atom's p-rows X_{mu p}, Y_{mu p} of the chunk are
// staged in LDS once, as (x, y) pairs per p with zeros for mu >= nmu; each
// thread keeps X_{mu q}, Y_{mu q} of its S q in registers (read once).  The
// loop runs over p and, inside, over the stripes, so a workgroup completes
// whole output rows (S * 256 >= the row length, or S stripes of it) at a time:
// partial 128-byte lines at stripe boundaries are merged in L2 before they
// reach HBM, which matters when rows are not line-aligned (the reference layout
// with an odd-sized orbital block).  LDS reads are wave-uniform broadcasts and
// the loop has no barrier.  Rows p in [P, Ppad) and columns q in [Q, Qpad) are
// written as zeros (the engine layouts' padding).  Dynamic LDS: 2 * NMU * pchunk
// doubles.
template <int NMU, int S>
__global__ __launch_bounds__(256) void tc_kernel(const double* __restrict__ Xp, const double* __restrict__ Yp, int P,
                                                 const double* __restrict__ Xq, const double* __restrict__ Yq, int Q,
                                                 const int* __restrict__ atom_off, int Ppad, int Qpad, int pchunk,
                                                 double* __restrict__ out, long sP, long sA) {
  extern __shared__ double sxy[];        // [pchunk][NMU][2]
  const int A = blockIdx.x;
  const int pbeg = blockIdx.y * pchunk;
  const int pend = min(Ppad, pbeg + pchunk);
  const int qb = blockIdx.z * S * 256 + threadIdx.x;
  const int mu0 = atom_off[A];
  const int nmu = atom_off[A + 1] - mu0;
  // staging: an index outside the atom's rows (or beyond P) is clamped to
  // element 0 and its value masked, so the loads carry no branch
  for (int t = threadIdx.x; t < (pend - pbeg) * NMU; t += blockDim.x) {
    const int pp = t / NMU, mu = t % NMU, p = pbeg + pp;
    const bool ok = (mu < nmu) && (p < P);
    const long idx = ok ? (long)(mu0 + mu) * P + p : 0;
    const double x = Xp[idx], y = Yp[idx];
    sxy[2 * t] = ok ? x : 0.0;
    sxy[2 * t + 1] = ok ? y : 0.0;
  }
  __syncthreads();
  double xq[S][NMU], yq[S][NMU];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int q = qb + 256 * j;
#pragma unroll
    for (int mu = 0; mu < NMU; ++mu) {
      const bool ok = (q < Q) && (mu < nmu);
      const long idx = ok ? (long)(mu0 + mu) * Q + q : 0;
      const double x = Xq[idx], y = Yq[idx];
      xq[j][mu] = ok ? x : 0.0;
      yq[j][mu] = ok ? y : 0.0;
    }
  }
  double* o = out + (long)A * sA + (long)pbeg * sP + qb;
  const double* r = sxy;
  for (int p = pbeg; p < pend; ++p, r += 2 * NMU, o += sP) {
    double rv[2 * NMU];
#pragma unroll
    for (int i = 0; i < 2 * NMU; ++i) rv[i] = r[i];
#pragma unroll
    for (int j = 0; j < S; ++j) {
      double s = 0.0;
#pragma unroll
      for (int mu = 0; mu < NMU; ++mu) s += rv[2 * mu] * yq[j][mu] + xq[j][mu] * rv[2 * mu + 1];
      if (qb + 256 * j < Qpad) o[256 * j] = 0.5 * s;   // zero for padded p (staged zeros) and q (zero registers)
    }
  }
}

template <int NMU, int S>
static void tc_launch(dim3 grid, size_t lds, hipStream_t st, const double* Xp, const double* Yp, int P,
                      const double* Xq, const double* Yq, int Q, const int* off, int Ppad, int Qpad, int pchunk,
                      double* out, long sP, long sA) {
  hipLaunchKernelGGL((tc_kernel<NMU, S>), grid, dim3(256), lds, st, Xp, Yp, P, Xq, Yq, Q, off, Ppad, Qpad, pchunk,
                     out, sP, sA);
}

// stripes per workgroup held in registers: up to 8 (NMU 4), 4 (NMU 9), 2 (NMU 16)
template <int NMU>
static void tc_dispatch(int S, dim3 grid, size_t lds, hipStream_t st, const double* Xp, const double* Yp, int P,
                        const double* Xq, const double* Yq, int Q, const int* off, int Ppad, int Qpad, int pchunk,
                        double* out, long sP, long sA) {
#define TC_CASE(n) \
  case n: tc_launch<NMU, (NMU * n <= 32 ? n : 1)>(grid, lds, st, Xp, Yp, P, Xq, Yq, Q, off, Ppad, Qpad, pchunk, out, sP, sA); break;
  switch (S) {
    TC_CASE(1) TC_CASE(2) TC_CASE(3) TC_CASE(4) TC_CASE(5) TC_CASE(6) TC_CASE(7) TC_CASE(8)
    default: break;
  }
#undef TC_CASE
}

struct TcPlan {
  int nat, norb, nocc, nvirt, maxmu;
  const double *Co, *Cv;
  double *Yo, *Yv;
  int* off;
};

static int tc_block(const TcPlan& t, bool p_occ, bool q_occ, int Ppad, int Qpad, double* out, long sP, long sA,
                    hipStream_t st) {
  if (!out) return DFTB_OK;
  const double* Xp = p_occ ? t.Co : t.Cv;
  const double* Yp = p_occ ? t.Yo : t.Yv;
  const double* Xq = q_occ ? t.Co : t.Cv;
  const double* Yq = q_occ ? t.Yo : t.Yv;
  int P = p_occ ? t.nocc : t.nvirt;
  int Q = q_occ ? t.nocc : t.nvirt;
  if (P == 0 || Ppad == 0 || Qpad == 0) return DFTB_OK;
  const int nmu = t.maxmu <= 4 ? 4 : (t.maxmu <= 9 ? 9 : 16);
  const int smax = 32 / nmu;                       // stripes of 256 q per workgroup (registers)
  const int stripes = ceil_div(Qpad, 256);
  const int S = stripes < smax ? stripes : smax;
  const int gz = ceil_div(stripes, S);
  // p chunks: output rows that are not 128-byte aligned (the reference layout
  // with nocc or nvirt not a multiple of 16) leave partial lines at row ends,
  // merged best by short chunks (16 rows: 4.8 TB/s against 3.5 TB/s with 624 at
  // the flake); aligned rows stream best in long chunks (>= ~1024 workgroups)
  const int pmax = (int)(TC_LDS_BYTES / (2 * nmu * sizeof(double)));
  int pchunk;
  if ((sP % 16) != 0 || (sA % 16) != 0) {
    pchunk = 16;
  } else {
    const long wg = (long)t.nat * gz;
    int psplit = (int)((1024 + wg - 1) / wg);
    if (psplit < 1) psplit = 1;
    pchunk = ceil_div(Ppad, psplit);
  }
  if (pchunk > pmax) pchunk = pmax;
  if (pchunk < 1) pchunk = 1;
  const size_t lds = (size_t)2 * nmu * pchunk * sizeof(double);
  dim3 grid(t.nat, ceil_div(Ppad, pchunk), gz);
  if (nmu == 4)
    tc_dispatch<4>(S, grid, lds, st, Xp, Yp, P, Xq, Yq, Q, t.off, Ppad, Qpad, pchunk, out, sP, sA);
  else if (nmu == 9)
    tc_dispatch<9>(S, grid, lds, st, Xp, Yp, P, Xq, Yq, Q, t.off, Ppad, Qpad, pchunk, out, sP, sA);
  else
    tc_dispatch<16>(S, grid, lds, st, Xp, Yp, P, Xq, Yq, Q, t.off, Ppad, Qpad, pchunk, out, sP, sA);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" size_t dftb_trans_charges_workspace(int norb, int nocc, int nvirt, int nat) {
  return align256((size_t)norb * (nocc + nvirt) * sizeof(double)) + align256((size_t)(nat + 1) * sizeof(int));
}

// Prepare S C products and atom offsets (orbs_per_atom is a host array, as the
// f2py argument it replaces).
static int tc_prepare(TcPlan& t, const double* Co, const double* Cv, const double* S, const int* orbs_per_atom,
                      int nat, int norb, int nocc, int nvirt, void* ws, size_t ws_bytes, hipStream_t st,
                      bool timed = true) {
  if (nat <= 0 || norb <= 0 || nocc < 0 || nvirt < 0 || !S || !orbs_per_atom || !ws) return DFTB_ERR_ARG;
  if (ws_bytes < dftb_trans_charges_workspace(norb, nocc, nvirt, nat)) return DFTB_ERR_WORKSPACE;
  std::vector<int> off(nat + 1, 0);
  int maxmu = 0;
  for (int A = 0; A < nat; ++A) {
    if (orbs_per_atom[A] < 0 || orbs_per_atom[A] > TC_MAXMU) return DFTB_ERR_ARG;
    off[A + 1] = off[A] + orbs_per_atom[A];
    if (orbs_per_atom[A] > maxmu) maxmu = orbs_per_atom[A];
  }
  if (off[nat] != norb) return DFTB_ERR_ARG;
  char* w = (char*)ws;
  t.nat = nat; t.norb = norb; t.nocc = nocc; t.nvirt = nvirt; t.maxmu = maxmu;
  t.Co = Co; t.Cv = Cv;
  t.Yo = (double*)w;
  t.Yv = t.Yo + (size_t)norb * nocc;
  t.off = (int*)(w + align256((size_t)norb * (nocc + nvirt) * sizeof(double)));
  if (hipMemcpyAsync(t.off, off.data(), (nat + 1) * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
    return DFTB_ERR_LAUNCH;
  // Y = S C  (S symmetric, row-major [norb][norb]; C row-major [norb][n])
  for (int pass = 0; pass < 2; ++pass) {
    int n = pass == 0 ? nocc : nvirt;
    if (n == 0) continue;
    GemmArgs g = gemm_args_default();
    g.A = S; g.lda = norb;
    g.B = pass == 0 ? Co : Cv; g.ldb = n;
    g.C = pass == 0 ? t.Yo : t.Yv; g.ldc = n;
    g.M = norb; g.N = n; g.K = norb;
    int rc = gemm_launch(g, st, timed);
    if (rc) return rc;
  }
  return DFTB_OK;
}

static int trans_charges_ref_layout(const double* orbs_occ, const double* orbs_virt, const double* S,
                                    const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt, double* q_oo,
                                    double* q_ov, double* q_vv, void* workspace, size_t ws_bytes, hipStream_t st,
                                    bool timed) {
  TcPlan t;
  int rc = tc_prepare(t, orbs_occ, orbs_virt, S, orbs_per_atom, nat, norb, nocc, nvirt, workspace, ws_bytes, st,
                      timed);
  if (rc) return rc;
  long o = nocc, v = nvirt;
  if ((rc = tc_block(t, true, true, nocc, nocc, q_oo, o, o * o, st))) return rc;
  if ((rc = tc_block(t, true, false, nocc, nvirt, q_ov, v, o * v, st))) return rc;
  if ((rc = tc_block(t, false, false, nvirt, nvirt, q_vv, v, v * v, st))) return rc;
  return DFTB_OK;
}

extern "C" int dftb_trans_charges(const double* orbs_occ, const double* orbs_virt, const double* S,
                                  const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt, double* q_oo,
                                  double* q_ov, double* q_vv, void* workspace, size_t ws_bytes, void* stream) {
  return trans_charges_ref_layout(orbs_occ, orbs_virt, S, orbs_per_atom, nat, norb, nocc, nvirt, q_oo, q_ov, q_vv,
                                  workspace, ws_bytes, (hipStream_t)stream, true);
}

// ---- allocation-free variant -------------------------------------------------
// dftb_trans_charges_noalloc keeps its scratch (the S C products, norb * (nocc +
// nvirt) doubles, and the atom offsets, nat + 1 ints) in this fixed device array
// instead of a caller workspace.  Each device holds its own copy for the life of
// the process.  Host code never passes `scratch` itself to a kernel or copy (on
// the host the name is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns, looked up once per device on that device's first
// call.  Every output element is overwritten, and the scratch contents of an
// earlier call are never read, so nothing needs clearing between calls.  Calls
// that overlap on one device (two host threads, or two streams without ordering
// between them) share the array and race.
#define TC_SCRATCH_BYTES ((size_t)64 << 20)
#define TC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[TC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[TC_MAX_DEVICES];

static int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= TC_MAX_DEVICES) {
    fprintf(stderr, "dftbaby_amd: dftb_trans_charges_noalloc: no usable current device (%d)\n", dev);
    return DFTB_ERR_LAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "dftbaby_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      return DFTB_ERR_LAUNCH;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DFTB_OK;
}

extern "C" int dftb_trans_charges_noalloc(const double* orbs_occ, const double* orbs_virt, const double* S,
                                          const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt,
                                          double* q_oo, double* q_ov, double* q_vv, void* stream) {
  if (nat <= 0 || norb <= 0 || nocc < 0 || nvirt < 0) return DFTB_ERR_ARG;
  const size_t need = dftb_trans_charges_workspace(norb, nocc, nvirt, nat);
  if (need > TC_SCRATCH_BYTES) {
    fprintf(stderr,
            "dftbaby_amd: dftb_trans_charges_noalloc: norb = %d, nocc + nvirt = %d, nat = %d need %zu bytes of "
            "scratch, above the fixed %zu (norb * (nocc + nvirt) <= %zu doubles); use dftb_trans_charges with a "
            "workspace\n",
            norb, nocc + nvirt, nat, need, TC_SCRATCH_BYTES, (TC_SCRATCH_BYTES - 65536) / sizeof(double));
    return DFTB_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  return trans_charges_ref_layout(orbs_occ, orbs_virt, S, orbs_per_atom, nat, norb, nocc, nvirt, q_oo, q_ov, q_vv,
                                  ws, TC_SCRATCH_BYTES, (hipStream_t)stream, false);
}

static inline int pad8(int n) { return (n + 7) & ~7; }

extern "C" int dftb_trans_charges_engine(const double* orbs_occ, const double* orbs_virt, const double* S,
                                         const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt,
                                         double* q_oo_e, double* q_ov_e, double* q_vv_e, double* q_c,
                                         void* workspace, size_t ws_bytes, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  TcPlan t;
  int rc = tc_prepare(t, orbs_occ, orbs_virt, S, orbs_per_atom, nat, norb, nocc, nvirt, workspace, ws_bytes, st);
  if (rc) return rc;
  long opad = pad8(nocc), vpad = pad8(nvirt), na = nat;
  if ((rc = tc_block(t, true, true, nocc, (int)opad, q_oo_e, na * opad, opad, st))) return rc;
  if ((rc = tc_block(t, true, false, nocc, (int)vpad, q_ov_e, na * vpad, vpad, st))) return rc;
  if ((rc = tc_block(t, false, false, (int)vpad, (int)vpad, q_vv_e, na * vpad, vpad, st))) return rc;
  if ((rc = tc_block(t, true, false, (int)opad, (int)vpad, q_c, vpad, opad * vpad, st))) return rc;
  return DFTB_OK;
}

extern "C" int dftb_fold_gamma(const double* gamma, int nat, const double* q, long np, long qpad, double* qf,
                               void* stream) {
  if (!gamma || !q || !qf || nat <= 0) return DFTB_ERR_ARG;
  GemmArgs g = gemm_args_default();
  g.A = gamma; g.lda = nat; g.sA = 0;
  g.B = q; g.ldb = qpad; g.sB = (long)nat * qpad;
  g.C = qf; g.ldc = qpad; g.sC = (long)nat * qpad;
  g.M = nat; g.N = (int)qpad; g.K = nat; g.batch = (int)np;
  return gemm_launch(g, (hipStream_t)stream);
}

__global__ void cvt_f32_kernel(const double* x, long n, float* y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}

__global__ void cvt_bf16pair_kernel(const double* x, long n, unsigned short* hi, unsigned short* lo) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float f = (float)x[i];
    unsigned short h = bf16_bits_rne(f);
    hi[i] = h;
    lo[i] = bf16_bits_rne(f - bf16_bits_to_float(h));
  }
}

static dim3 stream_grid(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

extern "C" int dftb_convert_f32(const double* x, long n, float* y, void* stream) {
  if (n <= 0) return DFTB_OK;
  hipLaunchKernelGGL(cvt_f32_kernel, stream_grid(n), dim3(256), 0, (hipStream_t)stream, x, n, y);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}

extern "C" int dftb_convert_bf16pair(const double* x, long n, unsigned short* hi, unsigned short* lo, void* stream) {
  if (n <= 0) return DFTB_OK;
  hipLaunchKernelGGL(cvt_bf16pair_kernel, stream_grid(n), dim3(256), 0, (hipStream_t)stream, x, n, hi, lo);
  DFTB_CHECK_LAUNCH();
  return DFTB_OK;
}
