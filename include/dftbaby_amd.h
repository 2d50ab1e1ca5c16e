/*
 * dftbaby_amd: C ABI of the MI355X (gfx950) LR-TD-DFTB excited-state engine.
 *
 * Every entry point takes device pointers, plain sizes and a hipStream_t
 * (passed as void*), returns an int status (0 = ok, see DFTB_ERR_* in
 * csrc/common.h: 1 invalid argument, 2 launch error, 3 subspace matrix not
 * positive definite, 4 workspace too small) and never allocates device memory:
 * the caller owns all buffers, including the workspace.  No exception crosses
 * the ABI.  Calls on one stream are ordered; the library keeps no state.
 *
 * Vector layout: an excitation vector is a dense [opad][vpad] FP64 block
 * (row i = occupied orbital, column a = virtual orbital, idx = i*vpad + a,
 * the reference's i-major flattening of Solver.py:242-247) whose padding
 * (i >= nocc or a >= nvirt) is zero.  opad and vpad are nocc and nvirt rounded
 * up to a multiple of 8.
 */
#ifndef DFTBABY_AMD_H
#define DFTBABY_AMD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generic batched GEMM ------------------------------------------------
 * C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b] on MFMA.  Element (m,k) of A
 * is A[b*sA + (a_trans ? k*lda + m : m*lda + k)], element (k,n) of B is
 * B[b*sB + (b_trans ? n*ldb + k : k*ldb + n)].  Output element (m, n) goes to
 * C[b*sC + row(m) + col(n)] with row(m) = m * ldc, or (m / rdiv) * rstr +
 * (m % rdiv) * rs2 when rdiv > 0, and col(n) = n, or (n / cdiv) * cstr +
 * (n % cdiv) * cs2 when cdiv > 0.  A non-NULL had scales each result by
 * had[(m % hmod) * ldh + n % hmod] (with hmap: had[hmap[m % hmod] * ldh +
 * hmap[n % hmod]]) before it is stored.  ksplit > 1 splits K
 * into chunks of kchunk and writes slice s at C + s*sCk (beta must be 0).
 * in_type GEMM_IN_BF16X3 takes (hi, lo) bf16 operand pairs in (A, A2) and
 * (B, B2) and forms hi*hi + hi*lo + lo*hi with FP32 accumulation; it needs
 * a_trans = 0, b_trans = 1 (both operands K-contiguous), K, lda, ldb and the
 * operand offsets multiples of 8 elements.
 * No reference counterpart: it carries the np.tensordot / MATMUL contractions
 * of Solver.py:583-635, 759-807 and tddftb.f90:114-287. */
enum { GEMM_IN_F64 = 0, GEMM_IN_F32 = 1, GEMM_IN_BF16X3 = 2 };
enum { GEMM_OUT_F64 = 0, GEMM_OUT_F32 = 1, GEMM_OUT_BF16PAIR = 2 };

typedef struct GemmArgs {
  const void* A; const void* A2;
  const void* B; const void* B2;
  void* C; void* C2;
  long lda, ldb, ldc;
  long sA, sB, sC;
  long cdiv, cstr, cs2;
  long rdiv, rstr, rs2;
  const double* had;   /* optional Hadamard factor: element (m, n) is scaled by had[(m % hmod) * ldh + n % hmod], */
  long hmod, ldh;      /* or, with hmap, by had[hmap[m % hmod] * ldh + hmap[n % hmod]] */
  const int* hmap;
  int M, N, K;
  int batch;
  int ksplit, kchunk;
  long sCk;
  double alpha, beta;
  int a_trans, b_trans;
  int in_type, out_type;
  int tile; /* 0 auto, 64 or 128; BF16x3 only: 256 = pipelined 256x128, 512 = pipelined 256x256 */
  void* Ct; void* Ct2; /* optional transposed copy (pipelined BF16x3, bf16-pair output, M, ldct, sCt */
  long ldct, sCt;      /* multiples of 8): element (m, n) of batch b also goes to Ct[b * sCt + n * ldct + m] */
} GemmArgs;

int dftb_gemm(const GemmArgs* p, void* stream);

/* ---- launch timing of the GEMM kernels (measurement only) -----------------
 * No reference counterpart.  dftb_gemm_timing_enable(1) clears the record and
 * brackets every following GEMM launch with two HIP events on its stream;
 * dftb_gemm_timing_read(cls, ...) waits for the recorded launches of kernel
 * class cls and returns their summed device time (ms), count and algorithmic
 * flops (2 M N K per batch entry, times 3 for the BF16x3 split product).
 * Classes: 0 FP64 128x128 (no transposed operand), 1 FP64 64x64, 2 FP32
 * 128x128, 3 FP32 64x64, 4 BF16x3 pipelined (256x128 or 256x256), 5 BF16x3 64/128 tiles,
 * 6 FP64 128x128 with a transposed operand. */
enum { GEMM_CLASS_F64_128 = 0, GEMM_CLASS_F64_64 = 1, GEMM_CLASS_F32_128 = 2, GEMM_CLASS_F32_64 = 3,
       GEMM_CLASS_BF16X3_PIPE = 4, GEMM_CLASS_BF16X3_TILE = 5, GEMM_CLASS_F64_128_T = 6 };
int dftb_gemm_timing_enable(int on);
int dftb_gemm_timing_read(int cls, double* ms_total, long* launches, double* flops);

/* ---- (a) Mulliken transition charges --------------------------------------
 * Replaces f2py tddftb.trans_charges(orbs_occ, orbs_virt, S, orbsPerAtom)
 * -> (q_oo, q_ov, q_vv)   (DFTB/extensions/tddftb.f90:12-112, called at
 * DFTB/LR_TDDFTB.py:303-306):
 *   q^{pq}_A = 1/2 sum_{mu in A} [C_{mu p} (SC)_{mu q} + C_{mu q} (SC)_{mu p}].
 * orbs_occ [norb][nocc], orbs_virt [norb][nvirt], S [norb][norb] row-major
 * device arrays; orbs_per_atom [nat] (int32) is a HOST array, like the numpy
 * array f2py receives.  Outputs in the reference's logical order
 * q_oo[A][i][j], q_ov[A][i][a], q_vv[A][a][b] (any output may be NULL).
 * workspace >= dftb_trans_charges_workspace(norb, nocc, nvirt, nat) bytes. */
size_t dftb_trans_charges_workspace(int norb, int nocc, int nvirt, int nat);
int dftb_trans_charges(const double* orbs_occ, const double* orbs_virt, const double* S,
                       const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt,
                       double* q_oo, double* q_ov, double* q_vv,
                       void* workspace, size_t ws_bytes, void* stream);

/* Allocation-free variant of dftb_trans_charges: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * S C products and atom offsets live in a fixed 64 MiB device array that the
 * library holds on each device for the life of the process, so the call takes
 * no workspace and calls no hipMalloc/hipFree.  Same arguments, layouts and
 * results as dftb_trans_charges.  Limit: dftb_trans_charges_workspace(norb,
 * nocc, nvirt, nat) <= 64 MiB (about norb * (nocc + nvirt) <= 8.38e6 doubles,
 * e.g. norb = nocc + nvirt <= 2896); larger calls return DFTB_ERR_WORKSPACE (4)
 * before any device work.  Calls on one device share the array: calls that
 * overlap there (other host threads, unordered streams) race. */
int dftb_trans_charges_noalloc(const double* orbs_occ, const double* orbs_virt, const double* S,
                               const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt,
                               double* q_oo, double* q_ov, double* q_vv, void* stream);

/* Same charges written in the engine's padded layouts (NULL = skip):
 *   q_oo_e [nocc][nat][opad], q_ov_e [nocc][nat][vpad], q_vv_e [vpad][nat][vpad],
 *   q_c    [nat][opad][vpad]  (the Coulomb layout).
 * Padding entries are written as zeros. */
int dftb_trans_charges_engine(const double* orbs_occ, const double* orbs_virt, const double* S,
                              const int* orbs_per_atom, int nat, int norb, int nocc, int nvirt,
                              double* q_oo_e, double* q_ov_e, double* q_vv_e, double* q_c,
                              void* workspace, size_t ws_bytes, void* stream);

/* ---- (b) Casida matvecs ---------------------------------------------------
 * The operator of Solver.py:583-635 (_ApBv/_AmBv, Fortran twins apbv/ambv at
 * tddftb.f90:114-287) with gamma_lr folded into q_ov and q_vv once:
 *   (A+B)v = omega*v + 4 q_ov^T gamma q_ov v - K1 v - K2 v
 *   (A-B)v = omega*v                          - K1 v + K2 v
 *   (K1 v)_ia = sum_{A,j} q^{ij}_A sum_b v_jb qf^{ba}_A ,  qf = gamma_lr . q_vv
 *   (K2 v)_ia = sum_{A,b} q^{ib}_A sum_j v_jb qf^{ja}_A ,  qf = gamma_lr . q_ov
 * Both products are formed for the same block of vectors, so the exchange
 * work is shared between them.  precision: 0 FP64, 1 FP32, 2 BF16x3 split
 * pairs (FP32 accumulation); results are always FP64.  Operand pointers are
 * FP64 (precision 0), FP32 (1) or (hi, lo) bf16 pairs (2; q_c stays FP32).
 * For precision 2 the folded charges are atom-major (K-contiguous for the NT
 * bf16 GEMM): qf_ov [nat][vpad][opad] and qf_vv [nat][vpad][vpad], as written
 * by dftb_convert_bf16pair_atom_major. */
typedef struct CasidaOperator {
  int nat, nocc, nvirt, opad, vpad;
  int lc;          /* 1: long-range exchange terms on */
  int precision;   /* 0 fp64, 1 fp32, 2 bf16x3 */
  int atom_chunk;  /* atoms per exchange slice (0 = auto) */
  int tile_expand, tile_contract; /* GEMM tile edge for the expand / contract launches (0 auto, 64, 128) */
  int atom_limit;  /* > 0: exchange over the first atom_limit atoms only (autotuner timing probe) */
  const double* omega;  /* [opad*vpad], padding 1.0 */
  const double* gamma;  /* [nat*nat] */
  const void* q_c;      /* [nat][opad][vpad]  f64 (precision 0) or f32; NULL in
                           dftb_casida_matvec: the Coulomb term is left out of
                           (A+B)V (the caller adds it) */
  const void* q_oo[2];  /* [nocc][nat][opad] */
  const void* q_ov[2];  /* [nocc][nat][vpad] */
  const void* qf_ov[2]; /* [nocc][nat][vpad] gamma_lr folded */
  const void* qf_vv[2]; /* [vpad][nat][vpad] gamma_lr folded */
  /* exchange = 1: long-range exchange in the AO basis (no q_oo/q_ov/qf_*): with
   * C_o, C_v the active MO coefficients, Y = S C, npad = Norb padded to 8 and
   * G = gamma_lr on orbital pairs:
   *   ao_v    = [Y_v; C_v]        [2 npad][vpad]
   *   ao_o    = [Y_o; C_o]        [2 npad][opad]
   *   ao_cyvt = [C_v; Y_v]^T      [vpad][2 npad]
   *   ao_w    = [C_o; Y_o]^T      [opad][2 npad]
   * in the operand type of the precision ([1] = lo halves for BF16x3). */
  int exchange;
  int npad;
  const void* ao_v[2];
  const void* ao_o[2];
  const void* ao_cyvt[2];
  const void* ao_w[2];
  const double* gam_ao;  /* gamma_lr [nat + 1][nat + 1], last row and column zero */
  const int* amap;       /* [npad] atom of each orbital (nat for the padding) */
} CasidaOperator;

size_t dftb_casida_matvec_workspace(const CasidaOperator* op, int nvec);
int dftb_casida_matvec(const CasidaOperator* op, const double* V, int nvec,
                       double* ApBV, double* AmBV, void* workspace, size_t ws_bytes, void* stream);

/* Hermitian operator of HermitianDavidson.Rv (Solver.py:431-444), used when
 * long_range_correction=0: u = om*v + 4 sqrt(omega) q^T gamma q (sqrt(omega) v),
 * om = omega^2 + 2 omega*shift.  sqrt_omega and om are [opad*vpad]. */
size_t dftb_hermitian_matvec_workspace(const CasidaOperator* op, int nvec);
int dftb_hermitian_matvec(const CasidaOperator* op, const double* sqrt_omega, const double* om,
                          const double* V, int nvec, double* RV,
                          void* workspace, size_t ws_bytes, void* stream);

/* Fold: qf[p][A][q] = sum_B gamma[A][B] q[p][B][q] for p < np, q < qpad
 * (the gamma_lr application of Solver.py:598-604 moved out of the matvec). */
int dftb_fold_gamma(const double* gamma, int nat, const double* q, long np, long qpad, double* qf,
                    void* stream);

/* Conversions of operands: FP64 -> FP32, FP64 -> (hi, lo) bf16 pairs. */
int dftb_convert_f32(const double* x, long n, float* y, void* stream);
int dftb_convert_bf16pair(const double* x, long n, unsigned short* hi, unsigned short* lo, void* stream);
/* q[p][A][x] (p < np, x < qpad) -> (hi, lo)[A][x][p] with p zero-padded to nppad. */
int dftb_convert_bf16pair_atom_major(const double* q, long np, long nppad, int nat, long qpad,
                                     unsigned short* hi, unsigned short* lo, void* stream);

/* ---- (c) Davidson support -------------------------------------------------
 * Residuals and preconditioner of Solver.py:809-884 / 480-530.
 *   W[j] = P[j] - w[j] X[j];  norms[j] = ||W[j]||_2            (j < k, length n;
 *                                   workspace >= dftb_norms_workspace(k) bytes)
 *   Q[t] = W[sel[t]] / (w[sel[t]] - omega) with |w - omega| < 1e-6 replaced by
 *          1e-6 * omega (Solver.py:865-868)                     (t < nsel)
 * w, sel and norms live in device memory. */
size_t dftb_norms_workspace(int k);
int dftb_residual(const double* P, const double* X, const double* w, int k, long n,
                  double* W, double* norms, void* workspace, size_t ws_bytes, void* stream);
int dftb_precondition(const double* W, const double* w, const int* sel, int nsel,
                      const double* omega, long n, double* Q, void* stream);
int dftb_row_norms(const double* X, int k, long n, double* norms, void* workspace, size_t ws_bytes,
                   void* stream);
/* Projections of the Davidson subspace matrices (M+- = bs^T (A+-B) bs,
 * Solver.py:757-761, and the tensordot overlaps of 890-897):
 *   out[p * ldo + q] = sum_{e < n} X[p * ldx + e] Y[q * ldy + e]   (p < P, q < Q)
 * split over n (workspace >= dftb_gram_workspace(P, Q, n) bytes), partial sums
 * added in a fixed order (deterministic). */
size_t dftb_gram_workspace(int P, int Q, long n);
int dftb_gram(const double* X, long ldx, int P, const double* Y, long ldy, int Q, long n, double* out, long ldo,
              void* workspace, size_t ws_bytes, void* stream);
/* X[j] *= s[j] */
int dftb_scale_rows(double* X, const double* s, int k, long n, void* stream);

/* Orthonormalise the m rows of Q against the l orthonormal rows of basis
 * (two passes of classical Gram-Schmidt) and among themselves (two passes of
 * Cholesky-QR with a rank tolerance).  Rows found linearly dependent
 * (relative norm < tol after projection) are flagged keep[t] = 0 and zeroed.
 * Replaces the economic QR of [bs | Q] at Solver.py:893-897. */
size_t dftb_orthonormalize_workspace(int l, int m, long n);
int dftb_orthonormalize(const double* basis, int l, double* Q, int m, long n, double tol,
                        int* keep, void* workspace, size_t ws_bytes, void* stream);

/* Subspace step of nonHermitianDavidson (Solver.py:759-798) for the lowest
 * nev roots: Mmsq = sqrtm(Mm), Mh = Mmsq Mp Mmsq, Mh Tb = Tb diag(w^2),
 * Rb = Mmsq Tb / sqrt(w), Lb = Mp Rb / w.  Mp, Mm: [l][ld] (device).
 * Outputs w [nev], Tb, Rb, Lb [l][nev] (column n = root n).  info[0] = 0 or
 * DFTB_ERR_NOT_SPD (Mm not positive definite or w^2 <= 0); diag[0] =
 * sum |Mh - Mh^T|, diag[1] = sum |Lb^T Rb - 1| (the two ExcitedStatesError
 * checks of Solver.py:766-770 and 795-798). */
size_t dftb_subspace_workspace(int l, int nev);
int dftb_subspace_nonhermitian(const double* Mp, const double* Mm, int l, int ld, int nev,
                               double* w, double* Tb, double* Rb, double* Lb, int* info, double* diag,
                               void* workspace, size_t ws_bytes, void* stream);
/* The same step by a Cholesky factor, for the iterations (Solver.py:759-798,
 * with the roots and Rb, Lb of Mh = Mmsq Mp Mmsq obtained without Mmsq): with
 * Mm = L L^T, H' = L^T Mp L has the roots w^2 of Mh; for H' T' = T' diag(w^2),
 * Rb = L T' / sqrt(w) and Lb = Mp Rb / w equal those of
 * dftb_subspace_nonhermitian.  Outputs w, Rb, Lb; info as there (info[1]: the
 * failing pivot); diag[0] = sum |H' - H'^T|, diag[1] = sum |Lb^T Rb - 1|.
 * Workspace: dftb_subspace_workspace(l, nev). */
int dftb_subspace_nonhermitian_rl(const double* Mp, const double* Mm, int l, int ld, int nev,
                                  double* w, double* Rb, double* Lb, int* info, double* diag,
                                  void* workspace, size_t ws_bytes, void* stream);
/* Tb = Mm^{-1/2} Rb diag(sqrt(w)): the eigenvectors of Mh belonging to the Rb of
 * the Cholesky step (C = bs Tb, Solver.py:906-910).  Workspace as above. */
int dftb_subspace_tb(const double* Mm, int l, int ld, int nev, const double* w, const double* Rb,
                     double* Tb, void* workspace, size_t ws_bytes, void* stream);
/* M[:l,:l] <- (M + M^T)/2; asym[0] = max|M - M^T| / max|M| (asym holds 3 doubles:
 * asym[1] and asym[2] receive the two maxima).  Used for the
 * subspace matrices in the mixed-precision modes, where the products carry
 * FP32-level rounding and the reference's 1e-10 Hermiticity test
 * (Solver.py:766-770) applies to FP64 only. */
int dftb_symmetrize(double* M, int l, int ld, double* asym, void* stream);
/* Symmetric square root and inverse square root of an SPD matrix (coupled
 * Newton-Schulz iteration on MFMA GEMMs; replaces scipy.linalg.sqrtm). */
size_t dftb_sqrtm_workspace(int l);
int dftb_sqrtm(const double* A, int l, int ld, double* Asq, double* Ainvsq, int* iters, double* err,
               void* workspace, size_t ws_bytes, void* stream);
/* Hermitian subspace step (Solver.py:482-484): H T = T diag(w2), lowest nev. */
int dftb_subspace_hermitian(const double* H, int l, int ld, int nev, double* w2, double* T,
                            int* info, void* workspace, size_t ws_bytes, void* stream);
/* Symmetric eigensolver used by both (lowest nev pairs, ascending):
 * Householder tridiagonalisation, bisection, inverse iteration. */
size_t dftb_syevx_workspace(int l, int nev);
int dftb_syevx(const double* A, int l, int ld, int nev, double* evals, double* evecs,
               void* workspace, size_t ws_bytes, void* stream);

/* out[j][e] = X[j][e] * rs[j] * cs[e]  (rs or cs may be NULL): the X+-Y
 * scalings of Solver.py:561-568. */
int dftb_scale2(const double* X, const double* rs, const double* cs, int k, long n, double* out, void* stream);
/* s[j] = sum_e X[j][e]^2 wgt[e]: Lambda2 diagnostic (LR_TDDFTB.py:561-565) and
 * the Lambda2 ordering of Solver.py:348-375. */
int dftb_weighted_sumsq(const double* X, const double* wgt, int k, long n, double* s, void* stream);
/* O[i][a] /= sqrt(Oii[i] Oaa[a]) on the padded grid (LR_TDDFTB.py:538-560). */
int dftb_oia_normalize(double* O, const double* Oii, const double* Oaa, int nocc, int nvirt, int opad, int vpad,
                       void* stream);

/* Unit guess vectors at the given flat indices (Solver.py:232-251). */
int dftb_unit_vectors(const long* idx, int k, long n, double* X, void* stream);

/* ---- batched ensemble: one persistent launch ---------------------------------
 * Solves the lowest k roots of the non-Hermitian Casida problem (the algorithm
 * of Solver.py:637-917) for many molecules of uneven size in one launch: one
 * workgroup per molecule runs its whole Davidson loop on the device.  The host
 * fills one EnsMolecule per molecule (device array, largest first): charges in
 * the engine layouts (FP64, plus FP32 copies q32 = {q_oo, q_ov, qf_ov, qf_vv}
 * for precision 1), nguess initial unit-vector indices, the expansion storage
 * (cap <= 1024 vectors), a workspace of dftb_ensemble_workspace() bytes and the
 * outputs.  status = {iterations, code, final subspace size}; code 1 =
 * converged, 0 = maxiter reached, -3 = (A-B) projection not positive definite,
 * -4 = Mh not Hermitian, -6 = expansion capacity exhausted. */
typedef struct EnsMolecule {
  int nat, nocc, nvirt, opad, vpad, k, cap, nguess, lc, precision;
  const double* omega;                  /* [opad*vpad], padding 1.0 */
  const double* gamma;                  /* [nat][nat] */
  const double* q_c;                    /* [nat][opad][vpad] */
  const double* q_oo;                   /* [nocc][nat][opad] */
  const double* q_ov;                   /* [nocc][nat][vpad] */
  const double* qf_ov;                  /* [nocc][nat][vpad] */
  const double* qf_vv;                  /* [vpad][nat][vpad] */
  const void* q32[4];                   /* FP32 copies (precision 1) */
  const long* guess;                    /* [nguess] flat padded indices */
  double *bs, *bp, *bm;                 /* [cap][opad*vpad] */
  double *Mp, *Mm;                      /* [cap][cap] */
  void* work;                           /* dftb_ensemble_workspace() bytes */
  double *w, *C, *XmY, *XpY, *resid;    /* outputs: [k], [k][N] x3, [k] */
  int* status;                          /* [3] */
} EnsMolecule;

size_t dftb_ensemble_workspace(const EnsMolecule* m);
int dftb_ensemble_solve(const EnsMolecule* mols_dev, int nmol, double conv, int maxiter, void* stream);

/* ---- learned initial space / preconditioner ---------------------------------
 * No counterpart in the reference: the engine's learned part (config of the
 * learned Davidson guess).  out[x] = sum_A P[A][x] Q[A][x] (P, Q: [nat][N]). */
int dftb_coldot(const double* P, long ldp, const double* Q, long ldq, int nat, long N, double* out,
                void* stream);
/* Per-pair features [N][nf] (float): omega, ||q||, A_ia,ia, B_ia,ia, A-B, sqrt((A+B)(A-B)). */
int dftb_pair_features(const double* omega, const double* qn2, const double* cc, const double* cl,
                       const double* xl, long N, int nf, float* F, void* stream);
/* (A+-B) restricted to M selected pairs (ip[p], ap[p]) from the engine-layout charges;
 * Kc = Coulomb block (p|q). */
int dftb_restricted_ab(const int* ip, const int* ap, int M, int nat, int opad, int vpad, const double* omega,
                       const double* Kc, const double* q_oo, const double* q_ov, const double* qf_ov,
                       const double* qf_vv, int lc, double* ApB, double* AmB, void* stream);
/* MLP steps (FP32): Z += b, H = silu(Z) (act) or Z; G *= silu'(Z); db = column sums of G;
 * weighted MSE loss partials and its gradient; Adam; feature normalisation. */
int dftb_mlp_bias_act(float* Z, const float* b, long rows, int cols, int act, float* H, void* stream);
int dftb_mlp_act_bwd(float* G, const float* Z, long n, void* stream);
int dftb_mlp_colsum(const float* G, long rows, int cols, float* db, void* stream);
int dftb_mlp_mse(const float* Y, const float* T, const float* w, long rows, int nout, float* G,
                 float* partial, int nblocks, void* stream);
int dftb_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
              int step, float wd, void* stream);
int dftb_mlp_normalize(float* X, long rows, int cols, const float* mean, const float* inv_std, void* stream);

/* ---- ground-state inputs: Slater-Koster matrices ---------------------------
 * S, H0 (Norb x Norb, row-major) and, if D != NULL, the dipole matrix as three
 * planes D[c][mu][nu] (c = x, y, z) for s and p valence orbitals: replaces the
 * f2py slako.h0ands and slako.dipolematrix (DFTB/extensions/slako.f90:77-394,
 * called at DFTB/SKMatrixElements.py:217-223, 346-352).  Device arrays:
 * pos [nat][3] (bohr), orb_atom / orb_l / orb_m [norb] (atom, l, real m of each
 * orbital, m = -1, 0, 1 -> y, z, x), atom_type [nat] (0 = lightest element),
 * orben [norb] (H0 diagonal), knots [npair][npts] and coefs
 * [npair][SLAKO_NTAB][npts] (rows 0-6 S tables, 7-13 H0 tables, 14-28 dipole
 * tables, by the reference's table index; B-spline knots and coefficients of
 * scipy's splrep, cubic) with pair index tmax (tmax + 1) / 2 + tmin.  Splines
 * are evaluated as splines.f90:25-94 does (interval l = INT((x - t1) / dt) + 1,
 * clamped to the end values outside the table), pairs at r >= cutoff are zero.
 * Elements of an unsupported input (l > 1, pair index >= npair) are NaN.
 * H0 = NULL skips H0; the matrices are exactly symmetric. */
enum { SLAKO_NTAB = 29 };
/* Overlap S[mu][nu] = <mu(geometry 1)|nu(geometry 2)> between the valence
 * orbitals of two geometries of one molecule (pos1, pos2 [nat][3]): replaces the
 * f2py slako.overlap12 (slako.f90:576-687, called from
 * ScalarCoupling.py:101-131).  Same tables and conventions as
 * dftb_slako_matrices; no proximity cutoff, not symmetric. */
int dftb_slako_overlap12(const double* pos1, const double* pos2, const int* orb_atom, const int* orb_l,
                         const int* orb_m, const int* atom_type, int nat, int norb, const double* knots,
                         const double* coefs, int npair, int npts, double* S, void* stream);
int dftb_slako_matrices(const double* pos, const int* orb_atom, const int* orb_l, const int* orb_m,
                        const int* atom_type, const double* orben, int nat, int norb, const double* knots,
                        const double* coefs, int npair, int npts, double cutoff, double* S, double* H0, double* D,
                        void* stream);

/* Nuclear-derivative contractions of the LC-TD-DFTB gradient: for every atom C,
 *   grad[3C + c] = <AS, dS/dR_Cc> + <AH, dH0/dR_Cc> + <BG, dgamma_AO/dR_Cc>
 *                + <BL, dgamma_lr_AO/dR_Cc>        (<X, Y> = sum_mn X_mn Y_mn)
 * with the derivative tensors of slako.f90 gradients_H0andS (402-569) and
 * GammaApproximation.gamma_AOwise (388-450), as ExcGradients.gradient_lc
 * (ExcGradients.py:286-711) contracts them, evaluated per atom pair and never
 * stored.  AS/AH/BG/BL: [norb][norb] device matrices (NULL = term absent);
 * atom_off [nat+1] orbital offsets; sigma [nat] Gaussian widths 1/(sqrt(pi) U);
 * rlr the long-range radius; knots / coefs as dftb_slako_matrices; dcoefs
 * [npair][14][npts] the first-derivative B-spline coefficients of the S (rows
 * 0-6) and H0 (7-13) tables (splder's recurrence, on the knots t[1..npts-2]).
 * grad [3 nat] is written (not accumulated).  No reference counterpart as one
 * function: it replaces the tensordot contractions of ExcGradients.py:477-496,
 * 694-711 over the [3 Nat][Norb][Norb] tensors. */
int dftb_grad_pairs(const double* pos, const int* atom_type, const int* atom_off, const int* orb_l, const int* orb_m,
                    const double* sigma, double rlr, const double* knots, const double* coefs, const double* dcoefs,
                    int npair, int npts, double cutoff, const double* AS, const double* AH, const double* BG,
                    const double* BL, int nat, int norb, double* grad, void* stream);

/* Gradient of the repulsive energy, replaces ExcGradients.gradient_Vrep
 * (ExcGradients.py:181-210): grad[3C + c] = sum_{D != C} Vrep'(R_CD)
 * (R_C - R_D)_c / R_CD with Vrep' = 0 for R >= dmax[pair] (RepulsivePotential.py
 * getVrep).  knots [npair][npts], dcoefs [npair][npts] the first-derivative
 * coefficients of the cubic B-spline (as for dftb_grad_pairs), dmax [npair]. */
int dftb_grad_vrep(const double* pos, const int* atom_type, const double* knots, const double* dcoefs,
                   const double* dmax, int npair, int npts, int nat, double* grad, void* stream);

#ifdef __cplusplus
}
#endif
#endif
