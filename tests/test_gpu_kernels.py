"""GPU parity tests of the HIP kernels against the CPU oracle (run with -m gpu)."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import ground_state
from oracle import tddftb_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda")


def T(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=_dev())


# ---------------------------------------------------------------------------
# generic MFMA GEMM
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("at,bt", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (37, 61, 29), (130, 257, 203), (64, 64, 1000)])
def test_gemm_f64(at, bt, M, N, K, rng):
    from dftbaby_amd.engine import gemm
    A = rng.standard_normal((K, M) if at else (M, K))
    B = rng.standard_normal((N, K) if bt else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 1.5 * (A.T if at else A) @ (B.T if bt else B) - 0.5 * C0
    C = T(C0)
    gemm(T(A), T(B), C, M, N, K, A.shape[1], B.shape[1], N, a_trans=at, b_trans=bt, alpha=1.5, beta=-0.5)
    np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("at,bt", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(20, 300, 70), (33, 130, 17), (40, 1000, 300), (300, 20, 250), (200, 45, 33),
                                   (700, 40, 3000)])
def test_gemm_f64_narrow_tiles(at, bt, M, N, K, rng):
    """The 32/48-wide tiles used when one side is 20-48 (Davidson basis products),
    including ragged edges and the edge rows/columns of partially filled tiles."""
    from dftbaby_amd.engine import gemm
    A = rng.standard_normal((K, M) if at else (M, K))
    B = rng.standard_normal((N, K) if bt else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = -1.0 * (A.T if at else A) @ (B.T if bt else B) + 1.0 * C0
    C = T(C0)
    gemm(T(A), T(B), C, M, N, K, A.shape[1], B.shape[1], N, a_trans=at, b_trans=bt, alpha=-1.0, beta=1.0, tile=64)
    np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-12, atol=1e-10)


def test_gram_narrow_splitk(rng):
    from dftbaby_amd.engine import gram, Workspace
    for P, Q in ((300, 40), (40, 40), (700, 20)):
        X = rng.standard_normal((P, 50021))
        Y = rng.standard_normal((Q, 50021))
        out = torch.zeros((P, Q), dtype=torch.float64, device=_dev())
        gram(T(X), P, T(Y), Q, 50021, out, Workspace(_dev()))
        np.testing.assert_allclose(out.cpu().numpy(), X @ Y.T, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("M,N,K,at", [(5, 20001, 7, 1), (20, 50000, 33, 1), (40, 50000, 100, 1), (64, 20001, 70, 0),
                                       (24, 50000, 5, 0), (33, 40000, 129, 1)])
def test_skinny_output_products(M, N, K, at, rng):
    """C = alpha op(A) B + beta C with M <= 64 rows and a long N (the basis
    products of the Davidson loop), batch 2 with strides."""
    from dftbaby_amd import _lib
    A = rng.standard_normal((2, K, M) if at else (2, M, K))
    B = rng.standard_normal((2, K, N))
    C0 = rng.standard_normal((2, M, N))
    Ad, Bd, Cd = T(A), T(B), T(C0)
    g = _lib.GemmArgs()
    g.A, g.B, g.C = Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr()
    g.lda, g.ldb, g.ldc = (M if at else K), N, N
    g.sA, g.sB, g.sC = A[0].size, B[0].size, C0[0].size
    g.cs2, g.M, g.N, g.K, g.batch, g.ksplit, g.kchunk = 1, M, N, K, 2, 1, 0
    g.alpha, g.beta, g.a_trans, g.hmod = -0.5, 2.0, at, 1
    _lib.call("dftb_gemm", ctypes.byref(g), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    for b in range(2):
        ref = -0.5 * ((A[b].T if at else A[b]) @ B[b]) + 2.0 * C0[b]
        np.testing.assert_allclose(Cd[b].cpu().numpy(), ref, rtol=1e-12, atol=1e-11)


def test_gram_strided_output_and_residual_norms(rng):
    """dftb_gram into a strided block of a larger matrix (the Davidson's new
    columns of M), and the sliced residual / row-norm kernels (deterministic:
    two calls give identical bits) against numpy."""
    from dftbaby_amd.engine import gram, residual, row_norms, Workspace
    ws = Workspace(_dev())
    l, m, n = 230, 80, 200003
    X = rng.standard_normal((l, n))
    Y = rng.standard_normal((m, n))
    M = torch.full((300, 400), 7.0, dtype=torch.float64, device=_dev())
    gram(T(X), l, T(Y), m, n, M, ws, ldo=400, o_off=5 * 400 + 11)
    Mh = M.cpu().numpy()
    np.testing.assert_allclose(Mh[5:5 + l, 11:11 + m], X @ Y.T, rtol=1e-11, atol=1e-9)
    assert np.all(Mh[:5] == 7.0) and np.all(Mh[5:5 + l, :11] == 7.0) and np.all(Mh[5:5 + l, 11 + m:] == 7.0)
    k = 40
    P, V = rng.standard_normal((k, n)), rng.standard_normal((k, n))
    w = rng.standard_normal(k)
    W = torch.empty((k, n), dtype=torch.float64, device=_dev())
    nr = torch.empty(k, dtype=torch.float64, device=_dev())
    residual(T(P), T(V), T(w), k, n, W, nr, ws)
    ref = P - w[:, None] * V
    np.testing.assert_allclose(W.cpu().numpy(), ref, rtol=0, atol=1e-14)
    np.testing.assert_allclose(nr.cpu().numpy(), np.linalg.norm(ref, axis=1), rtol=1e-13)
    a = nr.cpu().numpy().copy()
    residual(T(P), T(V), T(w), k, n, W, nr, ws)
    assert np.array_equal(a, nr.cpu().numpy())
    row_norms(T(V), k, n, nr, ws)
    np.testing.assert_allclose(nr.cpu().numpy(), np.linalg.norm(V, axis=1), rtol=1e-13)


def test_gemm_f64_batched_remap_splitk(rng):
    from dftbaby_amd.engine import gemm
    from dftbaby_amd import _lib
    # batch of 3, remapped output: col n -> (n // 5) * 1000 + row * 5 + n % 5
    M, N, K, nb = 7, 15, 33, 3
    A = rng.standard_normal((nb, M, K))
    B = rng.standard_normal((K, N))
    Cbig = torch.zeros(nb * 4000, dtype=torch.float64, device=_dev())
    g = _lib.GemmArgs()
    g.A, g.B, g.C = T(A).data_ptr(), T(B).data_ptr(), Cbig.data_ptr()
    Ad, Bd = T(A), T(B)
    g.A, g.B = Ad.data_ptr(), Bd.data_ptr()
    g.lda, g.ldb, g.ldc = K, N, 5
    g.sA, g.sB, g.sC = M * K, 0, 4000
    g.cdiv, g.cstr, g.cs2 = 5, 1000, 1
    g.M, g.N, g.K, g.batch, g.ksplit, g.kchunk = M, N, K, nb, 1, 0
    g.alpha, g.beta = 1.0, 0.0
    g.in_type, g.out_type, g.tile = 0, 0, 0
    _lib.call("dftb_gemm", ctypes.byref(g), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    C = Cbig.cpu().numpy().reshape(nb, 4000)
    for b in range(nb):
        ref = A[b] @ B
        for n in range(N):
            got = C[b, (n // 5) * 1000 + np.arange(M) * 5 + n % 5]
            np.testing.assert_allclose(got, ref[:, n], rtol=1e-12, atol=1e-12)
    # split-K partials
    from dftbaby_amd.engine import gram, Workspace
    X = rng.standard_normal((5, 100003))
    Y = rng.standard_normal((3, 100003))
    out = torch.zeros((5, 3), dtype=torch.float64, device=_dev())
    gram(T(X), 5, T(Y), 3, 100003, out, Workspace(_dev()))
    np.testing.assert_allclose(out.cpu().numpy(), X @ Y.T, rtol=1e-11, atol=1e-10)


@pytest.mark.parametrize("M,N,K", [(5, 9, 7), (130, 200, 96), (64, 64, 1024)])
def test_gemm_f32(M, N, K, rng):
    from dftbaby_amd.engine import gemm
    from dftbaby_amd import _lib
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = torch.zeros((M, N), dtype=torch.float64, device=_dev())
    gemm(T(A, torch.float32), T(B, torch.float32), C, M, N, K, K, N, N, in_type=_lib.GEMM_IN_F32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=2e-5, atol=2e-5 * np.sqrt(K))


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (70, 130, 96), (128, 256, 1024)])
def test_gemm_bf16x3(M, N, K, rng):
    """hi*hi + hi*lo + lo*hi split product: ~2^-16 relative per term, FP32 accumulation."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))          # NT form: B is [n][k]
    Ad, Bd = T(A), T(B)
    parts = []
    for X in (Ad, Bd):
        hi = torch.empty(X.shape, dtype=torch.int16, device=_dev())
        lo = torch.empty(X.shape, dtype=torch.int16, device=_dev())
        _lib.call("dftb_convert_bf16pair", _p(X), X.numel(), _p(hi), _p(lo), _stream())
        parts.append((hi, lo))
    C = torch.zeros((M, N), dtype=torch.float64, device=_dev())
    g = _lib.GemmArgs()
    g.A, g.A2 = parts[0][0].data_ptr(), parts[0][1].data_ptr()
    g.B, g.B2 = parts[1][0].data_ptr(), parts[1][1].data_ptr()
    g.C = C.data_ptr()
    g.lda, g.ldb, g.ldc = K, K, N
    g.cs2 = 1
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, 1, 1
    g.alpha, g.beta = 1.0, 0.0
    g.b_trans = 1
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_F64
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    ref = A @ B.T
    err = np.abs(C.cpu().numpy() - ref).max()
    assert err < 3e-5 * np.sqrt(K), err


def _bf_pair(X):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream
    hi = torch.empty(X.shape, dtype=torch.int16, device=_dev())
    lo = torch.empty(X.shape, dtype=torch.int16, device=_dev())
    _lib.call("dftb_convert_bf16pair", _p(X), X.numel(), _p(hi), _p(lo), _stream())
    return hi, lo


@pytest.mark.parametrize("M,N,K,batch,tile", [(256, 128, 128, 1, 256), (300, 260, 200, 2, 256), (513, 300, 1000, 3, 0),
                                              (260, 129, 136, 1, 0), (256, 256, 128, 1, 512),
                                              (300, 520, 200, 2, 512), (513, 300, 1000, 3, 512),
                                              (260, 129, 136, 1, 512)])
def test_gemm_bf16x3_pipelined(M, N, K, batch, tile, rng):
    """Large-tile pipelined BF16x3 kernel: partial tiles in M, N and K (K tail not a
    multiple of 32 loads zeros), strided batches, beta accumulation."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _stream
    lda, ldb = K + 8, K + 16
    A = rng.standard_normal((batch, M, lda))
    B = rng.standard_normal((batch, N, ldb))
    C0 = rng.standard_normal((batch, M, N))
    (ah, al), (bh, bl) = _bf_pair(T(A)), _bf_pair(T(B))
    C = T(C0.copy())
    g = _lib.GemmArgs()
    g.A, g.A2, g.B, g.B2, g.C = ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr(), C.data_ptr()
    g.lda, g.ldb, g.ldc = lda, ldb, N
    g.sA, g.sB, g.sC = M * lda, N * ldb, M * N
    g.cs2 = 1
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, batch, 1
    g.alpha, g.beta = 0.5, 1.0
    g.b_trans = 1
    g.tile = tile
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_F64
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    ref = 0.5 * np.einsum("bmk,bnk->bmn", A[:, :, :K], B[:, :, :K]) + C0
    err = np.abs(C.cpu().numpy() - ref).max()
    assert err < 3e-5 * np.sqrt(K), err


@pytest.mark.parametrize("tile", [256, 512])
def test_gemm_bf16x3_pipelined_transposed_copy(tile, rng):
    """bf16-pair output written twice: C[b][m][n] and the transposed copy
    Ct[b * sCt + n * ldct + m] (partial tiles in M and N, strided batches)."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _stream
    M, N, K, batch, ldct = 264, 300, 104, 2, 280
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((batch, N, K))
    (ah, al), (bh, bl) = _bf_pair(T(A)), _bf_pair(T(B))
    ch = torch.zeros(batch * M * N, dtype=torch.int16, device=_dev())
    cl = torch.zeros_like(ch)
    th = torch.full((batch * N * ldct,), 7, dtype=torch.int16, device=_dev())
    tl = torch.full_like(th, 7)
    g = _lib.GemmArgs()
    g.A, g.A2, g.B, g.B2, g.C, g.C2 = (ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr(), ch.data_ptr(),
                                       cl.data_ptr())
    g.Ct, g.Ct2, g.ldct, g.sCt = th.data_ptr(), tl.data_ptr(), ldct, N * ldct
    g.lda, g.ldb, g.ldc, g.cs2 = K, K, N, 1
    g.sA, g.sB, g.sC = 0, N * K, M * N
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, batch, 1
    g.alpha, g.beta = 1.0, 0.0
    g.b_trans = 1
    g.tile = tile
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_BF16PAIR
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    f = lambda t: (t.to(torch.int32) & 0xFFFF).cpu().numpy().astype(np.uint32) << 16
    got = (f(ch).view(np.float32).astype(np.float64) + f(cl).view(np.float32).astype(np.float64)).reshape(batch, M, N)
    gt = (f(th).view(np.float32).astype(np.float64) + f(tl).view(np.float32).astype(np.float64)).reshape(batch, N, ldct)
    ref = np.einsum("mk,bnk->bmn", A, B)
    assert (np.abs(got - ref) <= 3e-5 * np.sqrt(K) + 2.0 ** -15 * np.abs(ref)).all()
    assert np.array_equal(gt[:, :, :M], got.transpose(0, 2, 1))       # the same pairs, transposed
    pad = (th.view(batch, N, ldct)[:, :, M:] == 7).all() and (tl.view(batch, N, ldct)[:, :, M:] == 7).all()
    assert bool(pad)                                                  # columns past M untouched
    # the copy needs M, ldct, sCt multiples of 8 and the bf16-pair output
    g.M = M - 4
    with pytest.raises(_lib.HipError, match="invalid argument"):
        _lib.call("dftb_gemm", ctypes.byref(g), _stream())


@pytest.mark.parametrize("tile", [256, 512])
@pytest.mark.parametrize("keys", ["atoms", "scattered", "identity"])
def test_gemm_bf16x3_pipelined_hadamard(tile, keys, rng):
    """Hadamard factor in the epilogue of the pipelined kernel: out[r][c] =
    alpha had[key(r)][key(c)] (A B^T)[r][c] with key(x) = hmap[x % hmod] (or x % hmod).
    'atoms' (orbital -> atom map, rows wrapping at hmod) takes the per-tile LDS block;
    'scattered' keys span more than the LDS holds and take the global-memory reads."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _stream
    K = 160
    if keys == "atoms":
        nat, hmod = 61, 200
        per = np.concatenate([np.full(4, a) for a in range(40)] + [np.arange(40, 61)] + [np.full(19, 61)])
        hmap = per.astype(np.int32)                 # 160 + 21 + 19 = 200 orbitals, padding -> row nat
        M = N = 2 * hmod
    elif keys == "scattered":
        nat, hmod = 299, 520
        hmap = rng.integers(0, nat + 1, hmod).astype(np.int32)
        M, N = 520, 520
    else:
        nat, hmod, hmap = 63, 64, None
        M, N = 300, 264
    ldh = nat + 1
    gam = rng.standard_normal((ldh, ldh))
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    (ah, al), (bh, bl) = _bf_pair(T(A)), _bf_pair(T(B))
    C = torch.zeros((M, N), dtype=torch.float64, device=_dev())
    G = T(gam)
    Hm = torch.from_numpy(hmap).to(_dev()) if hmap is not None else None
    g = _lib.GemmArgs()
    g.A, g.A2, g.B, g.B2, g.C = ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr(), C.data_ptr()
    g.lda, g.ldb, g.ldc, g.cs2 = K, K, N, 1
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, 1, 1
    g.alpha, g.beta = 0.75, 0.0
    g.b_trans = 1
    g.tile = tile
    g.had, g.hmod, g.ldh = G.data_ptr(), hmod, ldh
    g.hmap = Hm.data_ptr() if Hm is not None else 0
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_F64
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    key = (lambda x: hmap[x % hmod]) if hmap is not None else (lambda x: x % hmod)
    fac = gam[key(np.arange(M))][:, key(np.arange(N))].astype(np.float32).astype(np.float64)
    ref = 0.75 * fac * (A @ B.T)
    err = np.abs(C.cpu().numpy() - ref).max()
    assert err < 3e-5 * np.sqrt(K) * np.abs(gam).max(), err


@pytest.mark.parametrize("tile", [256, 512])
def test_gemm_bf16x3_pipelined_remap_pair_output(tile, rng):
    """Column-remapped bf16-pair output (the exchange 'expand' form) of the pipelined kernel."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _stream
    M, K, cdiv, groups = 264, 160, 48, 4          # N = groups * cdiv columns -> out[g][c][row]
    N = groups * cdiv
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    (ah, al), (bh, bl) = _bf_pair(T(A)), _bf_pair(T(B))
    oh = torch.zeros(groups * cdiv * M, dtype=torch.int16, device=_dev())
    ol = torch.zeros_like(oh)
    g = _lib.GemmArgs()
    g.A, g.A2, g.B, g.B2, g.C, g.C2 = (ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr(), oh.data_ptr(),
                                       ol.data_ptr())
    g.lda, g.ldb, g.ldc = K, K, 1
    g.cdiv, g.cstr, g.cs2 = cdiv, M, groups * M     # col n -> (n // cdiv) * M + (n % cdiv) * groups * M + row
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, 1, 1
    g.alpha, g.beta = 1.0, 0.0
    g.b_trans = 1
    g.tile = tile
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_BF16PAIR
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    f = lambda t: (t.to(torch.int32) & 0xFFFF).cpu().numpy().astype(np.uint32) << 16
    got = f(oh).view(np.float32).astype(np.float64) + f(ol).view(np.float32).astype(np.float64)
    got = got.reshape(cdiv, groups, M)                 # [n % cdiv][n // cdiv][row]
    ref = A @ B.T                                      # [row][n]
    ref = ref.reshape(M, groups, cdiv).transpose(2, 1, 0)
    # product error plus the rounding of the stored hi/lo pair (~16 significant bits)
    assert (np.abs(got - ref) <= 3e-5 * np.sqrt(K) + 2.0 ** -15 * np.abs(ref)).all()


@pytest.mark.parametrize("tile", [256, 512])
def test_gemm_bf16x3_pipelined_row_remap(tile, rng):
    """Row-remapped bf16-pair output (the swapped exchange 'expand'): row (g, a) -> a * rs2 + g * rstr."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _stream
    rdiv, groups, N, K = 40, 9, 136, 144           # M = groups * rdiv rows
    M = groups * rdiv
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((N, K))
    (ah, al), (bh, bl) = _bf_pair(T(A)), _bf_pair(T(B))
    oh = torch.zeros(M * N, dtype=torch.int16, device=_dev())
    ol = torch.zeros_like(oh)
    g = _lib.GemmArgs()
    g.A, g.A2, g.B, g.B2, g.C, g.C2 = (ah.data_ptr(), al.data_ptr(), bh.data_ptr(), bl.data_ptr(), oh.data_ptr(),
                                       ol.data_ptr())
    g.lda, g.ldb = K, K
    g.rdiv, g.rstr, g.rs2 = rdiv, N, groups * N     # out[a][g][n]
    g.cs2 = 1
    g.M, g.N, g.K, g.batch, g.ksplit = M, N, K, 1, 1
    g.alpha, g.beta = 1.0, 0.0
    g.b_trans = 1
    g.tile = tile
    g.in_type, g.out_type = _lib.GEMM_IN_BF16X3, _lib.GEMM_OUT_BF16PAIR
    _lib.call("dftb_gemm", ctypes.byref(g), _stream())
    f = lambda t: (t.to(torch.int32) & 0xFFFF).cpu().numpy().astype(np.uint32) << 16
    got = (f(oh).view(np.float32).astype(np.float64) + f(ol).view(np.float32).astype(np.float64)).reshape(rdiv, groups, N)
    ref = (A @ B.T).reshape(groups, rdiv, N).transpose(1, 0, 2)
    assert (np.abs(got - ref) <= 3e-5 * np.sqrt(K) + 2.0 ** -15 * np.abs(ref)).all()


def test_mfma_f64_lane_layout():
    """A rank-one product with distinct row/column scales pins the MFMA output layout."""
    from dftbaby_amd.engine import gemm
    M = N = 16
    K = 4
    A = np.zeros((M, K)); A[:, 0] = np.arange(1, M + 1)
    B = np.zeros((K, N)); B[0, :] = 100.0 * np.arange(1, N + 1)
    C = torch.zeros((M, N), dtype=torch.float64, device=_dev())
    gemm(T(A), T(B), C, M, N, K, K, N, N)
    got = C.cpu().numpy()
    print("row ids seen per output row:", np.round(got[:, 0] / 100.0).astype(int).tolist())
    np.testing.assert_array_equal(got, np.outer(A[:, 0], B[0]))


def test_mfma_f32_lane_layout():
    from dftbaby_amd.engine import gemm
    from dftbaby_amd import _lib
    M = N = 16
    K = 4
    A = np.zeros((M, K), dtype=np.float32); A[:, 0] = np.arange(1, M + 1)
    B = np.zeros((K, N), dtype=np.float32); B[0, :] = 100.0 * np.arange(1, N + 1)
    C = torch.zeros((M, N), dtype=torch.float64, device=_dev())
    gemm(T(A, torch.float32), T(B, torch.float32), C, M, N, K, K, N, N, in_type=_lib.GEMM_IN_F32)
    got = C.cpu().numpy()
    print("f32 row ids seen per output row:", np.round(got[:, 0] / 100.0).astype(int).tolist())
    np.testing.assert_array_equal(got, np.outer(A[:, 0], B[0]))


# ---------------------------------------------------------------------------
# (a) transition charges
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,lc", [("ethylene", 1), ("C8H10", 1), ("C20H22", 0)])
def test_trans_charges_reference_layout(name, lc):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    gs = ground_state(name, lc)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    nat, o, v = q_ov.shape
    Co, Cv, St = T(C[:, ao]), T(C[:, av]), T(S)
    g_oo = torch.empty((nat, o, o), dtype=torch.float64, device=_dev())
    g_ov = torch.empty((nat, o, v), dtype=torch.float64, device=_dev())
    g_vv = torch.empty((nat, v, v), dtype=torch.float64, device=_dev())
    opa = np.ascontiguousarray(gs.orbsPerAtom, dtype=np.int32)
    ws = Workspace(_dev())
    w, wn = ws.get(_lib.load().dftb_trans_charges_workspace(S.shape[0], o, v, nat))
    _lib.call("dftb_trans_charges", _p(Co), _p(Cv), _p(St), opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
              nat, S.shape[0], o, v, _p(g_oo), _p(g_ov), _p(g_vv), _p(w), wn, _stream())
    for got, ref in ((g_oo, q_oo), (g_ov, q_ov), (g_vv, q_vv)):
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=0, atol=1e-13)


def test_trans_charges_flake_invariants():
    """Reference-layout charges of the flake C604H68 (Nat 672, o = v = 1242: 25 GB
    of outputs, prefilled with NaN): sum_A q^{pq}_A = (C^T S C)_{pq} = delta_pq for
    every (p, q) of the three blocks, q^{pq} = q^{qp}, and eight atoms (first,
    last, six random) against the closed form evaluated on the host."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    from dftbaby_amd.groundstate import geometry as g
    from dftbaby_amd.groundstate.dftb2 import DFTB2
    gs = DFTB2(g.flake_c604h68(), long_range_correction=1, long_range_radius=3.03)
    gs.getEnergy(scf_conv=1e-8)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    nat, norb, o, v = len(gs.orbsPerAtom), S.shape[0], len(ao), len(av)
    Co, Cv, St = T(C[:, ao]), T(C[:, av]), T(S)
    q = [torch.full(sz, float("nan"), dtype=torch.float64, device=_dev()) for sz in ((nat, o, o), (nat, o, v),
                                                                                    (nat, v, v))]
    opa = np.ascontiguousarray(gs.orbsPerAtom, dtype=np.int32)
    ws = Workspace(_dev())
    w, wn = ws.get(_lib.load().dftb_trans_charges_workspace(norb, o, v, nat))
    _lib.call("dftb_trans_charges", _p(Co), _p(Cv), _p(St), opa.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
              nat, norb, o, v, _p(q[0]), _p(q[1]), _p(q[2]), _p(w), wn, _stream())
    eye_o = torch.eye(o, dtype=torch.float64, device=_dev())
    eye_v = torch.eye(v, dtype=torch.float64, device=_dev())
    assert float((q[0].sum(0) - eye_o).abs().max()) < 1e-12
    assert float(q[1].sum(0).abs().max()) < 1e-12
    assert float((q[2].sum(0) - eye_v).abs().max()) < 1e-12
    assert float((q[0] - q[0].transpose(1, 2)).abs().max()) < 1e-15      # rounding of the FMA order only
    assert float((q[2] - q[2].transpose(1, 2)).abs().max()) < 1e-15
    off = np.concatenate([[0], np.cumsum(opa)])
    SC = S @ C
    rng = np.random.default_rng(3)
    for A in [0, nat - 1] + list(rng.integers(1, nat - 1, 6)):
        sl = slice(off[A], off[A + 1])
        for (blk, P_, Q_) in ((0, ao, ao), (1, ao, av), (2, av, av)):
            ref = 0.5 * (C[sl][:, P_].T @ SC[sl][:, Q_] + SC[sl][:, P_].T @ C[sl][:, Q_])
            np.testing.assert_allclose(q[blk][A].cpu().numpy(), ref, rtol=0, atol=1e-13)


def test_trans_charges_engine_layout():
    from dftbaby_amd.engine import CasidaOperatorGPU
    gs = ground_state("C8H10", 1, 0.05, 7)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], S, gs.orbsPerAtom, omega, gs.getGamma(),
                                         gs.getGamma_lr(), 1, exchange="charges")
    nat, o, v = q_ov.shape
    qc = op.q["c"].cpu().numpy()
    np.testing.assert_allclose(qc[:, :o, :v], q_ov, atol=1e-13)
    assert np.all(qc[:, o:, :] == 0) and np.all(qc[:, :, v:] == 0)
    np.testing.assert_allclose(op.q["oo"].cpu().numpy()[:, :, :o], q_oo.transpose(1, 0, 2), atol=1e-13)
    np.testing.assert_allclose(op.q["ov"].cpu().numpy()[:, :, :v], q_ov.transpose(1, 0, 2), atol=1e-13)
    glr = gs.getGamma_lr()
    qf_vv = np.einsum("AB,Bab->aAb", glr, q_vv)
    np.testing.assert_allclose(op.q["qf_vv"].cpu().numpy()[:v, :, :v], qf_vv, atol=1e-12)
    np.testing.assert_allclose(op.qaa.cpu().numpy(), np.einsum("Aaa->aA", q_vv), atol=1e-13)


# ---------------------------------------------------------------------------
# (b) matvecs
# ---------------------------------------------------------------------------
def _operator(name, lc, sigma=0.05, seed=3, exchange="ao"):
    from dftbaby_amd.engine import CasidaOperatorGPU
    gs = ground_state(name, lc, sigma, seed)
    f, orbe, C, S = gs.getOccupation(), gs.getKSEnergies(), gs.getKSCoefficients(), gs.getOverlapMatrix()
    _, _, ao, av = O.active_orbitals(f, orbe)
    q_oo, q_ov, q_vv = O.trans_charges(C[:, ao], C[:, av], S, gs.orbsPerAtom)
    omega, _ = O.orbital_differences(orbe, f, ao, av)
    g, glr = gs.getGamma(), gs.getGamma_lr()
    op = CasidaOperatorGPU.from_orbitals(C[:, ao], C[:, av], S, gs.orbsPerAtom, omega, g, glr, lc, exchange=exchange)
    return op, (g, glr, q_oo, q_ov, q_vv, omega)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (2, 3e-5)])
def test_casida_matvec_small_block_ao_coulomb(precision, tol, rng):
    """Blocks of one or two vectors take the Coulomb term through the AO operands
    (the library is asked to leave it out): the same products as the oracle, and in
    FP64 the same as the charge route of a larger block."""
    op, (g, glr, q_oo, q_ov, q_vv, omega) = _operator("C12H14", 1, exchange="ao")
    o, v = omega.shape
    vs = rng.standard_normal((o, v, 5))
    refP = O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    refM = O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    V = op.to_padded(np.moveaxis(vs, 2, 0))
    P5, M5 = op.matvec(V, precision)
    for sl in (slice(0, 1), slice(1, 3)):
        P, M = op.matvec(V[sl].contiguous(), precision)
        gotP = np.moveaxis(op.from_padded(P).cpu().numpy(), 0, 2)
        gotM = np.moveaxis(op.from_padded(M).cpu().numpy(), 0, 2)
        scale = np.abs(refP).max()
        np.testing.assert_allclose(gotP, refP[:, :, sl], rtol=0, atol=tol * scale)
        np.testing.assert_allclose(gotM, refM[:, :, sl], rtol=0, atol=tol * scale)
        if precision == 0:
            np.testing.assert_allclose(P.cpu().numpy(), P5[sl].cpu().numpy(), rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("exchange", ["ao", "charges"])
@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 3e-5), (2, 3e-5)])
def test_casida_matvec(exchange, precision, tol, rng):
    """Both exchange formulations (AO-basis GEMMs, transition-charge contractions) vs the oracle."""
    op, (g, glr, q_oo, q_ov, q_vv, omega) = _operator("C12H14", 1, exchange=exchange)
    o, v = omega.shape
    nvec = 5
    vs = rng.standard_normal((o, v, nvec))
    refP = O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    refM = O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    V = op.to_padded(np.moveaxis(vs, 2, 0))
    P, M = op.matvec(V, precision)
    gotP = op.from_padded(P).cpu().numpy()
    gotM = op.from_padded(M).cpu().numpy()
    scale = np.abs(refP).max()
    assert np.abs(gotP - np.moveaxis(refP, 2, 0)).max() < tol * scale
    assert np.abs(gotM - np.moveaxis(refM, 2, 0)).max() < tol * scale
    # padding stays zero
    Pp = P.reshape(nvec, op.opad, op.vpad).cpu().numpy()
    assert np.all(Pp[:, o:, :] == 0) and np.all(Pp[:, :, v:] == 0)


@pytest.mark.parametrize("exchange,precision,tol", [("charges", 0, 1e-12), ("charges", 2, 3e-5), ("ao", 0, 1e-12),
                                                    ("ao", 1, 3e-5), ("ao", 2, 3e-5)])
def test_casida_matvec_flake(exchange, precision, tol, rng):
    """Armchair flake C180H36 (o = v = 378): the exchange GEMMs take the large-tile
    pipelined path (charges: expands with row remap, contracts with M >= 256;
    AO: 2 Norb-sized GEMMs with the gamma_lr Hadamard epilogue)."""
    op, (g, glr, q_oo, q_ov, q_vv, omega) = _operator("C180H36", 1, 0.05, 5, exchange=exchange)
    o, v = omega.shape
    assert o >= 256 and v >= 256
    nvec = 3
    vs = rng.standard_normal((o, v, nvec))
    refP = O.ApBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    refM = O.AmBv(g, glr, q_oo, q_ov, q_vv, omega, vs, 1)
    P, M = op.matvec(op.to_padded(np.moveaxis(vs, 2, 0)), precision, atom_chunk=50)
    scale = np.abs(refP).max()
    assert np.abs(op.from_padded(P).cpu().numpy() - np.moveaxis(refP, 2, 0)).max() < tol * scale
    assert np.abs(op.from_padded(M).cpu().numpy() - np.moveaxis(refM, 2, 0)).max() < tol * scale


def test_casida_matvec_atom_chunks(rng):
    """Exchange contractions split over atom slices accumulate to the same product."""
    op, (g, glr, q_oo, q_ov, q_vv, omega) = _operator("C8H10", 1, exchange="charges")
    o, v = omega.shape
    V = op.to_padded(rng.standard_normal((3, o, v)))
    P1, M1 = op.matvec(V, 0, atom_chunk=0)
    P2, M2 = op.matvec(V, 0, atom_chunk=3)
    np.testing.assert_allclose(P1.cpu().numpy(), P2.cpu().numpy(), atol=1e-13)
    np.testing.assert_allclose(M1.cpu().numpy(), M2.cpu().numpy(), atol=1e-13)


def test_hermitian_matvec(rng):
    op, (g, glr, q_oo, q_ov, q_vv, omega) = _operator("C12H14", 0)
    o, v = omega.shape
    vs = rng.standard_normal((4, o, v))
    shift = np.zeros_like(omega)
    om = omega ** 2 + 2 * omega * shift
    wq = q_ov * np.sqrt(omega)
    ref = np.stack([om * x + 4 * np.tensordot(wq, g @ np.tensordot(wq, x, axes=([1, 2], [0, 1])), axes=(0, 0))
                    for x in vs])
    sq = torch.sqrt(op.omega)
    omt = op.omega * op.omega
    R = op.hermitian_matvec(op.to_padded(vs), sq, omt)
    np.testing.assert_allclose(op.from_padded(R).cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# (c) subspace kernels
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("l,nev", [(1, 1), (2, 2), (9, 4), (60, 20), (100, 20), (180, 12), (257, 30)])
def test_syevx(l, nev, rng):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    A = rng.standard_normal((l, l))
    A = A + A.T
    if l >= 9:
        # exact degeneracy in the wanted range
        Q, _ = np.linalg.qr(rng.standard_normal((l, l)))
        ev = np.sort(rng.standard_normal(l))
        ev[1] = ev[2] = ev[3]
        A = (Q * ev) @ Q.T
    ld = l + 3
    Ad = torch.zeros((l, ld), dtype=torch.float64, device=_dev())
    Ad[:, :l] = T(A)
    ev = torch.empty(nev, dtype=torch.float64, device=_dev())
    V = torch.empty((l, nev), dtype=torch.float64, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_syevx_workspace(l, nev))
    _lib.call("dftb_syevx", _p(Ad), l, ld, nev, _p(ev), _p(V), _p(w), wn, _stream())
    ref = np.linalg.eigvalsh(A)[:nev]
    e = ev.cpu().numpy()
    Vh = V.cpu().numpy()
    np.testing.assert_allclose(e, ref, atol=1e-12 * max(1, np.abs(ref).max()))
    np.testing.assert_allclose(Vh.T @ Vh, np.eye(nev), atol=1e-11)
    np.testing.assert_allclose(A @ Vh, Vh * e, atol=1e-10)


@pytest.mark.parametrize("l,nev", [(300, 20), (700, 20), (1100, 24), (2100, 8)])
def test_syevx_clustered_wanted_roots(l, nev, rng):
    """Subspace-like spectrum: the wanted roots lie in one cluster far below
    ||T|| (all of them within the 1e-3 ||T|| re-orthogonalisation window), with
    an exact degeneracy; covers the register and the workgroup back-transforms."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    Q, _ = np.linalg.qr(rng.standard_normal((l, l)))
    ev = np.concatenate([0.01 + 2e-4 * np.arange(30), np.linspace(0.5, 40.0, l - 30)])
    ev[4] = ev[5]
    ev = np.sort(ev)
    A = (Q * ev) @ Q.T
    A = 0.5 * (A + A.T)
    Ad = T(A)
    e_d = torch.empty(nev, dtype=torch.float64, device=_dev())
    V = torch.empty((l, nev), dtype=torch.float64, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_syevx_workspace(l, nev))
    _lib.call("dftb_syevx", _p(Ad), l, l, nev, _p(e_d), _p(V), _p(w), wn, _stream())
    e = e_d.cpu().numpy()
    Vh = V.cpu().numpy()
    ref = np.linalg.eigvalsh(A)[:nev]
    np.testing.assert_allclose(e, ref, atol=1e-12 * np.abs(ev).max())
    np.testing.assert_allclose(Vh.T @ Vh, np.eye(nev), atol=1e-11)
    np.testing.assert_allclose(A @ Vh, Vh * e, atol=1e-10 * np.abs(ev).max())


@pytest.mark.parametrize("layout", ["two_large", "isolated_then_large", "all_roots"])
def test_syevx_several_clusters(layout, rng):
    """Wanted roots in more than one re-orthogonalisation cluster (two clusters of
    10 members, isolated roots before a cluster, every root of a spread spectrum):
    every eigenvector is a unit residual-free vector, orthogonal to the others."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    l = 200
    if layout == "two_large":
        ev, nev = np.concatenate([0.01 + 1e-3 * np.arange(10), 0.6 + 1e-3 * np.arange(10),
                                  np.linspace(1.5, 40.0, l - 20)]), 24
    elif layout == "isolated_then_large":
        ev, nev = np.concatenate([[0.001, 0.2, 0.4], 0.6 + 1e-3 * np.arange(12), np.linspace(1.5, 40.0, l - 15)]), 24
    else:
        ev, nev = np.sort(np.concatenate([rng.uniform(0.0, 0.5, 60), rng.uniform(2.0, 3.0, l - 60)])), l
    Q, _ = np.linalg.qr(rng.standard_normal((l, l)))
    A = (Q * ev) @ Q.T
    A = 0.5 * (A + A.T)
    e_d = torch.empty(nev, dtype=torch.float64, device=_dev())
    V = torch.empty((l, nev), dtype=torch.float64, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_syevx_workspace(l, nev))
    _lib.call("dftb_syevx", _p(T(A)), l, l, nev, _p(e_d), _p(V), _p(w), wn, _stream())
    e, Vh = e_d.cpu().numpy(), V.cpu().numpy()
    np.testing.assert_allclose(e, np.sort(ev)[:nev], atol=1e-12 * np.abs(ev).max())
    np.testing.assert_allclose(Vh.T @ Vh, np.eye(nev), atol=1e-11)
    np.testing.assert_allclose(A @ Vh, Vh * e, atol=1e-10 * np.abs(ev).max())


@pytest.mark.parametrize("l,nev", [(40, 8), (150, 20), (333, 20)])
def test_subspace_cholesky_route_equals_sqrtm_route(l, nev, rng):
    """dftb_subspace_nonhermitian_rl (Cholesky factor of Mm) and dftb_subspace_tb
    give the roots, Rb, Lb and Tb of dftb_subspace_nonhermitian (square root of
    Mm, Solver.py:759-798), columns up to a common sign; an indefinite Mm is
    reported as not positive definite."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace

    def spd(lo, hi):
        Q, _ = np.linalg.qr(rng.standard_normal((l, l)))
        A = (Q * np.linspace(lo, hi, l)) @ Q.T
        return 0.5 * (A + A.T)

    Mp, Mm = spd(0.05, 3.0), spd(0.02, 2.0)
    ld = l + 5
    Mpd = torch.zeros((l, ld), dtype=torch.float64, device=_dev())
    Mmd = torch.zeros((l, ld), dtype=torch.float64, device=_dev())
    Mpd[:, :l] = T(Mp)
    Mmd[:, :l] = T(Mm)
    w, wn = Workspace(_dev()).get(_lib.load().dftb_subspace_workspace(l, nev))
    out = {}
    for route in ("sqrtm", "chol"):
        wd = torch.zeros(nev, dtype=torch.float64, device=_dev())
        Tb, Rb, Lb = (torch.zeros((l, nev), dtype=torch.float64, device=_dev()) for _ in range(3))
        info = torch.zeros(2, dtype=torch.int32, device=_dev())
        diag = torch.zeros(2, dtype=torch.float64, device=_dev())
        if route == "sqrtm":
            _lib.call("dftb_subspace_nonhermitian", _p(Mpd), _p(Mmd), l, ld, nev, _p(wd), _p(Tb), _p(Rb), _p(Lb),
                      _p(info), _p(diag), _p(w), wn, _stream())
        else:
            _lib.call("dftb_subspace_nonhermitian_rl", _p(Mpd), _p(Mmd), l, ld, nev, _p(wd), _p(Rb), _p(Lb),
                      _p(info), _p(diag), _p(w), wn, _stream())
            _lib.call("dftb_subspace_tb", _p(Mmd), l, ld, nev, _p(wd), _p(Rb), _p(Tb), _p(w), wn, _stream())
        assert list(info.cpu().numpy()) == [0, 0]
        assert diag.cpu().numpy()[1] < 1e-10
        out[route] = [x.cpu().numpy() for x in (wd, Tb, Rb, Lb)]
    ws_, Tq, Rq, Lq = out["sqrtm"]
    wc, Tc, Rc, Lc = out["chol"]
    ref = np.sqrt(np.sort(np.linalg.eigvals(Mp @ Mm).real)[:nev])
    np.testing.assert_allclose(ws_, ref, rtol=1e-12)
    np.testing.assert_allclose(wc, ref, rtol=1e-12)
    sgn = np.sign(np.sum(Tq * Tc, axis=0))
    for a, b in ((Tq, Tc), (Rq, Rc), (Lq, Lc)):
        np.testing.assert_allclose(a, b * sgn, atol=1e-10 * np.abs(a).max())
    # Tb orthonormal eigenvectors of Mh = Mm^1/2 Mp Mm^1/2
    np.testing.assert_allclose(Tc.T @ Tc, np.eye(nev), atol=1e-11)
    # indefinite Mm
    Mmd[:, :l] = T(Mm - 0.5 * np.eye(l))
    info = torch.zeros(2, dtype=torch.int32, device=_dev())
    diag = torch.zeros(2, dtype=torch.float64, device=_dev())
    Rb, Lb = (torch.zeros((l, nev), dtype=torch.float64, device=_dev()) for _ in range(2))
    wd = torch.zeros(nev, dtype=torch.float64, device=_dev())
    _lib.call("dftb_subspace_nonhermitian_rl", _p(Mpd), _p(Mmd), l, ld, nev, _p(wd), _p(Rb), _p(Lb),
              _p(info), _p(diag), _p(w), wn, _stream())
    assert info.cpu().numpy()[0] == _lib.DFTB_ERR_NOT_SPD


def test_sqrtm(rng):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    l = 80
    Q, _ = np.linalg.qr(rng.standard_normal((l, l)))
    A = (Q * np.linspace(0.02, 3.0, l)) @ Q.T
    Ad = T(A)
    S = torch.empty_like(Ad)
    Si = torch.empty_like(Ad)
    it = (ctypes.c_int * 1)()
    w, wn = Workspace(_dev()).get(_lib.load().dftb_sqrtm_workspace(l))
    err = torch.zeros(1, dtype=torch.float64)
    _lib.call("dftb_sqrtm", _p(Ad), l, l, _p(S), _p(Si), ctypes.cast(it, ctypes.c_void_p),
              ctypes.c_void_p(err.data_ptr()), _p(w), wn, _stream())
    ref = np.real(sla.sqrtm(A))
    np.testing.assert_allclose(S.cpu().numpy(), ref, atol=1e-12)
    np.testing.assert_allclose(S.cpu().numpy() @ Si.cpu().numpy(), np.eye(l), atol=1e-11)


def test_orthonormalize(rng):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    n, l, m = 5003, 7, 5
    B, _ = np.linalg.qr(rng.standard_normal((n, l)))
    Qh = rng.standard_normal((m, n))
    Qh[3] = Qh[1] + B[:, 2]           # dependent after projection on span(B) + Q1
    Bd = T(B.T.copy())
    Qd = T(Qh)
    keep = torch.empty(m, dtype=torch.int32, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_orthonormalize_workspace(l, m, n))
    _lib.call("dftb_orthonormalize", _p(Bd), l, _p(Qd), m, n, 1e-8, _p(keep), _p(w), wn, _stream())
    k = keep.cpu().numpy()
    assert list(k) == [1, 1, 1, 0, 1]
    Qr = Qd.cpu().numpy()[k == 1]
    np.testing.assert_allclose(Qr @ Qr.T, np.eye(4), atol=1e-12)
    np.testing.assert_allclose(Qr @ B, 0, atol=1e-12)
    # span preserved: original kept rows lie in span(B, Qr)
    basis = np.vstack([B.T, Qr])
    for row in Qh[k == 1]:
        res = row - basis.T @ (basis @ row)
        assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(row)


def test_orthonormalize_without_basis(rng):
    """l = 0 (a restart from given vectors): Cholesky-QR of the block alone."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    n, m = 9001, 12
    Qh = rng.standard_normal((m, n))
    Qh[4] = 2.0 * Qh[1] - Qh[0]
    Qd = T(Qh)
    keep = torch.empty(m, dtype=torch.int32, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_orthonormalize_workspace(0, m, n))
    _lib.call("dftb_orthonormalize", None, 0, _p(Qd), m, n, 1e-8, _p(keep), _p(w), wn, _stream())
    k = keep.cpu().numpy()
    assert list(k) == [1, 1, 1, 1, 0] + [1] * (m - 5)
    Qr = Qd.cpu().numpy()[k == 1]
    np.testing.assert_allclose(Qr @ Qr.T, np.eye(m - 1), atol=1e-12)
    for row in Qh[k == 1]:
        res = row - Qr.T @ (Qr @ row)
        assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(row)


@pytest.mark.parametrize("m", [40, 60])
def test_orthonormalize_block(m, rng):
    """Blocks of the Davidson's size (the LDS Cholesky-QR and the register
    triangular update, m <= 48) and beyond (global-memory factor, GEMM update)."""
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream, Workspace
    n, l = 20011, 100
    B, _ = np.linalg.qr(rng.standard_normal((n, l)))
    Qh = rng.standard_normal((m, n))
    Qh[7] = 0.5 * Qh[2] - Qh[5] + B[:, 9]      # dependent after the projection
    Qh[m - 1] = B[:, 3] * 1e-3                  # inside span(B)
    Bd = T(B.T.copy())
    Qd = T(Qh)
    keep = torch.empty(m, dtype=torch.int32, device=_dev())
    w, wn = Workspace(_dev()).get(_lib.load().dftb_orthonormalize_workspace(l, m, n))
    _lib.call("dftb_orthonormalize", _p(Bd), l, _p(Qd), m, n, 1e-8, _p(keep), _p(w), wn, _stream())
    k = keep.cpu().numpy()
    want = np.ones(m, dtype=int)
    want[7] = want[m - 1] = 0
    assert list(k) == list(want)
    Qr = Qd.cpu().numpy()[k == 1]
    np.testing.assert_allclose(Qr @ Qr.T, np.eye(m - 2), atol=1e-12)
    np.testing.assert_allclose(Qr @ B, 0, atol=1e-12)
    assert np.abs(Qd.cpu().numpy()[k == 0]).max() == 0.0
    basis = np.vstack([B.T, Qr])
    for row in Qh[k == 1]:
        res = row - basis.T @ (basis @ row)
        assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(row)


def test_symmetrize(rng):
    from dftbaby_amd import _lib
    from dftbaby_amd.engine import _p, _stream
    l, ld = 300, 310
    A = rng.standard_normal((l, l))
    S = A + A.T
    E = 1e-6 * rng.standard_normal((l, l))
    M = np.zeros((l, ld))
    M[:, :l] = S + E
    M[:, l:] = 7.0                                 # beyond the leading block: untouched
    Md = T(M)
    asym = torch.full((3,), 5.0, dtype=torch.float64, device=_dev())
    _lib.call("dftb_symmetrize", _p(Md), l, ld, _p(asym), _stream())
    out = Md.cpu().numpy()
    want = 0.5 * ((S + E) + (S + E).T)
    np.testing.assert_array_equal(out[:, :l], out[:, :l].T)
    np.testing.assert_allclose(out[:, :l], want, rtol=0, atol=1e-15 * np.abs(S).max())
    assert (out[:, l:] == 7.0).all()
    ref = np.abs((S + E) - (S + E).T).max() / np.abs(S + E).max()
    assert abs(asym.cpu().numpy()[0] - ref) <= 1e-15 * ref


def test_autotune_ao_tiles_are_used(tmp_path):
    """The project autotuner times the BF16x3 AO chain's pipelined tiles on full
    block matvecs, stores its pick in op.tuning (passed to every later matvec)
    and the cache; an operator of the same size takes it from the cache without
    measuring; tuned products equal the FP64 products to the BF16x3 tolerance;
    LR_TDDFTB(autotune) runs it before the Davidson."""
    from dftbaby_amd.autotune import autotune
    from dftbaby_amd.LR_TDDFTB import LR_TDDFTB
    op, _ = _operator("C180H36", 1, exchange="ao")
    cache = str(tmp_path / "tuning.json")
    lines = []
    cfg = autotune(op, 2, nvec=8, cache=cache, log=lines.append)
    assert len(lines) == 4 and cfg["tile_expand"] in (256, 512) and cfg["tile_contract"] in (256, 512)
    assert op.tuning[2] == cfg
    s = op._struct(2)
    assert (s.tile_expand, s.tile_contract) == (cfg["tile_expand"], cfg["tile_contract"])
    op2, _ = _operator("C180H36", 1, exchange="ao")
    lines2 = []
    assert autotune(op2, 2, nvec=8, cache=cache, log=lines2.append) == cfg and not lines2
    V = torch.zeros((6, op.N), dtype=torch.float64, device=_dev())
    V.view(6, op.opad, op.vpad)[:, :op.nocc, :op.nvirt] = torch.randn((6, op.nocc, op.nvirt), dtype=torch.float64,
                                                                      device=_dev())
    P2, M2 = op.matvec(V, 2)
    P0, M0 = op.matvec(V, 0)
    for a, b in ((P2, P0), (M2, M0)):
        assert float((a - b).abs().max()) < 3e-5 * float(b.abs().max())
    td = LR_TDDFTB(ground_state=ground_state("C12H14", 1, 0.05, 3))
    td.autotune = True
    td.getEnergies(nstates=4, diag_conv=1e-6, diag_maxiter=100, precision="bf16x3")
    assert td.op.tuning.get(2)
