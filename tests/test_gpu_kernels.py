"""GPU: the kernels in isolation against numpy -- the batched MFMA GEMM
(complex128 and complex64 paths, tile edges, B split, row scatter, beta) and
the pivoted batched inverse (whole-matrix LDS path and blocked panel path)."""
import ctypes

import numpy as np
import pytest

import robinsolver_amd as rsa

pytestmark = pytest.mark.gpu

N_TILES = 16   # GEMM tile choices of rs_check_gemm_tile (0 .. 15)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _lib():
    L = rsa.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.rs_check_gemm.argtypes = [ctypes.c_int] * 4 + [dp, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, dp,
                                                     ctypes.c_int, ip, dp, dp]
    L.rs_check_inverse.argtypes = [ctypes.c_int, ctypes.c_int, dp]
    L.rs_last_error.restype = ctypes.c_char_p
    return L


def _rc(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _run_gemm(prec, A, B, C, alpha, beta, B2=None, ksplit=None, rmap=None, ldc=None):
    L = _lib()
    M, K = A.shape
    N = B.shape[1]
    Af = np.asfortranarray(A).astype(np.complex128)
    Bf = np.asfortranarray(B).astype(np.complex128)
    B2f = None if B2 is None else np.asfortranarray(B2).astype(np.complex128)
    Cf = np.asfortranarray(C).astype(np.complex128)
    al = np.array([alpha.real, alpha.imag])
    be = np.array([beta.real, beta.imag])
    rm = None
    if rmap is not None:
        rm = np.ascontiguousarray(rmap, dtype=np.int32)
    a_flat = Af.reshape(-1, order="F").copy()
    b_flat = Bf.reshape(-1, order="F").copy()
    b2_flat = None if B2f is None else B2f.reshape(-1, order="F").copy()
    c_flat = Cf.reshape(-1, order="F").copy()
    rc = L.rs_check_gemm(prec, M, N, K, _dp(a_flat.view(np.float64)), M, _dp(b_flat.view(np.float64)),
                         None if b2_flat is None else _dp(b2_flat.view(np.float64)), Bf.shape[0],
                         ksplit if ksplit is not None else K, _dp(c_flat.view(np.float64)), Cf.shape[0],
                         None if rm is None else rm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(al), _dp(be))
    assert rc == 0, L.rs_last_error(None)
    return c_flat.reshape(Cf.shape, order="F")


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED, rsa.RS_MIXED_MERGE])
@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (64, 64, 16), (37, 5, 23), (130, 200, 70), (48, 1, 96),
                                   (7, 65, 3), (64, 20, 2048), (200, 1, 1000), (100, 70, 777), (16, 40, 64), (30, 100, 50), (12, 300, 2000)])
def test_gemm_matches_numpy(prec, M, N, K, gpu):
    print("case: gemm", prec, M, N, K, flush=True)
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A, B, C = _rc(rng, M, K), _rc(rng, K, N), _rc(rng, M, N)
    alpha, beta = 0.7 - 0.3j, 1.1 + 0.2j
    out = _run_gemm(prec, A, B, C, alpha, beta)
    ref = alpha * (A @ B) + beta * C
    tol = 1e-13 if prec == rsa.RS_FP64 else 2e-6
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED])
def test_gemm_column_results_do_not_depend_on_width(prec, gpu):
    """Each output column gets the same bits whatever the number of columns
    (the K split and the tile width are chosen without looking at N)."""
    rng = np.random.default_rng(3)
    M, K = 96, 1536
    A, B = _rc(rng, M, K), _rc(rng, K, 80)
    wide = _run_gemm(prec, A, B, np.zeros((M, 80), complex), 1.0 + 0j, 0j)
    for c in (0, 17, 79):
        one = _run_gemm(prec, A, B[:, c:c + 1], np.zeros((M, 1), complex), 1.0 + 0j, 0j)
        assert np.array_equal(one[:, 0], wide[:, c])


@pytest.mark.parametrize("M,N,K", [(96, 40, 1536), (64, 20, 2048), (37, 5, 23)])
def test_mixed_merge_gemm_is_complex64_products_with_complex128_result(M, N, K, gpu):
    """The merge form of the mixed factorization: with alpha = 1, beta = 0 its
    result is bitwise the complex64 GEMM's (same operand rounding, same FP32
    accumulation sequence, K split or not); its error against complex128 is
    that of complex64 products; and its epilogue keeps complex128 (alpha = 0,
    beta = 1 returns C bit for bit, which complex64 storage cannot)."""
    print("case: mixed merge gemm", M, N, K, flush=True)
    rng = np.random.default_rng(M + 7 * N + K)
    A, B, C = _rc(rng, M, K), _rc(rng, K, N), _rc(rng, M, N)
    merge = _run_gemm(rsa.RS_MIXED_MERGE, A, B, np.zeros((M, N), complex), 1.0 + 0j, 0j)
    c64 = _run_gemm(rsa.RS_MIXED, A, B, np.zeros((M, N), complex), 1.0 + 0j, 0j)
    assert np.array_equal(merge, c64)
    exact = A @ B
    err = np.linalg.norm(merge - exact) / np.linalg.norm(exact)
    assert 1e-9 < err < 2e-6, err
    keep = _run_gemm(rsa.RS_MIXED_MERGE, A, B, C, 0j, 1.0 + 0j)
    assert np.array_equal(keep, C)
    rounded = _run_gemm(rsa.RS_MIXED, A, B, C, 0j, 1.0 + 0j)
    assert not np.array_equal(rounded, C)


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED])
@pytest.mark.parametrize("N", [20, 3])
def test_gemm_split_and_scatter(prec, N, gpu):
    print("case: gemm split/scatter", prec, N, flush=True)
    rng = np.random.default_rng(1)
    M, K, ks = 40, 30, 12
    A = _rc(rng, M, K)
    B1, B2 = _rc(rng, K, N), _rc(rng, K, N)
    Bfull = np.vstack([B1[:ks], B2[:K - ks]])
    rmap = rng.permutation(M + 9)[:M]
    C = _rc(rng, M + 9, N)
    out = _run_gemm(prec, A, B1, C, 1.0 + 0j, 0.5 + 0j, B2=B2, ksplit=ks, rmap=rmap)
    ref = C.copy()
    ref[rmap] = A @ Bfull + 0.5 * C[rmap]
    tol = 1e-13 if prec == rsa.RS_FP64 else 2e-6
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("n,batch", [(16, 5), (48, 3), (64, 2), (128, 2), (200, 1), (202, 2), (512, 2), (1004, 1),
                                     (1024, 2), (1100, 2), (1536, 1), (2048, 1), (3072, 1), (4096, 1), (4100, 1)])
def test_inverse_matches_numpy(n, batch, gpu):
    """Up to 4096 points the panels live in registers (4096: the largest row the
    pivot key's 12-bit row field holds); above 1024 points the panels are grouped
    in blocks of 64 columns (1100: 4-column panels, a last block of 12 columns);
    4100 takes the LDS panel kernel with one-column panels and the
    one-thread-per-column interchanges."""
    print("case: inverse", n, batch, flush=True)
    rng = np.random.default_rng(n)
    mats = _rc(rng, batch, n, n) + 0.1 * np.eye(n)
    # permute rows so that partial pivoting has work to do
    for b in range(batch):
        mats[b] = mats[b][rng.permutation(n)]
    flat = np.concatenate([m.reshape(-1, order="F") for m in mats]).copy()
    L = _lib()
    assert L.rs_check_inverse(n, batch, _dp(flat.view(np.float64))) == 0
    for b in range(batch):
        inv = flat[b * n * n:(b + 1) * n * n].reshape(n, n, order="F")
        if n <= 3072:
            err = np.linalg.norm(inv @ mats[b] - np.eye(n)) / np.sqrt(n)
            assert err < 1e-10, (b, err)
        else:
            # products with a few vectors keep the check O(n^2)
            x = _rc(rng, 1, n, 4)[0]
            err = np.linalg.norm(inv @ (mats[b] @ x) - x) / np.linalg.norm(x)
            assert err < 1e-9, (b, err)


def test_inverse_refuses_sizes_no_kernel_takes(gpu):
    """Above ~5.1k points even a one-column LDS panel exceeds the 160 KB LDS: the
    entry point refuses the size before any allocation or launch."""
    n = 5200
    flat = np.zeros(n * n, np.complex128)
    L = _lib()
    assert L.rs_check_inverse(n, 1, _dp(flat.view(np.float64))) == 1
    assert b"too large" in L.rs_last_error(None)


@pytest.mark.parametrize("n", [40, 256])
def test_inverse_needs_pivoting(n, gpu):
    """A zero diagonal (a permutation plus a small perturbation): without row
    interchanges the first pivot would be 0."""
    print("case: inverse pivoting", n, flush=True)
    rng = np.random.default_rng(n + 1)
    perm = np.roll(np.arange(n), 1)
    M = np.zeros((n, n), np.complex128)
    M[np.arange(n), perm] = 1.0 + 0.5j
    M += 1e-3 * _rc(rng, 1, n, n)[0]
    np.fill_diagonal(M, 0.0)
    flat = M.reshape(-1, order="F").copy()
    assert _lib().rs_check_inverse(n, 1, _dp(flat.view(np.float64))) == 0
    inv = flat.reshape(n, n, order="F")
    assert np.linalg.norm(inv @ M - np.eye(n)) / np.sqrt(n) < 1e-10


@pytest.mark.parametrize("n", [64, 256, 1024])
def test_inverse_with_tied_pivots(n, gpu):
    """A Sylvester-Hadamard matrix times (1 + i): every candidate of the first
    column has the same |re|+|im|, and ties recur in later columns, within a wave
    and across the waves of the pivot search (the smallest row index wins, as
    zgetrf's idamax-style choice)."""
    print("case: inverse tied pivots", n, flush=True)
    H = np.array([[1.0]])
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    M = (1.0 + 1.0j) * H
    flat = M.reshape(-1, order="F").copy()
    assert _lib().rs_check_inverse(n, 1, _dp(flat.view(np.float64))) == 0
    inv = flat.reshape(n, n, order="F")
    assert np.linalg.norm(inv @ M - np.eye(n)) / np.sqrt(n) < 1e-12
    # the exact inverse is H^T / (n (1 + i))
    assert np.abs(inv - H.T / (n * (1.0 + 1.0j))).max() < 1e-12


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED, rsa.RS_MIXED_MERGE])
@pytest.mark.parametrize("M,N,K,scatter", [(48, 256, 16, True), (100, 70, 777, False), (200, 33, 1000, False),
                                           (16, 64, 32, True), (2048, 1, 2048, False), (130, 5, 77, True),
                                           (64, 16, 4096, True), (64, 1, 4096, True), (37, 3, 1001, False),
                                           (300, 4, 50, True)])
def test_gemm_tile_shapes_give_identical_bits(prec, M, N, K, scatter, gpu):
    """The autotuner may pick any tile shape: every shape must give the same bits."""
    print("case: gemm tiles", prec, M, N, K, scatter, flush=True)
    L = _lib()
    L.rs_check_gemm_tile.argtypes = L.rs_check_gemm.argtypes + [ctypes.c_int]
    rng = np.random.default_rng(M + N + K)
    A = _rc(rng, M, K)
    B = _rc(rng, K, N)
    ldc = M + 5 if scatter else M
    C0 = _rc(rng, ldc, N)
    rmap = np.ascontiguousarray(rng.permutation(ldc)[:M].astype(np.int32)) if scatter else None
    outs = []
    for tile in range(-1, N_TILES):
        C = np.asfortranarray(C0.copy())
        a = np.asfortranarray(A)
        b = np.asfortranarray(B)
        alpha = np.array([1.0, 0.5])
        beta = np.array([0.25, -1.0])
        rc = L.rs_check_gemm_tile(prec, M, N, K, _dp(a.reshape(-1, order="F").view(np.float64)), M,
                                  _dp(b.reshape(-1, order="F").view(np.float64)), None, K, K,
                                  _dp(C.reshape(-1, order="F").view(np.float64)), ldc,
                                  rmap.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if scatter else None,
                                  _dp(alpha), _dp(beta), tile)
        assert rc == 0
        outs.append(C.copy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED, rsa.RS_MIXED_MERGE])
@pytest.mark.parametrize("M,N,K,ksplit", [(128, 64, 96, 32), (96, 40, 80, 33), (64, 256, 512, 256)])
def test_gemm_tiles_with_split_b(prec, M, N, K, ksplit, gpu):
    """B's rows k >= ksplit from a second array (the two children's interface
    vectors): every tile gives the same bits, and the product is A [B; B2]
    (an odd ksplit moves the operands in 4-byte pieces on the direct-to-LDS tiles)."""
    print("case: gemm tiles split B", prec, M, N, K, ksplit, flush=True)
    L = _lib()
    L.rs_check_gemm_tile.argtypes = L.rs_check_gemm.argtypes + [ctypes.c_int]
    rng = np.random.default_rng(7 * M + K)
    A = np.asfortranarray(_rc(rng, M, K))
    B1 = np.asfortranarray(_rc(rng, K, N))
    B2 = np.asfortranarray(_rc(rng, K, N))
    C0 = _rc(rng, M, N)
    alpha = np.array([1.0, 0.5])
    beta = np.array([0.25, -1.0])
    outs = []
    for tile in range(-1, N_TILES):
        C = np.asfortranarray(C0.copy())
        rc = L.rs_check_gemm_tile(prec, M, N, K, _dp(A.reshape(-1, order="F").view(np.float64)), M,
                                  _dp(B1.reshape(-1, order="F").view(np.float64)),
                                  _dp(B2.reshape(-1, order="F").view(np.float64)), K, ksplit,
                                  _dp(C.reshape(-1, order="F").view(np.float64)), M, None, _dp(alpha), _dp(beta),
                                  tile)
        assert rc == 0
        outs.append(C.copy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    Bfull = np.vstack([B1[:ksplit], B2[:K - ksplit]])
    ref = (1.0 + 0.5j) * (A @ Bfull) + (0.25 - 1.0j) * C0
    tol = 1e-12 if prec == rsa.RS_FP64 else 1e-4
    assert np.linalg.norm(outs[0] - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED, rsa.RS_MIXED_MERGE])
@pytest.mark.parametrize("M,N,K,ldb,pairs", [(64, 64, 48, 100, False), (130, 70, 96, 96, False),
                                             (256, 256, 128, 300, False), (128, 256, 192, 400, True),
                                             (64, 96, 1100, 1200, True)])
def test_gemm_tiles_with_row_mapped_b(prec, M, N, K, ldb, pairs, gpu):
    """B's row k read at row bmap[k] (the down pass's incoming rows where the
    level above wrote them): every tile gives the same bits and the product
    A B[bmap].  With pairs, the map sends rows 2j, 2j + 1 to consecutive rows
    starting at an even one (as the tree's down-pass maps do), which the
    direct-to-LDS tiles read in 16-byte units; K = 1100 exceeds the map range
    they hold in LDS."""
    print("case: gemm tiles row-mapped B", prec, M, N, K, ldb, pairs, flush=True)
    L = _lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.rs_check_gemm_bmap.argtypes = [ctypes.c_int] * 4 + [dp, ctypes.c_int, dp, ctypes.c_int, ip, dp, ctypes.c_int,
                                                          dp, dp, ctypes.c_int]
    rng = np.random.default_rng(11 * M + K)
    A = np.asfortranarray(_rc(rng, M, K))
    B = np.asfortranarray(_rc(rng, ldb, N))
    if pairs:
        bmap = np.ascontiguousarray((2 * rng.permutation(ldb // 2)[:K // 2, None] + np.arange(2)).reshape(-1)
                                    .astype(np.int32))
    else:
        bmap = np.ascontiguousarray(rng.permutation(ldb)[:K].astype(np.int32))
    C0 = _rc(rng, M, N)
    alpha = np.array([1.0, 0.5])
    beta = np.array([0.25, -1.0])
    outs = []
    for tile in range(-1, N_TILES):
        C = np.asfortranarray(C0.copy())
        rc = L.rs_check_gemm_bmap(prec, M, N, K, _dp(A.reshape(-1, order="F").view(np.float64)), M,
                                  _dp(B.reshape(-1, order="F").view(np.float64)), ldb, bmap.ctypes.data_as(ip),
                                  _dp(C.reshape(-1, order="F").view(np.float64)), M, _dp(alpha), _dp(beta), tile)
        assert rc == 0
        outs.append(C.copy())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = (1.0 + 0.5j) * (A @ B[bmap]) + (0.25 - 1.0j) * C0
    tol = 1e-12 if prec == rsa.RS_FP64 else 1e-4
    assert np.linalg.norm(outs[0] - ref) / np.linalg.norm(ref) < tol


def test_autotuned_solve_matches_untuned_bits(gpu):
    print("case: autotuned solve bits", flush=True)
    import robinsolver_amd as rsa2
    res = []
    for tune in (True, False):
        t = rsa2.GpuTree(2, 6, [[0, 1], [0, 1], [0, 1]], 1.0 / 128, -12.0, precision=rsa2.RS_MIXED, n_refine=1,
                         autotune=tune)
        pts, _ = t.all_points()
        t.factor(144.0 * (1 + 0.2 * np.sin(3 * pts[:, t.nb:, 0])) + 1.0j)
        rng = np.random.default_rng(2)
        f = _rc(rng, 20, t.n_leaves, t.n_int)
        g = _rc(rng, 20, t.n_leaves, t.nb)
        res.append(t.solve(f, g))
        if tune:
            ch = t.tile_choices(rsa2.RS_MIXED)
            print("tiles:", ch.tolist(), flush=True)
            assert ch.shape == (6, 4) and (ch[1:] >= 0).all()
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("n", [40, 256, 1024])
def test_inverse_reports_singular_matrix(n, gpu):
    """A column with no nonzero candidate: the elimination makes no interchange
    (no out-of-range row index) and the call reports the singular matrix."""
    print("case: inverse singular", n, flush=True)
    rng = np.random.default_rng(n + 7)
    M = _rc(rng, 1, n, n)[0]
    M[:, n // 3] = 0.0
    flat = np.ascontiguousarray(np.tile(M.reshape(-1, order="F"), 2))
    assert _lib().rs_check_inverse(n, 2, _dp(flat.view(np.float64))) == 2


@pytest.mark.parametrize("d,L,h,domain", [(2, 4, 1 / 24, [[0, 1], [0, 1], [0, 1]]),
                                          (2, 5, 1 / 24, [[0.25, 1.5], [-1.0, 0.5], [0, 1]]),
                                          (3, 4, 1 / 12, [[0, 1], [0, 2], [-0.5, 0.5]])])
def test_points_device_match_host_points(d, L, h, domain, gpu):
    """rs_points_device writes the bits of rs_leaf_points for every leaf."""
    print("case: points device", d, L, h, flush=True)
    import torch
    t = rsa.GpuTree(d, L, domain, h, -5.0)
    pts, face = t.all_points()
    xyz = torch.zeros((t.n_leaves, t.m, 3), dtype=torch.float64, device="cuda:0")
    fc = torch.zeros((t.n_leaves, t.m), dtype=torch.int32, device="cuda:0")
    t.points_device(xyz.data_ptr(), fc.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(xyz.cpu().numpy(), pts)
    assert np.array_equal(fc.cpu().numpy()[0], face)
    assert np.array_equal(fc.cpu().numpy()[-1], face)


def test_factor_device_matches_host_factor(gpu):
    print("case: factor device", flush=True)
    import torch
    trees = [rsa.GpuTree(2, 4, [[0, 1], [0, 1], [0, 1]], 1.0 / 64, -10.0) for _ in range(2)]
    pts, _ = trees[0].all_points()
    x = pts[:, trees[0].nb:]
    k = 100.0 * (1 + 0.3 * np.sin(3 * x[..., 0]) * np.cos(2 * x[..., 1])) + 1j
    trees[0].factor(k)
    trees[1].factor_device(torch.from_numpy(np.ascontiguousarray(k)).to("cuda:0").data_ptr())
    rng = np.random.default_rng(2)
    f = rng.standard_normal((2, trees[0].n_leaves, trees[0].n_int)) + 0j
    g = np.zeros((2, trees[0].n_leaves, trees[0].nb), complex)
    assert np.array_equal(trees[0].solve(f, g), trees[1].solve(f, g))
