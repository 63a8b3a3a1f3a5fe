"""GPU: the mixed-precision path (complex64 solve passes on the matrix cores,
FP64 residual and iterative refinement) against the FP64 oracle.
Tolerance: rel. l2 <= 1e-8 per RHS after refinement (the mixed-path bound)."""
import numpy as np
import pytest

import robinsolver_amd as rsa
from oracle.pyoracle import OracleTree

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]
TOL_MIXED = 1e-8


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _k(x, kb, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, 8)
    p = rng.integers(1, 9, size=(8, 2))
    ph = rng.uniform(0, 2 * np.pi, 8)
    q = sum(a[j] * np.cos(2 * np.pi * (x[..., :2] @ p[j]) + ph[j]) for j in range(8)) / np.abs(a).sum()
    return kb ** 2 * (1 + 0.3 * q) + 0.01j * kb ** 2


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("d,n,L,kb,n_refine", [(2, 64, 4, 10.0, 2), (2, 128, 6, 30.0, 2), (3, 16, 6, 4.0, 2),
                                                 (2, 64, 4, 10.0, 0),
                                                 # 2D leaves of 8 x 16 points (the generic leaf kernel)
                                                 (2, 64, 5, 10.0, 2),
                                                 # 3D leaves of 8^3 (64-point layers) and 6^3 (36, padded)
                                                 (3, 16, 3, 4.0, 2), (3, 12, 3, 4.0, 2), (3, 16, 3, 4.0, 0)])
def test_mixed_matches_oracle(d, n, L, kb, n_refine, gpu):
    print("case: mixed", d, n, L, kb, n_refine, flush=True)
    o = OracleTree(d, L, UNIT, 1.0 / n, -kb)
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=n_refine)
    pts, face = o.all_points()
    k = _k(pts[:, o.nb:], kb)
    o.factor(k)
    t.factor(k)
    rng = np.random.default_rng(3)
    nrhs = 4
    f = _rand_c(rng, (nrhs, o.n_leaves, o.n_int))
    g = _rand_c(rng, (nrhs, o.n_leaves, o.nb))
    u = t.solve(f, g)
    rels = [_rel(u[r], o.solve(f[r], g[r])) for r in range(nrhs)]
    print("rel l2:", rels, flush=True)
    if n_refine == 0:
        assert max(rels) < 1e-3      # a complex64 solve alone
    else:
        assert max(rels) <= TOL_MIXED, rels


@pytest.mark.parametrize("n,L", [(64, 4), (56, 4), (48, 4)])   # 16x16, 14x14, 12x12 leaves
def test_mixed_wide_batch_matches_oracle(n, L, gpu):
    """More RHS than one 64-column leaf tile, on each leaf size of the
    matrix-core leaf kernel."""
    print("case: mixed wide batch", n, L, flush=True)
    kb = 12.0
    o = OracleTree(2, L, UNIT, 1.0 / n, -kb)
    t = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=1)
    pts, face = o.all_points()
    k = _k(pts[:, o.nb:], kb)
    o.factor(k)
    t.factor(k)
    rng = np.random.default_rng(8)
    nrhs = 70
    f = _rand_c(rng, (nrhs, o.n_leaves, o.n_int))
    g = _rand_c(rng, (nrhs, o.n_leaves, o.nb))
    u = t.solve(f, g)
    for r in (0, 33, 64, 69):
        rel = _rel(u[r], o.solve(f[r], g[r]))
        assert rel <= TOL_MIXED, (r, rel)


def test_mixed_batch_invariance(gpu):
    print("case: mixed batch invariance", flush=True)
    t = rsa.GpuTree(2, 4, UNIT, 1.0 / 64, -10.0, precision=rsa.RS_MIXED)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], 10.0))
    rng = np.random.default_rng(4)
    nrhs = 20
    f = _rand_c(rng, (nrhs, t.n_leaves, t.n_int))
    g = _rand_c(rng, (nrhs, t.n_leaves, t.nb))
    u = t.solve(f, g)
    for r in (0, 11, 19):
        assert np.array_equal(t.solve(f[r], g[r]), u[r])



@pytest.mark.parametrize("fp64_levels", [0, 2])
def test_complex64_merges_and_precision_rules(fp64_levels, gpu):
    """merge_fp64_levels = k >= 0: the merges below the top k levels run in
    complex64, so the RtR maps differ from an FP64 tree's by complex64 rounding
    (not more); such a tree refuses FP64 solves and its refined mixed solves
    reach the mixed-path bound.  A mixed tree that factors in FP64
    (merge_fp64_levels -1): its RtR maps are bitwise an FP64 tree's; it keeps
    complex64 solve data only and refuses FP64 solves, unless it keeps the FP64
    data too (RS_RETAIN_FP64), and then its FP64 solves are bitwise an FP64
    tree's."""
    print("case: complex64 merges, fp64 levels", fp64_levels, flush=True)
    kb, n, L = 12.0, 64, 4
    o = OracleTree(2, L, UNIT, 1.0 / n, -kb)
    maps = rsa.RS_RETAIN_MAPS
    tm = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=2, merge_fp64_levels=fp64_levels,
                     retain=maps)
    tdef = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=1,
                       retain=maps | rsa.RS_RETAIN_FP64, merge_fp64_levels=-1)
    tplain = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=1, merge_fp64_levels=-1)
    td = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_FP64, retain=maps)
    pts, _ = o.all_points()
    k = _k(pts[:, o.nb:], kb)
    for t in (o, tm, tdef, tplain, td):
        t.factor(k)
    for lvl, idx in ((L - 1, 0), (1, 1)):
        a, b = tm.node_rtr(lvl, idx), td.node_rtr(lvl, idx)
        d = np.linalg.norm(a - b) / np.linalg.norm(b)
        print("RtR level", lvl, "complex64 merges vs FP64:", d, flush=True)
        assert 1e-12 < d < 1e-4, (lvl, d)
        assert np.array_equal(tdef.node_rtr(lvl, idx), b)
    rng = np.random.default_rng(12)
    f = _rand_c(rng, (2, o.n_leaves, o.n_int))
    g = _rand_c(rng, (2, o.n_leaves, o.nb))
    with pytest.raises(rsa.RobinError, match="FP64 solve"):
        tm.solve(f, g, precision=rsa.RS_FP64)
    with pytest.raises(rsa.RobinError, match="complex64 solve data only"):
        tplain.solve(f, g, precision=rsa.RS_FP64)
    assert np.array_equal(tplain.solve(f, g), tdef.solve(f, g))
    assert np.array_equal(tdef.solve(f, g, precision=rsa.RS_FP64), td.solve(f, g))
    for u in (tm.solve(f, g), td.solve(f, g, precision=rsa.RS_MIXED)):
        for r in range(2):
            assert _rel(u[r], o.solve(f[r], g[r])) <= TOL_MIXED


@pytest.mark.parametrize("n,L", [(16, 3), (16, 6)])   # 8^3 and 4^3 leaves
def test_mixed_3d_batch_invariance(n, L, gpu):
    """3D leaves: more columns than one workgroup's 16; each column alone gives
    the bits it gets in the batch."""
    print("case: mixed 3d batch invariance", n, L, flush=True)
    t = rsa.GpuTree(3, L, UNIT, 1.0 / n, -4.0, precision=rsa.RS_MIXED, n_refine=1)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], 4.0))
    rng = np.random.default_rng(14)
    nrhs = 37
    f = _rand_c(rng, (nrhs, t.n_leaves, t.n_int))
    g = _rand_c(rng, (nrhs, t.n_leaves, t.nb))
    u = t.solve(f, g)
    for r in (0, 16, 36):
        assert np.array_equal(t.solve(f[r], g[r]), u[r])
