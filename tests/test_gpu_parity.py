"""GPU parity: the MI355X factor and solve against the CPU oracle, through the
C ABI.  FP64 path: rel. l2 <= 1e-10 per RHS (tolerance of the FP64 GPU path);
intermediate RtR maps to 1e-10 relative.  The mixed path is checked in
test_gpu_mixed.py."""
import sys

import numpy as np
import pytest

import robinsolver_amd as rsa
from oracle.pyoracle import OracleTree
from tests import fdsystem as fds

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]
TOL64 = 1e-10


def _case(tag):
    print("case:", tag, flush=True)
    sys.stdout.flush()


def _rng(seed):
    return np.random.default_rng(seed)


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _k(pts_int, kb):
    x = pts_int
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1) * np.cos(2 * x[..., 1])) + 0.05j * kb ** 2


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CASES = [
    (2, 16, 2, 5.0, None),
    (2, 64, 4, 10.0, None),        # cfg1-sized: 64^2, L=4, 16x16 leaves
    (2, 24, 3, 7.0, None),
    (2, 32, 0, 3.0, None),         # a single leaf (root = leaf)
    (2, 48, 1, 4.0, None),
    (2, 128, 6, 20.0, None),
    (3, 8, 3, 3.0, None),
    (3, 16, 6, 4.0, None),
    (2, 16, 3, 5.0, np.array([[1.3j * 5, 0.7], [-0.8j * 5, 1.1]])),
    (3, 8, 2, 3.0, np.array([[2.0 + 1j, 1.0], [0.5, 1.0 - 0.5j]])),
]


def _pair(d, n, L, kb, bp, skip_root=True):
    o = OracleTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp)
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp, skip_root_rtr=skip_root, retain=rsa.RS_RETAIN_MAPS)
    pts, face = o.all_points()
    k = _k(pts[:, o.nb:], kb)
    o.factor(k)
    t.factor(k)
    return o, t, pts, face


@pytest.mark.parametrize("d,n,L,kb,bp", CASES)
def test_fp64_solve_matches_oracle(d, n, L, kb, bp, gpu):
    """Both the tree that keeps its full maps and a tree with the default
    retention, whose level 1 forms its maps only where the root merge reads
    them (the root interface's rows and columns)."""
    _case(("fp64_solve", d, n, L, kb, bp is not None))
    o, t, pts, face = _pair(d, n, L, kb, bp)
    t2 = rsa.GpuTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp)
    t2.factor(_k(pts[:, o.nb:], kb))
    rng = _rng(11 + n)
    nrhs = 3
    f = _rand_c(rng, (nrhs, o.n_leaves, o.n_int))
    g = _rand_c(rng, (nrhs, o.n_leaves, o.nb))
    u = t.solve(f, g)
    u2 = t2.solve(f, g)
    # one column alone (the one-RHS leaf passes: layers of 6 points at 24^2, L = 3)
    u1 = t.solve(f[1:2], g[1:2])
    for r in range(nrhs):
        u_ref = o.solve(f[r], g[r])
        rel = _rel(u[r], u_ref)
        assert rel <= TOL64, (r, rel)
        rel2 = _rel(u2[r], u_ref)
        assert rel2 <= TOL64, (r, rel2)
        if r == 1:
            assert _rel(u1[0], u_ref) <= TOL64
            if d == 2:
                assert np.array_equal(u1[0], u[1])


@pytest.mark.parametrize("d,n,L,kb,bp", [CASES[0], CASES[2], CASES[4], CASES[6], CASES[8], CASES[9]])
def test_fp64_rtr_maps_match_oracle(d, n, L, kb, bp, gpu):
    _case(("fp64_rtr", d, n, L, kb))
    o, t, _, _ = _pair(d, n, L, kb, bp, skip_root=False)
    for lvl in range(L, -1, -1):
        for idx in sorted({0, (1 << lvl) - 1, min(1, (1 << lvl) - 1)}):
            T_gpu = t.node_rtr(lvl, idx)
            T_ref = o.node_rtr(o.level_node_id(lvl, idx))
            assert _rel(T_gpu, T_ref) <= TOL64, (lvl, idx, _rel(T_gpu, T_ref))


def test_fp64_manufactured_2d(gpu):
    _case("manufactured_2d")
    errs = []
    for n in (32, 64):
        t = rsa.GpuTree(2, 2, UNIT, 1.0 / n, 1.0)
        pts, face = t.all_points()
        t.factor(np.ones((t.n_leaves, t.n_int)))
        g = fds.sample_g(pts, face, 2, fds.g2d)
        u = t.solve(np.zeros((t.n_leaves, t.n_int)), g)
        errs.append(_rel(u, fds.true_sol2d(pts)))
    assert errs[-1] < 1e-4 and 3.2 < errs[0] / errs[1] < 4.8, errs


def test_batch_composition_is_bitwise_invariant(gpu):
    """A right-hand side gets the same bits alone and inside a batch."""
    _case("batch_invariance")
    o, t, _, _ = _pair(2, 64, 4, 10.0, None)
    rng = _rng(5)
    nrhs = 37
    f = _rand_c(rng, (nrhs, t.n_leaves, t.n_int))
    g = _rand_c(rng, (nrhs, t.n_leaves, t.nb))
    u_all = t.solve(f, g)
    for r in (0, 17, 36):
        u1 = t.solve(f[r], g[r])
        assert np.array_equal(u1, u_all[r]), r
    # three columns: the few-column leaf and product kernels (one wave per leaf,
    # matrix-core-free products) against the wide ones
    assert np.array_equal(t.solve(f[5:8], g[5:8]), u_all[5:8])


def test_repeated_solves_and_refactor(gpu):
    _case("repeat")
    o, t, pts, _ = _pair(2, 32, 2, 6.0, None)
    rng = _rng(9)
    f = _rand_c(rng, (t.n_leaves, t.n_int))
    g = _rand_c(rng, (t.n_leaves, t.nb))
    u1 = t.solve(f, g)
    u2 = t.solve(f, g)
    assert np.array_equal(u1, u2)
    k2 = _k(pts[:, t.nb:], 6.0) * 1.1
    t.factor(k2)
    o.factor(k2)
    assert _rel(t.solve(f, g), o.solve(f, g)) <= TOL64
