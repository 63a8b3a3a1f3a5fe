"""GPU: the Python mirror of the reference's Fortran API (robin_tree, factor,
solve, get_solution; robin_tree_module.f90:55-233, robin_output_module.f90:48-75)
gives the same bits as the C ABI driven directly with arrays sampled at the
same points, and samples the boundary callbacks at the physical faces only,
with the tangential coordinates packed first (fd_module.f90:645-651)."""
import numpy as np
import pytest

import robinsolver_amd as rsa

pytestmark = pytest.mark.gpu


def _k(x):
    return 25.0 * (1 + 0.2 * np.sin(3 * x[0] + x[1])) + 0.1j


def _f(x):
    return np.exp(-((x[0] - 0.4) ** 2 + (x[1] - 0.6) ** 2) / 0.01) + 0j


def _g_face(j):
    return lambda x: (1.0 + 0.3 * j) * np.exp(1j * 2.0 * x[0]) + 0.1j * x[1]


@pytest.mark.parametrize("prec", [rsa.RS_FP64, rsa.RS_MIXED])
def test_python_mirror_gives_the_c_abi_bits(prec, gpu):
    print("case: python mirror", prec, flush=True)
    kb, n, L = 5.0, 32, 3
    op = rsa.elliptic_operator(d=2, domain=[[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], k=_k)
    opts = rsa.solve_opts(h_tgt=1.0 / n, kb=-kb, precision=prec)
    T = rsa.robin_tree()
    T.bndry_params = np.array([[1j * opts.kb, 1.0], [-1j * opts.kb, 1.0]])
    calls = []
    g = [(lambda j: (lambda x: (calls.append(j), _g_face(j)(x))[1]))(j) for j in range(4)] + [None, None]
    T.factor(L, op, opts)
    T.solve(rsa.elliptic_rhs(d=2, f=_f, g=g))
    sol = rsa.get_solution(T)

    t = rsa.GpuTree(2, L, [[0, 1], [0, 1], [0, 1]], 1.0 / n, -kb, T.bndry_params, precision=prec)
    pts, face = t.all_points()
    k = np.array([[_k(pts[q, p]) for p in range(t.nb, t.m)] for q in range(t.n_leaves)])
    f = np.array([[_f(pts[q, p]) for p in range(t.nb, t.m)] for q in range(t.n_leaves)])
    gb = np.zeros((t.n_leaves, t.nb), np.complex128)
    nphys = 0
    for q in range(t.n_leaves):
        box = t.leaf_box(q)                     # rows: axis; columns: lo, hi
        for p in range(t.nb):
            fc = int(face[p])
            ax, side = (fc - 1) // 2, (fc - 1) % 2
            if box[ax, side] != (0.0 if side == 0 else 1.0):
                continue
            x = pts[q, p]
            gb[q, p] = _g_face(fc - 1)(np.array([x[1 - ax], 0.0, 0.0]))
            nphys += 1
    t.factor(k)
    u = t.solve(f, gb)
    assert len(calls) == nphys       # one callback per physical boundary point
    assert len(sol) == t.n_leaves
    for q in range(t.n_leaves):
        np.testing.assert_array_equal(sol[q].u, u[q])
        np.testing.assert_array_equal(sol[q].box, t.leaf_box(q).T)
