"""Pins the CPU oracle (oracle/rs_oracle.c) against what the mathematics fixes.

The reference ships no golden vectors and cannot be built here (it needs an
external LAPACK/BLAS), so every function of the oracle is checked against an
independent route:
  * leaf FD matrix: exactness on quadratic polynomials (each interior and
    boundary stencil of fd_module.f90:104-224 is exact for quadratics);
  * leaf RtR: I + (sigma - alpha) (A^-1)_BB from a numpy inverse of the pinned A;
  * merge + root RtR and both solve passes: the global coupled system of
    tests/fdsystem.py solved by scipy, at rounding level;
  * apply_invm: inv(M) M = I with M from the children's T_II;
  * end to end: O(h^2) convergence to the manufactured solutions of
    example.f90 in 2D and 3D (this also pins the g_face coordinate packing).
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleTree
from tests import fdsystem as fds

UNIT = [[0, 1], [0, 1], [0, 1]]


def _rng(seed):
    return np.random.default_rng(seed)


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _k_field(pts_int, kb, rng):
    x = pts_int
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1) * np.cos(2 * x[..., 1])) + 0.05j * kb ** 2


def _quad(x, c):
    # generic quadratic in 3 variables
    return (c[0] + c[1] * x[..., 0] + c[2] * x[..., 1] + c[3] * x[..., 2] + c[4] * x[..., 0] ** 2
            + c[5] * x[..., 1] ** 2 + c[6] * x[..., 2] ** 2 + c[7] * x[..., 0] * x[..., 1]
            + c[8] * x[..., 1] * x[..., 2] + c[9] * x[..., 0] * x[..., 2])


def _quad_grad_lap(x, c):
    gx = c[1] + 2 * c[4] * x[..., 0] + c[7] * x[..., 1] + c[9] * x[..., 2]
    gy = c[2] + 2 * c[5] * x[..., 1] + c[7] * x[..., 0] + c[8] * x[..., 2]
    gz = c[3] + 2 * c[6] * x[..., 2] + c[8] * x[..., 1] + c[9] * x[..., 0]
    return np.stack([gx, gy, gz], -1)


@pytest.mark.parametrize("d,n,L", [(2, 12, 2), (2, 16, 3), (3, 8, 3)])
def test_leaf_matrix_exact_on_quadratics(d, n, L):
    kb = 3.0
    t = OracleTree(d, L, UNIT, 1.0 / n, kb)
    rng = _rng(1)
    c = rng.standard_normal(10)
    if d == 2:
        c[[3, 6, 8, 9]] = 0.0
    for leaf in (0, t.n_leaves - 1):
        xyz, face = t.leaf_points(leaf)
        k = _rand_c(rng, t.n_int)
        A = t.leaf_matrix(leaf, k)
        u = _quad(xyz, c)
        lap = 2 * (c[4] + c[5] + (c[6] if d == 3 else 0.0))
        grad = _quad_grad_lap(xyz, c)
        exp = np.zeros(t.m, np.complex128)
        nbp = t.nb
        for p in range(nbp):
            ax, sgn = fds.face_normal(face[p])
            exp[p] = 1j * kb * u[p] + sgn * grad[p, ax]
        exp[nbp:] = lap + k * u[nbp:]
        assert np.allclose(A @ u, exp, rtol=1e-9, atol=1e-8 * np.abs(exp).max())


@pytest.mark.parametrize("d,n,L", [(2, 16, 2), (3, 8, 3)])
def test_leaf_rtr_identity(d, n, L):
    kb = 4.0
    t = OracleTree(d, L, UNIT, 1.0 / n, kb)
    rng = _rng(2)
    pts, face = t.all_points()
    k = _k_field(pts[:, t.nb:], kb, rng)
    t.factor(k)
    for leaf in range(t.n_leaves):
        A = t.leaf_matrix(leaf, k[leaf])
        Ainv = np.linalg.inv(A)
        T_exp = np.eye(t.nb) + (-2j * kb) * Ainv[:t.nb, :t.nb]
        T = t.node_rtr(t.level_node_id(L, leaf))
        assert np.allclose(T, T_exp, rtol=1e-10, atol=1e-10 * np.abs(T_exp).max())


def _setup_global(t, k):
    pts, face = t.all_points()
    mats = [t.leaf_matrix(q, k[q]) for q in range(t.n_leaves)]
    K, nbl = fds.global_system(mats, pts, face, t.d, t.kb, t.bndry_params)
    return pts, face, K, nbl


CASES = [
    (2, 16, 2, 5.0, None),
    (2, 32, 4, 12.0, None),
    (2, 24, 3, 7.0, None),
    (3, 8, 3, 3.0, None),
    (3, 12, 4, 2.0, None),
    (2, 16, 3, 5.0, np.array([[1.3j * 5, 0.7], [-0.8j * 5, 1.1]])),
    (3, 8, 2, 3.0, np.array([[2.0 + 1j, 1.0], [0.5, 1.0 - 0.5j]])),
]


@pytest.mark.parametrize("d,n,L,kb,bp", CASES)
def test_solve_matches_global_system(d, n, L, kb, bp):
    t = OracleTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp)
    rng = _rng(3 + n + L)
    pts, face = t.all_points()
    k = _k_field(pts[:, t.nb:], kb, rng)
    t.factor(k)
    _, _, K, nbl = _setup_global(t, k)
    f = _rand_c(rng, (t.n_leaves, t.n_int))
    g = _rand_c(rng, (t.n_leaves, t.nb))
    u = t.solve(f, g)
    b = fds.global_rhs(f, g, nbl, face)
    u_ref = fds.solve_global(K, b).reshape(t.n_leaves, t.m)
    rel = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    assert rel < 1e-10, rel
    # the tree solution satisfies the global system to rounding
    res = np.linalg.norm(K @ u.reshape(-1) - b) / np.linalg.norm(b)
    assert res < 1e-10, res


def _root_face_order(t, pts, face):
    """Root boundary points in concatenation order: for each root face, the
    leaves touching it in DFS order, each with its own face points in order
    (apply_bc concatenation, robin_tree_module.f90:1205-1230)."""
    d = t.d
    order = []
    for fc in range(1, 2 * d + 1):
        ax = (fc - 1) // 2
        side = 1 if fc % 2 == 0 else 0
        for q in range(t.n_leaves):
            box = t.leaf_box(q)
            if abs(box[ax, side] - float(side)) > 1e-12:
                continue
            for p in np.nonzero(face == fc)[0]:
                order.append((q, p))
    return order


@pytest.mark.parametrize("d,n,L,kb,bp", CASES[:5] + CASES[6:])
def test_root_rtr_matches_global_system(d, n, L, kb, bp):
    t = OracleTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp)
    rng = _rng(5)
    pts, face = t.all_points()
    k = _k_field(pts[:, t.nb:], kb, rng)
    t.factor(k)
    _, _, K, nbl = _setup_global(t, k)
    order = _root_face_order(t, pts, face)
    T = t.node_rtr(0)
    assert T.shape[0] == len(order)
    nb_root = len(order)
    B = np.zeros((t.n_leaves * t.m, nb_root), np.complex128)
    for j, (q, p) in enumerate(order):
        B[q * t.m + p, j] = 1.0
    U = spla_solve(K, B)
    mats = [t.leaf_matrix(q, k[q]) for q in range(t.n_leaves)]
    T_ref = np.zeros_like(T)
    for j in range(nb_root):
        u = U[:, j].reshape(t.n_leaves, t.m)
        for r, (q, p) in enumerate(order):
            T_ref[r, j] = mats[q][p] @ u[q] + (-2j * kb) * u[q, p]
    assert np.linalg.norm(T - T_ref) / np.linalg.norm(T_ref) < 1e-10


def spla_solve(K, B):
    import scipy.sparse.linalg as spla
    lu = spla.splu(K.tocsc())
    return lu.solve(B)


@pytest.mark.parametrize("d,n,L,bp", [(2, 16, 3, None), (3, 8, 3, None),
                                        (2, 16, 2, np.array([[1.5j, 0.9], [-0.4j, 1.2]]))])
def test_apply_invm_inverts_M(d, n, L, bp):
    kb = 4.0
    t = OracleTree(d, L, UNIT, 1.0 / n, kb, bndry_params=bp)
    rng = _rng(7)
    pts, _ = t.all_points()
    t.factor(_k_field(pts[:, t.nb:], kb, rng))
    nu, zeta = fds.nu_zeta(t.bndry_params)
    for lvl in range(L):
        nid = t.level_node_id(lvl, 0)
        isz, i1, i2 = t.node_iface(nid)
        c1 = t.level_node_id(lvl + 1, 0)
        c2 = t.level_node_id(lvl + 1, 1)
        T1, T2 = t.node_rtr(c1), t.node_rtr(c2)
        o1 = t.node_face_offsets(c1)[i1 - 1]
        o2 = t.node_face_offsets(c2)[i2 - 1]
        T1ii = T1[o1:o1 + isz, o1:o1 + isz]
        T2ii = T2[o2:o2 + isz, o2:o2 + isz]
        I = np.eye(isz)
        M = np.block([[zeta * I - T1ii, nu * I], [nu * I, zeta * I - T2ii]])
        x = _rand_c(rng, 2 * isz)
        y = t.apply_invm(nid, M @ x)
        assert np.linalg.norm(y - x) / np.linalg.norm(x) < 1e-11


def _conv_errors_2d(ns, L):
    errs = []
    for n in ns:
        t = OracleTree(2, L, UNIT, 1.0 / n, 1.0)
        pts, face = t.all_points()
        t.factor(np.ones((t.n_leaves, t.n_int)))
        g = fds.sample_g(pts, face, 2, fds.g2d)
        u = t.solve(np.zeros((t.n_leaves, t.n_int)), g)
        ue = fds.true_sol2d(pts)
        errs.append(np.linalg.norm(u - ue) / np.linalg.norm(ue))
    return errs


def test_manufactured_2d_converges_second_order():
    errs = _conv_errors_2d([16, 32, 64], 2)
    assert errs[-1] < 1e-4
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.2 < r1 < 4.8 and 3.2 < r2 < 4.8, errs


def test_manufactured_3d_converges_second_order():
    errs = []
    for n, L in [(8, 3), (16, 6)]:
        t = OracleTree(3, L, UNIT, 1.0 / n, 1.0)
        pts, face = t.all_points()
        t.factor(np.ones((t.n_leaves, t.n_int)))
        g = fds.sample_g(pts, face, 3, fds.g3d)
        u = t.solve(np.zeros((t.n_leaves, t.n_int)), g)
        ue = fds.truesol_nonsym(pts)
        errs.append(np.linalg.norm(u - ue) / np.linalg.norm(ue))
    assert errs[-1] < 2e-3
    assert 3.0 < errs[0] / errs[1] < 6.0, errs


def test_wrong_packing_breaks_manufactured_2d():
    # If g_face received the normal coordinate instead of the tangential one,
    # the manufactured solution would not be reproduced: the packing matters.
    t = OracleTree(2, 2, UNIT, 1.0 / 32, 1.0)
    pts, face = t.all_points()
    t.factor(np.ones((t.n_leaves, t.n_int)))
    nbp = t.nb
    g = np.zeros((t.n_leaves, nbp), np.complex128)
    for p in range(nbp):
        fc = int(face[p]); ax = (fc - 1) // 2
        g[:, p] = fds.g2d(fc, pts[:, p, ax])
    u = t.solve(np.zeros((t.n_leaves, t.n_int)), g)
    ue = fds.true_sol2d(pts)
    assert np.linalg.norm(u - ue) / np.linalg.norm(ue) > 1e-2


def test_geometry_and_output_order():
    t = OracleTree(2, 4, UNIT, 1.0 / 64, 10.0)
    assert (t.n_leaves, t.m, t.n_int, t.nb) == (16, 320, 256, 64)
    # DFS order: first cut is x (ties go to x), child1 is the lower half
    b0, b1 = t.leaf_box(0), t.leaf_box(t.n_leaves - 1)
    assert np.allclose(b0[:2], [[0, 0.25], [0, 0.25]])
    assert np.allclose(b1[:2], [[0.75, 1.0], [0.75, 1.0]])
    assert np.allclose(t.leaf_box(8)[0], [0.5, 0.75]) and np.allclose(t.leaf_box(8)[1], [0, 0.25])
    xyz, face = t.leaf_points(0)
    # boundary first: x1 face at x = 0, then x2, y1, y2; interior cell-centred, x fastest
    assert np.all(face[:16] == 1) and np.allclose(xyz[:16, 0], 0.0)
    assert np.allclose(xyz[:16, 1], (np.arange(16) + 0.5) / 64)
    assert np.all(face[48:64] == 4) and np.allclose(xyz[48:64, 1], 0.25)
    assert np.allclose(xyz[64:80, 0], (np.arange(16) + 0.5) / 64)
    assert np.allclose(xyz[80, :2], [0.5 / 64, 1.5 / 64])


def test_get_h_clamp_and_ceiling():
    # get_h: n = max(ceil(w / h_tgt), 3) (fd_module.f90:1075-1082)
    t = OracleTree(2, 6, UNIT, 1.0 / 64, 1.0)
    assert t.ptbox[:2] == (8, 8)
    t = OracleTree(2, 8, UNIT, 1.0 / 32, 1.0)   # leaves 1/16 wide -> 2 < 3 -> clamp
    assert t.ptbox[:2] == (3, 3)


@pytest.mark.parametrize("args,msg", [
    ((2, 2, [[0, 1], [1, 0], [0, 1]], 0.1, 1.0, None), "Invalid operator domain"),
    ((4, 2, UNIT, 0.1, 1.0, None), "Dimension must be 2 or 3"),
    ((2, 2, UNIT, 0.1, 1.0, np.array([[1.0, 1.0], [1.0, 1.0]])), "nearly singular"),
])
def test_errors(args, msg):
    with pytest.raises(RuntimeError, match=msg):
        OracleTree(*args[:5], bndry_params=args[5])
