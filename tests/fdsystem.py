"""Independent checks for the HPS tree: the global coupled FD system and the
manufactured solutions of the reference driver.

The tree of robinsolver is an exact direct solver of one global linear system
(leaf FD equations plus interface transmission conditions).  This module builds
that system from the leaf matrices alone -- without the tree, the merge, or the
solve passes -- and solves it with numpy/scipy LAPACK.  A mistake in the tree
(a sign, an index, a block placement, nu/zeta) shows up as a mismatch at
rounding level.

Transmission condition at a leaf boundary point p that coincides with point p'
of the neighbouring leaf (derived from the merge system M[g1; g2] = [h1; h2],
robin_tree_module.f90:425-467, 868-873):
    zeta * g_l(p) + nu * g_nb(p') - h_l(p) = 0,
with g = (alpha + D_n) u the incoming data (the leaf's boundary row) and
h = (sigma + D_n) u the outgoing data, alpha = +i kb, sigma = -i kb
(fd_module.f90:131, 187-216; robin_tree_module.f90:683).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def nu_zeta(bp):
    a11, a12, a21, a22 = bp[0, 0], bp[0, 1], bp[1, 0], bp[1, 1]
    nu = (-a11 * a22 + a12 * a21) / (2 * a11 * a12)
    zeta = (a11 * a22 + a12 * a21) / (2 * a11 * a12)
    return nu, zeta


def face_normal(face):
    """Outward unit normal axis and sign for face id 1..6."""
    ax = (face - 1) // 2
    sgn = -1.0 if face % 2 == 1 else 1.0
    return ax, sgn


def pair_points(pts, faces, d, tol=1e-9):
    """For every (leaf, boundary point) find the coincident point of another
    leaf, matching by coordinates.  Returns nb_leaf, nb_idx (-1 = physical)."""
    nl, m, _ = pts.shape
    bidx = np.nonzero(faces > 0)[0]
    keys = {}
    for q in range(nl):
        for p in bidx:
            key = tuple(np.round(pts[q, p, :d] / tol).astype(np.int64))
            keys.setdefault(key, []).append((q, p))
    nbl = -np.ones((nl, m), dtype=np.int64)
    nbi = -np.ones((nl, m), dtype=np.int64)
    for lst in keys.values():
        if len(lst) == 2:
            (q1, p1), (q2, p2) = lst
            nbl[q1, p1], nbi[q1, p1] = q2, p2
            nbl[q2, p2], nbi[q2, p2] = q1, p1
        elif len(lst) > 2:
            raise AssertionError("more than two coincident boundary points")
    return nbl, nbi


def global_system(leaf_mats, pts, faces, d, kb, bp):
    """Sparse global matrix K over all leaf unknowns (leaf-major)."""
    nl, m, _ = pts.shape
    nu, zeta = nu_zeta(bp)
    alpha, sigma = 1j * kb, -1j * kb
    nbl, nbi = pair_points(pts, faces, d)
    rows, cols, vals = [], [], []
    for q in range(nl):
        A = leaf_mats[q]
        for p in range(m):
            nz = np.nonzero(A[p])[0]
            if faces[p] == 0 or nbl[q, p] < 0:
                rows += [q * m + p] * len(nz); cols += list(q * m + nz); vals += list(A[p, nz])
                continue
            # zeta*g_l + nu*g_nb - h_l, h_l = g_l + (sigma - alpha) u_l(p)
            r = q * m + p
            rows += [r] * len(nz); cols += list(q * m + nz); vals += list((zeta - 1.0) * A[p, nz])
            rows.append(r); cols.append(q * m + p); vals.append(-(sigma - alpha))
            q2, p2 = nbl[q, p], nbi[q, p]
            nz2 = np.nonzero(leaf_mats[q2][p2])[0]
            rows += [r] * len(nz2); cols += list(q2 * m + nz2); vals += list(nu * leaf_mats[q2][p2, nz2])
    K = sp.csr_matrix((vals, (rows, cols)), shape=(nl * m, nl * m), dtype=np.complex128)
    return K, nbl


def global_rhs(f_int, g_bnd, nbl, faces):
    """Right-hand side: f on interior rows, g on physical boundary rows,
    zero on interface rows."""
    nl = f_int.shape[0]
    nbp = g_bnd.shape[1]
    b = np.zeros((nl, nbp + f_int.shape[1]), np.complex128)
    b[:, nbp:] = f_int
    phys = nbl[:, :nbp] < 0
    b[:, :nbp] = np.where(phys, g_bnd, 0.0)
    return b.reshape(-1)


def solve_global(K, b):
    return spla.spsolve(K.tocsc(), b)


# Manufactured solutions of the reference driver (example.f90:107-276).
def true_sol2d(x):
    return np.cos(x[..., 0] / np.sqrt(2.0)) * np.cos(x[..., 1] / np.sqrt(2.0))


def g2d(face, t):
    """2D boundary data of example.f90 (zero2D / one2D), evaluated at the packed
    tangential coordinate t = x(1) (fd_module.f90:645-651)."""
    s2 = np.sqrt(2.0)
    if face in (1, 3):
        return 1j * np.cos(t / s2)
    return 1j * np.cos(1 / s2) * np.cos(t / s2) - 1 / s2 * np.sin(1 / s2) * np.cos(t / s2)


def truesol_nonsym(x):
    s3 = np.sqrt(3.0)
    return np.cos(x[..., 0] / s3) * np.sin(x[..., 1] / s3) * np.sin(x[..., 2] / s3 + 1.0)


def g3d(face, a, b):
    """3D boundary data of example.f90 (x0bc..z1bc) at packed (x(1), x(2))."""
    s3 = np.sqrt(3.0)
    if face == 1:
        return 1j * np.sin(a / s3) * np.sin(b / s3 + 1.0)
    if face == 2:
        return (1j * np.cos(1 / s3) * np.sin(a / s3) * np.sin(b / s3 + 1.0)
                - 1 / s3 * np.sin(1 / s3) * np.sin(a / s3) * np.sin(b / s3 + 1.0))
    if face == 3:
        return -1 / s3 * np.cos(a / s3) * np.sin(b / s3 + 1.0)
    if face == 4:
        return (1j * np.cos(a / s3) * np.sin(1 / s3) * np.sin(b / s3 + 1.0)
                + 1 / s3 * np.cos(a / s3) * np.cos(1 / s3) * np.sin(b / s3 + 1.0))
    if face == 5:
        return (1j * np.cos(a / s3) * np.sin(b / s3) * np.sin(1.0)
                - 1 / s3 * np.cos(a / s3) * np.sin(b / s3) * np.cos(1.0))
    return (1j * np.cos(a / s3) * np.sin(b / s3) * np.sin(1 / s3 + 1.0)
            + 1 / s3 * np.cos(a / s3) * np.sin(b / s3) * np.cos(1 / s3 + 1.0))


def packed_face_coords(xyz, face, d):
    """The coordinates g_face receives (fd_module.f90:645-651): the tangential
    coordinates in axis order; for d = 2 the single tangential coordinate in
    x(1)."""
    ax = (face - 1) // 2
    tang = [j for j in range(d) if j != ax]
    return [xyz[..., j] for j in tang]


def sample_g(pts, faces, d, gfun):
    """Sample boundary data gfun(face, *packed) at every leaf boundary point."""
    nbp = int(np.count_nonzero(faces))
    g = np.zeros((pts.shape[0], nbp), np.complex128)
    for p in range(nbp):
        fc = int(faces[p])
        g[:, p] = gfun(fc, *packed_face_coords(pts[:, p, :], fc, d))
    return g
