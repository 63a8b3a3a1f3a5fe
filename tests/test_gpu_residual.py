"""GPU: the FD residual check rs_residual_norms.  Its definition is pinned
against the independent global coupled system (tests/fdsystem.py, built from
the oracle's leaf matrices): for a perturbed solution both give the same
||b - K u|| / ||b||.  Then the check is applied to the solvers' own results:
the FP64 path leaves a residual at rounding level, the complex64 solve alone
does not, one refinement step restores it."""
import numpy as np
import pytest

import robinsolver_amd as rsa
from oracle.pyoracle import OracleTree
from tests import fdsystem as fds

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _k(x, kb):
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1) * np.cos(2 * x[..., 1])) + 0.05j * kb ** 2


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda:0")


@pytest.mark.parametrize("d,n,L,kb", [(2, 64, 4, 10.0), (3, 16, 6, 4.0)])
def test_residual_norm_matches_global_system(d, n, L, kb, gpu):
    print("case: residual norms", d, n, L, kb, flush=True)
    o = OracleTree(d, L, UNIT, 1.0 / n, -kb)
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb)
    pts, face = o.all_points()
    k = _k(pts[:, o.nb:], kb)
    t.factor(k)
    mats = [o.leaf_matrix(q, k[q]) for q in range(o.n_leaves)]
    K, nbl = fds.global_system(mats, pts, face, d, o.kb, o.bndry_params)
    rng = np.random.default_rng(21)
    nrhs = 3
    f = _rand_c(rng, (nrhs, o.n_leaves, o.n_int))
    g = _rand_c(rng, (nrhs, o.n_leaves, o.nb))
    u = t.solve(f, g)
    up = u + 1e-6 * np.abs(u).max() * _rand_c(rng, u.shape)
    fd, gd, ud = _dev(f), _dev(g), _dev(up)
    rel = t.residual_norms(nrhs, fd.data_ptr(), gd.data_ptr(), ud.data_ptr())
    for r in range(nrhs):
        b = fds.global_rhs(f[r], g[r], nbl, face)
        ref = np.linalg.norm(b - K @ up[r].reshape(-1)) / np.linalg.norm(b)
        assert abs(rel[r] - ref) <= 1e-6 * ref, (r, rel[r], ref)
    # the FP64 solution itself: residual at rounding level
    ud = _dev(u)
    rel0 = t.residual_norms(nrhs, fd.data_ptr(), gd.data_ptr(), ud.data_ptr())
    assert rel0.max() < 1e-11, rel0


def test_mixed_residual_before_and_after_refinement(gpu):
    print("case: residual mixed", flush=True)
    import torch
    kb = 10.0
    t = rsa.GpuTree(2, 4, UNIT, 1.0 / 64, -kb, precision=rsa.RS_MIXED, n_refine=0)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], kb))
    rng = np.random.default_rng(22)
    f = _dev(_rand_c(rng, (2, t.n_leaves, t.n_int)))
    g = _dev(_rand_c(rng, (2, t.n_leaves, t.nb)))
    u = torch.zeros((2, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    t.solve_device(2, f.data_ptr(), g.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    r0 = t.residual_norms(2, f.data_ptr(), g.data_ptr(), u.data_ptr())
    t.set_refine(1)
    t.solve_device(2, f.data_ptr(), g.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    r1 = t.residual_norms(2, f.data_ptr(), g.data_ptr(), u.data_ptr())
    print("residual n_refine 0:", r0, "1:", r1, flush=True)
    assert r0.min() > 1e-9
    assert r1.max() < 1e-10
