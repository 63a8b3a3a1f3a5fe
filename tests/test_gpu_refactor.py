"""GPU: refactoring in place.  A tree that keeps its factorization's working
memory (RS_RETAIN_SCRATCH, as served trees do) refactors with new potentials
into the same buffers; from its second factorization on (tiles tuned, nothing
reallocated) the launch sequence is captured once and replayed.  Every
factorization must give the bits of a fresh tree factored with the same
potential."""
import numpy as np
import pytest

import robinsolver_amd as rsa

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _k(x, kb, j):
    return kb ** 2 * (1 + 0.2 * np.sin((3 + j) * x[..., 0]) * np.cos(2 * x[..., 1])) + (0.02 + 0.01 * j) * 1j * kb ** 2


@pytest.mark.parametrize("d,n,L,kb,prec", [(2, 128, 6, 12.0, rsa.RS_FP64), (2, 128, 6, 12.0, rsa.RS_MIXED),
                                           (3, 16, 3, 4.0, rsa.RS_FP64)])
def test_refactor_in_place_matches_fresh_factorization(d, n, L, kb, prec, gpu):
    print("case: refactor in place", d, n, L, prec, flush=True)
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=prec, n_refine=1, retain=rsa.RS_RETAIN_SCRATCH)
    pts, _ = t.all_points()
    rng = np.random.default_rng(5)
    f = rng.standard_normal((3, t.n_leaves, t.n_int)) + 1j * rng.standard_normal((3, t.n_leaves, t.n_int))
    g = rng.standard_normal((3, t.n_leaves, t.nb)) + 1j * rng.standard_normal((3, t.n_leaves, t.nb))
    # factorizations 1 (tiles tuned), 2 (captured), 3 and 4 (replayed)
    for j in range(4):
        k = _k(pts[:, t.nb:], kb, j)
        t.factor(k)
        u = t.solve(f, g)
        fresh = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=prec, n_refine=1)
        fresh.factor(k)
        u_ref = fresh.solve(f, g)
        fresh.close()
        assert np.array_equal(u, u_ref), (j, np.abs(u - u_ref).max())
    t.close()
