"""GPU: solves replayed from captured hipGraphs give the same bits as direct
launches.  Each graph is replayed several times with new inputs, with other
work (solves of another shape, other precision, a refactor in place) between
the replays, and every replay is checked."""
import numpy as np
import pytest

import robinsolver_amd as rsa

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _rc(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda:0")


def _pair(prec, d=2, n=64, L=4, kb=10.0):
    trees = []
    for graphs in (True, False):
        t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=prec, n_refine=1, graphs=graphs)
        pts, _ = t.all_points()
        t.factor(kb ** 2 * (1 + 0.3 * np.sin(4 * pts[:, t.nb:, 0]) * np.cos(3 * pts[:, t.nb:, 1])) + 0.1j * kb)
        trees.append(t)
    return trees


def _solve(t, f, g):
    import torch
    u = torch.zeros((f.shape[0], t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    t.solve_device(f.shape[0], f.data_ptr(), g.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    return u.cpu().numpy()


@pytest.mark.parametrize("prec", [rsa.RS_MIXED, rsa.RS_FP64])
def test_graph_replays_match_direct_launches(prec, gpu):
    print("case: graph replays, precision", prec, flush=True)
    tg, td = _pair(prec)
    rng = np.random.default_rng(31)
    for rep in range(4):
        for nrhs in (1, 5, 1, 12):
            f = _dev(_rc(rng, nrhs, tg.n_leaves, tg.n_int))
            g = _dev(_rc(rng, nrhs, tg.n_leaves, tg.nb))
            assert np.array_equal(_solve(tg, f, g), _solve(td, f, g)), (rep, nrhs)
        if rep == 1:
            # other work between replays: a mixed solve of an FP64 tree (a mixed
            # tree takes no FP64 solves) or another shape, then a refactor in place
            f = _dev(_rc(rng, 3, tg.n_leaves, tg.n_int))
            g = _dev(_rc(rng, 3, tg.n_leaves, tg.nb))
            import torch
            u = torch.zeros((3, tg.n_leaves, tg.m), dtype=torch.complex128, device="cuda:0")
            tg.solve_device(3, f.data_ptr(), g.data_ptr(), u.data_ptr(), precision=rsa.RS_MIXED)
            torch.cuda.synchronize()
        if rep == 2:
            for t in (tg, td):
                pts, _ = t.all_points()
                t.factor(90.0 * (1 + 0.2 * np.cos(5 * pts[:, t.nb:, 1])) + 0.5j)


def test_graph_host_api_and_refine_change(gpu):
    print("case: graph host api", flush=True)
    tg, td = _pair(rsa.RS_MIXED, n=48, L=4)
    rng = np.random.default_rng(32)
    f = _rc(rng, 4, tg.n_leaves, tg.n_int)
    g = _rc(rng, 4, tg.n_leaves, tg.nb)
    for nref in (1, 2, 1):
        tg.set_refine(nref)
        td.set_refine(nref)
        for _ in range(2):
            assert np.array_equal(tg.solve(f, g), td.solve(f, g)), nref
