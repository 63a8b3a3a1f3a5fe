"""GPU: the serving loop.  Every request's result is bitwise the same when it
is solved alone (rs_solve_device, one request per call) and when the server
coalesces it with other requests of several trees on several streams, or
runs it in a pass grouped across trees; a
factorization queued through the server (including a refactor in place with a
new potential) gives the same results as a fresh tree; mixed-path results stay
within rel. l2 1e-8 of the FP64 oracle; the host-side memory estimates match
what the tree allocates."""
import numpy as np
import pytest

import robinsolver_amd as rsa
from robinsolver_amd.serve import Server, RS_REQ_DONE, lane_bytes
from oracle.pyoracle import OracleTree

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _torch():
    import torch
    return torch


def _k(pts_int, kb, phase=0.0):
    x = pts_int
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1 + phase) * np.cos(2 * x[..., 1])) + 0.02j * kb ** 2


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda:0")


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _make(d, n, L, kb, prec, phase=0.0, autotune=True):
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=prec, n_refine=1, autotune=autotune)
    pts, _ = t.all_points()
    k = _k(pts[:, t.nb:], kb, phase)
    return t, k


def _alone(t, f, g):
    torch = _torch()
    u = torch.empty((f.shape[0], t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    t.solve_device(f.shape[0], f.data_ptr(), g.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    return u.cpu().numpy()


@pytest.mark.parametrize("prec,group", [(rsa.RS_MIXED, 1), (rsa.RS_FP64, 1), (rsa.RS_MIXED, 4), (rsa.RS_FP64, 4)])
def test_results_same_alone_and_in_the_mix(prec, group, gpu):
    """group 4: passes that hold several trees' right-hand sides, every level's
    products launched once across them, give every request its own bits."""
    print("case: serve mix, precision", prec, "group", group, flush=True)
    torch = _torch()
    specs = [(2, 64, 4, 10.0), (2, 128, 6, 20.0), (3, 16, 6, 4.0)]
    trees = []
    for d, n, L, kb in specs:
        t, k = _make(d, n, L, kb, prec)
        t.factor(k)
        trees.append(t)
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(18):
        ti = i % 3 if i < 9 else int(rng.integers(0, 3))
        t = trees[ti]
        nr = int(rng.integers(1, 9))
        f = _dev(_rand_c(rng, (nr, t.n_leaves, t.n_int)))
        g = _dev(_rand_c(rng, (nr, t.n_leaves, t.nb)))
        reqs.append((ti, f, g))
    alone = [_alone(trees[ti], f, g) for ti, f, g in reqs]
    with Server(max_cols=16, lanes=2 if group == 1 else 1, group=group) as srv:
        ids = [srv.add_tree(t) for t in trees]
        outs, rids = [], []
        for ti, f, g in reqs:
            t = trees[ti]
            u = torch.zeros((f.shape[0], t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
            outs.append(u)
            rids.append(srv.solve(ids[ti], f.shape[0], f.data_ptr(), g.data_ptr(), u.data_ptr()))
        srv.drain()
        for r in rids:
            assert srv.state(r) == RS_REQ_DONE, srv.error(r)
        st = srv.stats()
        print("stats:", st, flush=True)
        assert st["passes"] < len(reqs)          # some requests were coalesced
        if group > 1:
            assert st["mean_trees"] > 1.0        # and some passes held several trees
    for (ti, f, g), u, a in zip(reqs, outs, alone):
        assert np.array_equal(u.cpu().numpy(), a)


def test_mixed_served_results_match_the_oracle(gpu):
    print("case: serve vs oracle", flush=True)
    torch = _torch()
    kb = 10.0
    o = OracleTree(2, 4, UNIT, 1.0 / 64, -kb)
    t, k = _make(2, 64, 4, kb, rsa.RS_MIXED)
    o.factor(k)
    t.factor(k)
    rng = np.random.default_rng(5)
    f = _rand_c(rng, (3, t.n_leaves, t.n_int))
    g = _rand_c(rng, (3, t.n_leaves, t.nb))
    fd, gd = _dev(f), _dev(g)
    u = torch.zeros((3, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    with Server(max_cols=8, lanes=2) as srv:
        tid = srv.add_tree(t)
        r = srv.solve(tid, 3, fd.data_ptr(), gd.data_ptr(), u.data_ptr())
        srv.wait(r)
    uh = u.cpu().numpy()
    for c in range(3):
        ref = o.solve(f[c], g[c])
        assert np.linalg.norm(uh[c] - ref) / np.linalg.norm(ref) <= 1e-8


def test_factor_through_the_server_and_refactor_in_place(gpu):
    print("case: serve factor + refactor", flush=True)
    torch = _torch()
    kb = 12.0
    t, k1 = _make(2, 128, 6, kb, rsa.RS_MIXED)
    _, k2 = _make(2, 128, 6, kb, rsa.RS_MIXED, phase=0.7)
    fresh1, _ = _make(2, 128, 6, kb, rsa.RS_MIXED)
    fresh2, _ = _make(2, 128, 6, kb, rsa.RS_MIXED)
    fresh1.factor(k1)
    fresh2.factor(k2)
    rng = np.random.default_rng(9)
    f = _dev(_rand_c(rng, (4, t.n_leaves, t.n_int)))
    g = _dev(_rand_c(rng, (4, t.n_leaves, t.nb)))
    ref1, ref2 = _alone(fresh1, f, g), _alone(fresh2, f, g)
    k1d, k2d = _dev(k1), _dev(k2)
    u1 = torch.zeros((4, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    u2 = torch.zeros_like(u1)
    with Server(max_cols=8, lanes=2) as srv:
        tid = srv.add_tree(t)
        a = srv.factor(tid, k1d.data_ptr())
        s1 = srv.solve(tid, 4, f.data_ptr(), g.data_ptr(), u1.data_ptr())
        b = srv.factor(tid, k2d.data_ptr())
        s2 = srv.solve(tid, 4, f.data_ptr(), g.data_ptr(), u2.data_ptr())
        srv.drain()
        for r in (a, s1, b, s2):
            assert srv.state(r) == RS_REQ_DONE, srv.error(r)
    assert np.array_equal(u1.cpu().numpy(), ref1)
    assert np.array_equal(u2.cpu().numpy(), ref2)


@pytest.mark.parametrize("prec", [rsa.RS_MIXED, rsa.RS_FP64])
def test_memory_estimates_match_allocations(prec, gpu):
    print("case: memory estimates, precision", prec, flush=True)
    torch = _torch()
    t, k = _make(2, 128, 6, 12.0, prec, autotune=False)   # no tuning workspace after the factor
    kd = _dev(k)
    with Server(max_cols=24, lanes=1) as srv:
        tid = srv.add_tree(t)
        # a served tree keeps its factorization's working memory
        res, peak, work = t.memory_estimate(24)
        assert res == peak
        srv.wait(srv.factor(tid, kd.data_ptr()))
        held, most = t.memory_used()
        assert held == res
        assert abs(most - peak) <= 0.01 * peak, (most, peak)
        f = _dev(np.ones((2, t.n_leaves, t.n_int)))
        g = _dev(np.zeros((2, t.n_leaves, t.nb)))
        u = torch.zeros((2, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
        srv.wait(srv.solve(tid, 2, f.data_ptr(), g.data_ptr(), u.data_ptr()))
        held2, _ = t.memory_used()
        # the tree holds only its descriptors on the lane; the lane holds the
        # workspace arena (the tree's own workspace without its column tables
        # and descriptors)
        arena, descs = lane_bytes(t, 24)
        assert held2 - held == descs, (held2 - held, descs)
        assert srv.stats()["held_bytes"] == res + arena + descs
        assert arena + descs <= work - 3 * 24 * 8 <= arena + descs + 4096
