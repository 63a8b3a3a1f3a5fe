"""The serving loop's scheduling policy on the host (simulated device): per-tree
arrival order, coalescing of queued right-hand sides into passes of at most
max_cols, factor requests ordered against the tree's solves, admission by
memory (one workspace arena per lane, shared by the trees) with eviction of
idle arenas, and the latency statistics.  Trees
are really planned (the memory figures come from their plans); passes and
factorizations take modelled times instead of running on a GPU."""
import time

import pytest

import robinsolver_amd as rsa
from robinsolver_amd.serve import Server, RS_REQ_DONE, RS_REQ_FAILED, lane_bytes

UNIT = [[0, 1], [0, 1], [0, 1]]


def _tree(n=64, L=4, prec=rsa.RS_MIXED):
    return rsa.GpuTree(2, L, UNIT, 1.0 / n, -10.0, precision=prec, n_refine=1)


def _server(**kw):
    args = dict(max_cols=32, lanes=2, hbm_limit=1 << 40, simulate=True, sim_pass_s=0.01, sim_col_s=0.0,
                sim_factor_s=0.02)
    args.update(kw)
    return Server(**args)


def _factor(srv, tid):
    r = srv.factor(tid, 0)
    srv.wait(r)


def test_coalescing_and_arrival_order():
    t = _tree()
    with _server() as srv:
        tid = srv.add_tree(t)
        _factor(srv, tid)
        sizes = [1, 5, 7, 3, 16, 9, 2, 30, 4, 1, 8, 12]
        reqs = [srv.solve(tid, n, 0, 0, 0) for n in sizes]
        srv.drain()
        info = [srv.info(r) for r in reqs]
        passes = {}
        for n, (ts, tb, td, p) in zip(sizes, info):
            assert ts <= tb <= td
            passes.setdefault(p, []).append(n)
        # every pass within max_cols; requests of one tree run in arrival order
        assert all(sum(v) <= 32 for v in passes.values())
        pids = [p for (_, _, _, p) in info]
        assert pids == sorted(pids)
        # requests queued while the first pass ran were coalesced
        assert len(passes) < len(sizes)
        st = srv.stats()
        assert st["rhs_done"] == sum(sizes) and st["requests_done"] == len(sizes)
        assert st["lat_p50_s"] <= st["lat_p99_s"] <= st["lat_max_s"]
        assert st["mean_cols"] > 1.0


def test_factor_is_ordered_against_the_trees_solves():
    t = _tree()
    with _server() as srv:
        tid = srv.add_tree(t)
        _factor(srv, tid)
        a = srv.solve(tid, 4, 0, 0, 0)
        f = srv.factor(tid, 0)
        b = srv.solve(tid, 4, 0, 0, 0)
        srv.drain()
        _, _, a_done, pa = srv.info(a)
        _, f_start, f_done, _ = srv.info(f)
        _, b_start, _, pb = srv.info(b)
        assert a_done <= f_start and f_done <= b_start
        assert pa != pb          # b was not coalesced across the factorization


def test_solve_on_unfactored_tree_fails():
    t = _tree()
    with _server() as srv:
        tid = srv.add_tree(t)
        r = srv.solve(tid, 2, 0, 0, 0)
        srv.drain()
        assert srv.state(r) == RS_REQ_FAILED
        assert "factored first" in srv.error(r)
        with pytest.raises(rsa.RobinError, match="factored first"):
            srv.wait(r)


def test_nrhs_bound_and_bad_tree():
    t = _tree()
    with _server(max_cols=8) as srv:
        tid = srv.add_tree(t)
        with pytest.raises(rsa.RobinError, match="max_cols"):
            srv.solve(tid, 9, 0, 0, 0)
        with pytest.raises(rsa.RobinError, match="no such tree"):
            srv.solve(tid + 1, 1, 0, 0, 0)
        with pytest.raises(rsa.RobinError, match="already registered"):
            srv.add_tree(t)


def test_lanes_overlap_across_trees():
    t1, t2 = _tree(), _tree(128, 6)
    with _server(sim_pass_s=0.05) as srv:
        a, b = srv.add_tree(t1), srv.add_tree(t2)
        _factor(srv, a)
        _factor(srv, b)
        r1 = srv.solve(a, 3, 0, 0, 0)
        r2 = srv.solve(b, 3, 0, 0, 0)
        srv.drain()
        _, s1, d1, _ = srv.info(r1)
        _, s2, d2, _ = srv.info(r2)
        assert s2 < d1 and s1 < d2          # both passes were in flight together


def test_lane_arena_is_shared_across_trees():
    # three trees of different sizes: a lane holds one arena, the size of the
    # largest tree it ran, and each tree its descriptors on the lanes it ran on
    ts = [_tree(64, 4), _tree(128, 6), _tree(256, 8)]
    res = sum(t.memory_estimate(16)[1] for t in ts)   # served: resident = peak
    lb = [lane_bytes(t, 16) for t in ts]
    arena = max(a for a, _ in lb)
    descs = sum(d for _, d in lb)
    own = [t.memory_estimate(16)[2] for t in ts]
    assert arena <= max(own) and sum(own) > 1.2 * arena    # not the sum over trees
    for lanes in (1, 2):
        with _server(max_cols=16, lanes=lanes, sim_pass_s=0.02) as srv:
            ids = [srv.add_tree(t) for t in ts]
            for i in ids:
                _factor(srv, i)
            for _ in range(3):
                for i in ids:
                    srv.solve(i, 4, 0, 0, 0)
                time.sleep(0.005)
            srv.drain()
            st = srv.stats()
        assert st["evictions"] == 0
        if lanes == 1:
            assert st["held_bytes"] == res + arena + descs, (st["held_bytes"], res, arena, descs)
        else:
            assert res + arena + descs <= st["held_bytes"] <= res + 2 * arena + 2 * descs


def test_grouped_passes_hold_several_trees():
    ts = [_tree(64, 4), _tree(128, 6), _tree(256, 8)]
    sizes = [3, 4, 5, 2, 6, 1, 7, 3, 2]
    with _server(max_cols=16, lanes=1, group=3, sim_pass_s=0.03) as srv:
        ids = [srv.add_tree(t) for t in ts]
        for i in ids:
            _factor(srv, i)
        first = srv.solve(ids[0], 2, 0, 0, 0)        # occupies the lane meanwhile
        time.sleep(0.005)
        reqs = [(k % 3, srv.solve(ids[k % 3], n, 0, 0, 0)) for k, n in enumerate(sizes)]
        srv.drain()
        st = srv.stats()
        assert srv.state(first) == RS_REQ_DONE
        pid = [srv.info(r)[3] for _, r in reqs]
    per_pass = {}
    for (ti, _), n, p in zip(reqs, sizes, pid):
        per_pass.setdefault(p, []).append((ti, n))
    for ti in range(3):                               # per tree, arrival order
        mine = [p for (t, _), p in zip(reqs, pid) if t == ti]
        assert mine == sorted(mine)
    assert all(sum(n for _, n in v) <= 16 for v in per_pass.values())
    assert max(len({ti for ti, _ in v}) for v in per_pass.values()) > 1
    assert st["mean_trees"] > 1.0
    res = sum(t.memory_estimate(16)[1] for t in ts)   # served: resident = peak
    arena = max(sum(lane_bytes(t, 16)) for t in ts)   # the arena holds the descriptors too
    assert st["held_bytes"] == res + arena


def test_admission_evicts_idle_lane_arena():
    # a second tree's factorization that needs the memory of the idle lane arena
    t1, t2 = _tree(), _tree()
    with _server(max_cols=16, lanes=1, hbm_limit=1 << 40) as srv:
        srv.add_tree(t1)                        # (registration keeps the factor's working memory)
    r, p, _ = t1.memory_estimate(16)
    assert r == p
    a, d = lane_bytes(t1, 16)
    limit = 2 * p + a + d - a // 2
    with _server(max_cols=16, lanes=1, hbm_limit=limit) as srv:
        i1, i2 = srv.add_tree(t1), srv.add_tree(t2)
        _factor(srv, i1)
        srv.wait(srv.solve(i1, 4, 0, 0, 0))
        assert srv.stats()["held_bytes"] == p + a + d
        _factor(srv, i2)                        # needs p: the arena goes (the descriptors may stay)
        st = srv.stats()
        assert st["evictions"] >= 1 and st["held_bytes"] in (2 * p, 2 * p + d)
        assert st["held_bytes"] <= limit


def test_served_tree_refactors_in_place():
    # registered trees keep their factorization's working memory: resident = peak
    t = _tree()
    r0, p0, _ = t.memory_estimate(16)
    assert r0 < p0
    with _server(max_cols=16, lanes=1) as srv:
        tid = srv.add_tree(t)
        r, p, _ = t.memory_estimate(16)
        assert r == p == p0
        _factor(srv, tid)
        _factor(srv, tid)
        assert srv.stats()["held_bytes"] == p


def test_admission_refuses_what_can_never_fit():
    t = rsa.GpuTree(2, 4, UNIT, 1.0 / 64, -10.0, precision=rsa.RS_MIXED, n_refine=1, retain=rsa.RS_RETAIN_SCRATCH)
    r, p, w = t.memory_estimate(16)
    a, d = lane_bytes(t, 16)
    assert p == r and 0 < d < a < w
    with _server(max_cols=16, lanes=1, hbm_limit=r + a // 2) as srv:
        tid = srv.add_tree(t)
        _factor(srv, tid)
        q = srv.solve(tid, 4, 0, 0, 0)
        srv.drain()
        assert srv.state(q) == RS_REQ_FAILED
        assert "not enough device memory" in srv.error(q)
    with _server(max_cols=16, lanes=1, hbm_limit=p // 2) as srv:
        tid = srv.add_tree(t)
        f = srv.factor(tid, 0)
        srv.drain()
        assert srv.state(f) == RS_REQ_FAILED
        assert "not enough device memory" in srv.error(f)


def test_memory_estimates_scale():
    t = _tree(128, 6)
    r, p, w16 = t.memory_estimate(16)
    _, _, w32 = t.memory_estimate(32)
    assert 0 < r <= p
    assert 1.9 < w32 / w16 < 2.1         # the workspace is per right-hand side
    t64 = rsa.GpuTree(2, 6, UNIT, 1.0 / 128, -10.0, precision=rsa.RS_FP64)
    r64, _, w64 = t64.memory_estimate(16)
    assert r64 > r                        # complex128 solve data (a mixed tree keeps complex64 only)
    assert w64 < w16 * 2.0                # no refinement buffers in FP64


def test_release_and_stats_reset():
    t = _tree()
    with _server() as srv:
        tid = srv.add_tree(t)
        _factor(srv, tid)
        r = srv.solve(tid, 2, 0, 0, 0)
        srv.wait(r)
        srv.release(r)
        with pytest.raises(rsa.RobinError, match="no such request"):
            srv.state(r)
        srv.reset_stats()
        assert srv.stats()["requests_done"] == 0
