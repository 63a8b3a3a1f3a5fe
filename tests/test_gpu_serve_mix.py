"""GPU: the serving configuration at full size -- eight resident 2D trees
(128^2 .. 1024^2, leaves of 12..20 points per side, kb 10..80, seeded
potentials), a Poisson stream of 240 requests of 1..32 right-hand sides
(Gaussian sources plus the impedance data of a random-angle plane wave), 2% of
them refactorizations with the other of two potentials, mixed precision with
one refinement step, two lanes.

Checks: every request completes; each request's solution is bitwise the same
as the same columns solved alone on a tree factored with the potential that was
current when the request arrived; and within rel. l2 1e-8 of the FP64 GPU
solution of those columns.
"""
import time

import numpy as np
import pytest

import robinsolver_amd as rsa
from robinsolver_amd import workloads as wl
from robinsolver_amd.serve import Server, RS_REQ_DONE

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]
POOL = 32


def _rel(a, b):
    return float(wl.rel_l2_cols(a, b).max())


@pytest.mark.parametrize("group", [1, 8])
def test_serving_mix_of_eight_trees(group, gpu):
    import torch
    print("case: serving mix, 8 trees, group", group, flush=True)
    rng = np.random.default_rng(55)
    trees = []
    for gi, (n, L, kb) in enumerate(wl.MIX_TREES):
        t = rsa.GpuTree(2, L, UNIT, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=1)
        xyz, face = wl.device_points(torch, t)
        pint = xyz[:, t.nb:, :]
        ks = [wl.potential_dev(torch, pint, kb, 500 + 10 * gi + j).contiguous() for j in range(2)]
        f = wl.gaussian_sources(torch, pint, POOL, 700 + gi).contiguous()
        g = wl.plane_wave_g_dev(torch, t, xyz, face, kb, rng.uniform(0, 2 * np.pi, POOL)).contiguous()
        torch.cuda.synchronize()
        t.factor_device(ks[0].data_ptr())
        trees.append({"t": t, "ks": ks, "f": f, "g": g, "L": L, "n": n, "kb": kb,
                      "cf": t.n_leaves * t.n_int * 16, "cg": t.n_leaves * t.nb * 16})
        del xyz, face, pint
        print("  tree %d^2 L=%d factored" % (n, L), flush=True)

    nreq = 240
    jobs = []
    version = [0] * len(trees)
    for _ in range(nreq):
        ti = int(rng.integers(0, len(trees)))
        if rng.uniform() < 0.02:
            version[ti] += 1
            jobs.append({"kind": "factor", "ti": ti, "v": version[ti]})
            continue
        nr = int(rng.integers(1, 33))
        off = int(rng.integers(0, POOL - nr + 1))
        T = trees[ti]
        u = torch.zeros((nr, T["t"].n_leaves, T["t"].m), dtype=torch.complex128, device="cuda:0")
        jobs.append({"kind": "solve", "ti": ti, "v": version[ti], "nr": nr, "off": off, "u": u})
    torch.cuda.synchronize()
    print("  %d requests, %d refactorizations" % (nreq, sum(j["kind"] == "factor" for j in jobs)), flush=True)

    arrivals = np.cumsum(rng.exponential(1.0 / 1500.0, nreq))
    with Server(max_cols=256, lanes=2, group=group) as srv:
        ids = [srv.add_tree(T["t"]) for T in trees]
        rids = []
        t0 = time.perf_counter()
        for i, j in enumerate(jobs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            T = trees[j["ti"]]
            if j["kind"] == "factor":
                rids.append(srv.factor(ids[j["ti"]], T["ks"][j["v"] % 2].data_ptr()))
            else:
                rids.append(srv.solve(ids[j["ti"]], j["nr"], T["f"].data_ptr() + j["off"] * T["cf"],
                                      T["g"].data_ptr() + j["off"] * T["cg"], j["u"].data_ptr()))
        srv.drain()
        for r in rids:
            assert srv.state(r) == RS_REQ_DONE, srv.error(r)
        st = srv.stats()
        print("  served: %d passes, %.1f cols per pass, p99 %.1f ms, %d factors" %
              (st["passes"], st["mean_cols"], 1e3 * st["lat_p99_s"], st["factors_done"]), flush=True)
        for r in rids:
            srv.release(r)

    # alone: each request's columns on its own, with the potential current at its arrival
    worst = 0.0
    for ti, T in enumerate(trees):
        mine = [j for j in jobs if j["kind"] == "solve" and j["ti"] == ti]
        for v in sorted({j["v"] for j in mine}):
            T["t"].factor_device(T["ks"][v % 2].data_ptr())
            t64 = rsa.GpuTree(2, T["L"], UNIT, 1.0 / T["n"], -T["kb"], precision=rsa.RS_FP64)
            t64.factor_device(T["ks"][v % 2].data_ptr())
            for j in (j for j in mine if j["v"] == v):
                a, b = j["off"], j["off"] + j["nr"]
                fa, ga = T["f"][a:b].contiguous(), T["g"][a:b].contiguous()
                ua = torch.zeros_like(j["u"])
                u64 = torch.zeros_like(j["u"])
                T["t"].solve_device(j["nr"], fa.data_ptr(), ga.data_ptr(), ua.data_ptr())
                t64.solve_device(j["nr"], fa.data_ptr(), ga.data_ptr(), u64.data_ptr())
                torch.cuda.synchronize()
                assert wl.same_bits(ua, j["u"]), ("tree", T["n"], "version", v, "cols", a, b)
                rel = _rel(j["u"], u64)
                if rel > 1e-8:
                    # which side is off: the FD residual of each solution
                    r64 = t64.residual_norms(j["nr"], fa.data_ptr(), ga.data_ptr(), u64.data_ptr())
                    rmx = T["t"].residual_norms(j["nr"], fa.data_ptr(), ga.data_ptr(), j["u"].data_ptr())
                    print("  tree %d^2 version %d cols %d:%d rel %.2e; FD residual FP64 %.2e, served %.2e" %
                          (T["n"], v, a, b, rel, r64.max(), rmx.max()), flush=True)
                worst = max(worst, rel)
            t64.close()
            torch.cuda.synchronize()
    print("  every request bitwise equal alone; max rel l2 vs FP64 %.2e" % worst, flush=True)
    assert worst <= 1e-8, worst
