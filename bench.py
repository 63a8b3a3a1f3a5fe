"""Benchmark: solves/s and factor seconds of the HPS Helmholtz solver on one
MI355X, 2D 1024^2 (L=12, 16x16 leaves), kb = 80, seeded random potential,
256 Gaussian-source right-hand sides; mixed precision (complex64 solve passes)
against the plain FP64 GPU path.

One "step" is one multi-RHS solve of the whole batch (up pass + down pass over
all 12 levels) with every input already resident in HBM.  `value` is the
mixed-path throughput in solves/s.  Beside it: factor seconds (FP64 factor;
the mixed path adds the complex64 conversion), the plain FP64 path (one RHS
per call, the reference's semantics), the batched FP64 path, the rel. l2
error of every mixed RHS against the FP64 solution, the roofline of the
dominant kernel (HIP events inside the library, on the stream the kernels run
on) and the CPU oracle timed on this host on the same tree (its factor and two
of the right-hand sides), whose solutions also give the GPU paths' rel. l2
error against the oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from robinsolver_amd.workloads import (MIX_TREES, device_points, gaussian_sources, plane_wave_g_dev,  # noqa: E402
                                      potential_dev, rel_l2_cols, same_bits)

PEAK = {"hbm_GBs": 8000.0, "fp64_TF": 78.6, "fp32_TF": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class URing:
    """Output columns of one tree's requests: a ring of `cap` solution columns.
    A request takes the next free run of columns; before a run is reused the
    requests that held it must have completed (their u is the caller's data
    until then)."""

    def __init__(self, torch, tree, cap):
        self.u = torch.zeros((cap, tree.n_leaves, tree.m), dtype=torch.complex128, device="cuda:0")
        self.cap, self.pos, self.col_bytes = cap, 0, tree.n_leaves * tree.m * 16
        self.out = []                   # (request id, start, end), oldest first
        self.owner = [-1] * cap
        self.waits = 0

    def take(self, srv, nrhs, job):
        if self.pos + nrhs > self.cap:
            self.pos = 0
        s, e = self.pos, self.pos + nrhs
        while any(a < e and s < b for _, a, b in self.out):
            rid, _, _ = self.out.pop(0)
            if srv.state(rid) < 2:          # still queued or running: the client blocks
                self.waits += 1
            srv.wait(rid)
        self.pos = e
        for c in range(s, e):
            self.owner[c] = job
        return s

    def ptr(self, col):
        return self.u.data_ptr() + col * self.col_bytes


def drive(srv, jobs, rate, seed):
    """Submit jobs (dicts: tree id, kind, nrhs, input pointers, ring) at Poisson
    arrival times of the given rate (None: as fast as the output rings allow),
    then drain.  Returns wall seconds from the first submission to the last
    completion and the request ids."""
    rng = np.random.default_rng(seed)
    t_arr = np.cumsum(rng.exponential(1.0 / rate, len(jobs))) if rate else np.zeros(len(jobs))
    ids = []
    t0 = time.perf_counter()
    for i, jb in enumerate(jobs):
        if rate:
            dt = t0 + t_arr[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        if jb["kind"] == "factor":
            rid = srv.factor(jb["tree"], jb["k"])
        else:
            ring = jb["ring"]
            c0 = ring.take(srv, jb["nrhs"], i)
            jb["ucol"] = c0
            rid = srv.solve(jb["tree"], jb["nrhs"], jb["f"], jb["g"], ring.ptr(c0))
            ring.out.append((rid, c0, c0 + jb["nrhs"]))
        ids.append(rid)
    srv.drain()
    for jb in jobs:
        if "ring" in jb:
            jb["ring"].out.clear()       # every request has completed
    t_end = max(srv.info(r)[2] for r in ids)
    t_first = min(srv.info(r)[0] for r in ids)
    return t_end - t_first, ids


def serve_cfg3(torch, rsa, t, f, g, u32, u64, args, sz_f, sz_g):
    """The served path on the same tree and right-hand sides: requests of 1..32
    columns of the 256-source batch (random offsets), a saturated client (at
    most the output ring's columns outstanding) for sustained solves/s, then a
    Poisson client at 70% of that rate for p50/p99 latency.  Served results of
    requests whose output columns were not reused are compared bit for bit
    with the batched results of the same columns."""
    from robinsolver_amd.serve import Server
    rng = np.random.default_rng(args.seed + 17)
    out = {}
    with Server(max_cols=args.nrhs, lanes=2) as srv:
        tid = srv.add_tree(t)
        ring = URing(torch, t, 3 * args.nrhs)

        def jobs(n):
            js = []
            for _ in range(n):
                nr = int(rng.integers(1, 33))
                off = int(rng.integers(0, args.nrhs - nr + 1))
                js.append({"tree": tid, "kind": "solve", "nrhs": nr, "off": off, "ring": ring,
                           "f": f.data_ptr() + off * sz_f * 16, "g": g.data_ptr() + off * sz_g * 16})
            return js

        drive(srv, jobs(40), None, 1)            # warm: workspaces, both lanes
        for name, frac in (("saturated", None), ("poisson70", 0.7)):
            js = jobs(args.serve_requests)
            srv.reset_stats()
            rate = None if frac is None else frac * out["saturated"]["solves_per_s"] / 16.5
            wall, ids = drive(srv, js, rate, 2)
            st = srv.stats()
            rhs = sum(j["nrhs"] for j in js)
            # bitwise: served columns against the batched mixed result
            same = tot = 0
            worst = 0.0
            for i, j in enumerate(js):
                if any(ring.owner[c] != i for c in range(j["ucol"], j["ucol"] + j["nrhs"])):
                    continue
                us = ring.u[j["ucol"]:j["ucol"] + j["nrhs"]]
                ub = u32[j["off"]:j["off"] + j["nrhs"]]
                tot += 1
                same += int(same_bits(us, ub))
                ref = u64[j["off"]:j["off"] + j["nrhs"]]
                worst = max(worst, float(rel_l2_cols(us, ref).max()))
            out[name] = {"requests": len(js), "rhs": rhs, "wall_s": wall, "solves_per_s": rhs / wall,
                         "offered_requests_per_s": rate, "lat_p50_ms": 1e3 * st["lat_p50_s"],
                         "lat_p99_ms": 1e3 * st["lat_p99_s"], "lat_max_ms": 1e3 * st["lat_max_s"],
                         "passes": st["passes"], "mean_cols_per_pass": st["mean_cols"],
                         "busy_s": st["busy_s"], "ring_waits": ring.waits, "failed": st["failed"],
                         "checked_requests": tot, "bitwise_equal_to_batched": same,
                         "max_rel_l2_vs_fp64": worst}
            log("served %s: %.1f solves/s, p50 %.1f ms, p99 %.1f ms, %d passes (%.1f cols), "
                "%d/%d bitwise equal, max rel l2 %.2e" %
                (name, rhs / wall, 1e3 * st["lat_p50_s"], 1e3 * st["lat_p99_s"], st["passes"], st["mean_cols"],
                 same, tot, worst))
            for r in ids:
                srv.release(r)
    return out


def serve_mix(args):
    """Eight resident trees (2D 128^2 .. 1024^2, leaves near 16^2, kb by grid),
    requests of 1..32 right-hand sides (Gaussian sources plus the impedance data
    of a random-angle plane wave) on a uniformly chosen tree,
    2% of requests refactorizations with a new potential; a saturated client
    for sustained solves/s, then Poisson arrivals at 70% of it for latency."""
    import torch
    import robinsolver_amd as rsa
    from robinsolver_amd.serve import Server, lane_bytes
    dev = torch.device("cuda:0")
    unit = [[0, 1], [0, 1], [0, 1]]
    rng = np.random.default_rng(args.seed + 5)
    trees, info = [], []
    for gi, (n, L, kb) in enumerate(MIX_TREES):
        t = rsa.GpuTree(2, L, unit, 1.0 / n, -kb, precision=rsa.RS_MIXED, n_refine=1)
        xyz, face = device_points(torch, t)
        ks = [potential_dev(torch, xyz[:, t.nb:], kb, args.seed + 100 * gi + j).contiguous() for j in range(2)]
        torch.cuda.synchronize()
        t.factor_device(ks[0].data_ptr())
        gpool = plane_wave_g_dev(torch, t, xyz, face, kb, rng.uniform(0, 2 * np.pi, 32))
        fpool = gaussian_sources(torch, xyz[:, t.nb:, :], 32, args.seed + 100 * gi).contiguous()
        # accuracy of one refinement step on this tree, against its FP64 solve
        t64 = rsa.GpuTree(2, L, unit, 1.0 / n, -kb, precision=rsa.RS_FP64)
        t64.factor_device(ks[0].data_ptr())
        ua = torch.zeros((4, t.n_leaves, t.m), dtype=torch.complex128, device=dev)
        ub = torch.zeros_like(ua)
        torch.cuda.synchronize()
        t.solve_device(4, fpool.data_ptr(), gpool.data_ptr(), ua.data_ptr())
        t64.solve_device(4, fpool.data_ptr(), gpool.data_ptr(), ub.data_ptr())
        torch.cuda.synchronize()
        rel = float(rel_l2_cols(ua, ub).max())
        if rel > 1e-8:
            t.set_refine(2)
        del t64, ua, ub
        r, p, w = t.memory_estimate(args.nrhs)
        arena, descs = lane_bytes(t, args.nrhs)
        info.append({"grid": n, "L": L, "kb": kb, "leaf": int(round((t.n_int) ** 0.5)), "rel_l2_1step": rel,
                     "factor_s": t.timings()[0], "resident_GB": r / 1e9, "lane_arena_GB": arena / 1e9,
                     "descriptors_MB_per_lane": descs / 1e6})
        trees.append((t, ks, fpool, gpool))
        log("tree %d^2 L=%d: factor %.3fs, rel l2 (1 step) %.2e" % (n, L, t.timings()[0], rel))
    out = {"trees": info}
    # the same request streams (same seeds) through both pass forms: one tree
    # per pass on two lanes, and passes grouped across up to 8 trees
    for mode, group in (("per_tree", 1), ("grouped", 8)):
        out[mode] = serve_mix_run(args, torch, Server, trees, group)
        log("mix %s: held %.2f GB" % (mode, out[mode]["held_GB"]))
    best = max(("per_tree", "grouped"), key=lambda m: out[m]["saturated"]["solves_per_s"])
    out["best"] = best
    for k in ("saturated", "poisson70", "held_GB"):
        out[k] = out[best][k]
    return out


def serve_mix_run(args, torch, Server, trees, group):
    """One pass form of the serving scenario over the resident trees."""
    rng = np.random.default_rng(args.seed + 6)
    out = {"group": group}
    with Server(max_cols=args.nrhs, lanes=2, group=group) as srv:
        ids = [srv.add_tree(t) for t, _, _, _ in trees]
        rings = [URing(torch, t, args.nrhs) for t, _, _, _ in trees]
        col = [(t.n_leaves * t.n_int * 16, t.n_leaves * t.nb * 16) for t, _, _, _ in trees]
        nfac = [0] * len(trees)

        def jobs(n):
            js = []
            for _ in range(n):
                ti = int(rng.integers(0, len(trees)))
                if rng.uniform() < 0.02:
                    nfac[ti] += 1
                    js.append({"tree": ids[ti], "kind": "factor", "k": trees[ti][1][nfac[ti] % 2].data_ptr()})
                    continue
                nr = int(rng.integers(1, 33))
                off = int(rng.integers(0, 32 - nr + 1))
                js.append({"tree": ids[ti], "ti": ti, "kind": "solve", "nrhs": nr, "ring": rings[ti],
                           "f": trees[ti][2].data_ptr() + off * col[ti][0],
                           "g": trees[ti][3].data_ptr() + off * col[ti][1]})
            return js

        drive(srv, jobs(60), None, 3)
        for name, frac in (("saturated", None), ("poisson70", 0.7)):
            js = jobs(args.mix_requests)
            srv.reset_stats()
            nreq = len(js)
            rate = None if frac is None else frac * out["saturated"]["requests_per_s"]
            wall, rids = drive(srv, js, rate, 4)
            st = srv.stats()
            rhs = sum(j.get("nrhs", 0) for j in js)
            per = {}
            for j, r in zip(js, rids):
                if j["kind"] != "solve":
                    continue
                a, _, c, _ = srv.info(r)
                per.setdefault(MIX_TREES[j["ti"]][0], []).append(c - a)
            out[name] = {"requests": nreq, "factor_requests": sum(1 for j in js if j["kind"] == "factor"),
                         "rhs": rhs, "wall_s": wall, "solves_per_s": rhs / wall, "requests_per_s": nreq / wall,
                         "offered_requests_per_s": rate, "lat_p50_ms": 1e3 * st["lat_p50_s"],
                         "lat_p99_ms": 1e3 * st["lat_p99_s"], "lat_max_ms": 1e3 * st["lat_max_s"],
                         "passes": st["passes"], "mean_cols_per_pass": st["mean_cols"], "failed": st["failed"],
                         "factors_done": st["factors_done"], "admission_waits": st["admission_waits"],
                         "ring_waits": sum(r.waits for r in rings),
                         "p99_ms_by_grid": {str(k): 1e3 * float(np.quantile(v, 0.99)) for k, v in sorted(per.items())}}
            log("mix %s: %.1f solves/s, %.1f req/s, p50 %.1f ms, p99 %.1f ms, %d passes (%.1f cols), %d factors" %
                (name, rhs / wall, nreq / wall, 1e3 * st["lat_p50_s"], 1e3 * st["lat_p99_s"], st["passes"],
                 st["mean_cols"], st["factors_done"]))
            for r in rids:
                srv.release(r)
        out["held_GB"] = srv.stats()["held_bytes"] / 1e9
    for t, ks, _, _ in trees:            # the next run starts from the first potentials
        t.factor_device(ks[0].data_ptr())
    return out


def cpu_baseline(args, kb_opt, k_dev, f, u64, u32, ncols=2):
    """The oracle (the reference algorithm restated in C: dense leaf LU, zgetrf /
    zgetrs / zgemm semantics, blocked and threaded over this host's cores) on
    the full configuration: factor, then `ncols` of the batch's right-hand
    sides, one solve each (the reference's one-RHS solve), timed.  Their
    solutions are the parity check of the GPU's FP64 and mixed solutions of the
    same columns."""
    from oracle.pyoracle import OracleTree
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    threads = max(1, min(avail, int(os.environ.get("OMP_NUM_THREADS", avail))))
    o = OracleTree(args.dim, args.levels, [[0, 1], [0, 1], [0, 1]], 1.0 / args.grid, kb_opt, threads=threads)
    t0 = time.perf_counter()
    o.factor(k_dev.cpu().numpy())
    tf = time.perf_counter() - t0
    fh = f[:ncols].cpu().numpy()
    g0 = np.zeros((o.n_leaves, o.nb), np.complex128)
    ts, uo = [], []
    for r in range(ncols):
        t0 = time.perf_counter()
        uo.append(o.solve(fh[r], g0))
        ts.append(time.perf_counter() - t0)
    o.close()
    uo = np.stack(uo).reshape(ncols, -1)

    def rel(u):
        a = u[:ncols].cpu().numpy().reshape(ncols, -1)
        return (np.linalg.norm(a - uo, axis=1) / np.linalg.norm(uo, axis=1)).tolist()

    r64, rmx = rel(u64), rel(u32)
    log("oracle: factor %.2f s, solve %.3f s/RHS (%d threads); rel l2 vs oracle: fp64 %s, mixed %s"
        % (tf, float(np.mean(ts)), threads, r64, rmx))
    return {
        "value": ncols / sum(ts),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": ("the oracle on the full %d^2 L=%d tree: factor (%.2f s), then %d of the %d right-hand sides, "
                   "one solve each (%s s); %d threads, nproc %d"
                   % (args.grid, args.levels, tf, ncols, args.nrhs, ", ".join("%.3f" % x for x in ts), threads,
                      nproc)),
        "factor_s": tf,
        "solve_s": ts,
        "nproc": nproc,
        "fp64_rel_l2_vs_oracle": r64,
        "mixed_rel_l2_vs_oracle": rmx,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--levels", type=int, default=12)
    ap.add_argument("--kb", type=float, default=80.0)
    ap.add_argument("--nrhs", type=int, default=256)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--dim", type=int, default=2, choices=(2, 3),
                    help="3: the 3D workload (e.g. --dim 3 --grid 64 --levels 9 --kb 8 --nrhs 16)")
    ap.add_argument("--r1-solves", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=600,
                    help="requests per phase of the single-tree serving scenario")
    ap.add_argument("--mix-requests", type=int, default=2000,
                    help="requests per phase of the 8-tree serving scenario (--serve-mix)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-mix", action="store_true",
                    help="only the multi-tree serving scenario (8 resident trees 128^2..1024^2)")
    ap.add_argument("--r1-only", action="store_true",
                    help="only the FP64 one-RHS solves (--r1-solves of them) of the 2D configuration")
    ap.add_argument("--factor-only", type=int, default=0, metavar="REPS",
                    help="only time REPS factorizations of the workload tree (FP64), one JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)." % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.factor_only:
        import robinsolver_amd as rsa
        import torch
        t = rsa.GpuTree(2, args.levels, [[0, 1], [0, 1], [0, 1]], 1.0 / args.grid, -args.kb, precision=rsa.RS_FP64)
        xyz, _ = device_points(torch, t)
        k_dev = potential_dev(torch, xyz[:, t.nb:], args.kb, args.seed).contiguous()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.factor_only):
            t.factor_device(k_dev.data_ptr())
            times.append(t.timings()[0])
            log("factor %.4f s" % times[-1])
        ff, _ = t.algorithmic_counts()
        print(json.dumps({"metric": "factor seconds, 2D %d^2 Helmholtz (FP64)" % args.grid, "value": min(times),
                          "unit": "s", "higher_is_better": False, "times": times,
                          "reference_factor_GF": ff / 1e9}))
        return

    if args.r1_only:
        # the reference's call shape alone: an FP64 tree, one right-hand side per
        # solve (for kernel traces of that path)
        import robinsolver_amd as rsa
        import torch
        t = rsa.GpuTree(2, args.levels, [[0, 1], [0, 1], [0, 1]], 1.0 / args.grid, -args.kb, precision=rsa.RS_FP64)
        xyz, _ = device_points(torch, t)
        k_dev = potential_dev(torch, xyz[:, t.nb:], args.kb, args.seed).contiguous()
        n = args.r1_solves
        f = gaussian_sources(torch, xyz[:, t.nb:, :], n, args.seed).contiguous()
        g = torch.zeros((n, t.n_leaves, t.nb), dtype=torch.complex128, device="cuda:0")
        u = torch.zeros((n, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
        t.factor_device(k_dev.data_ptr())
        sf, sg, su = t.n_leaves * t.n_int * 16, t.n_leaves * t.nb * 16, t.n_leaves * t.m * 16
        for _ in range(2):
            t.solve_device(1, f.data_ptr(), g.data_ptr(), u.data_ptr(), precision=rsa.RS_FP64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(n):
            t.solve_device(1, f.data_ptr() + r * sf, g.data_ptr() + r * sg, u.data_ptr() + r * su,
                           precision=rsa.RS_FP64)
        torch.cuda.synchronize()
        rate = n / (time.perf_counter() - t0)
        log("fp64 R=1: %.1f solves/s" % rate)
        # a plain copy of known size, the calibration point for hardware byte counters
        cal = torch.empty(1 << 25, dtype=torch.float64, device="cuda:0")   # 256 MiB
        torch.empty_like(cal).copy_(cal)
        torch.cuda.synchronize()
        print(json.dumps({"metric": "FP64 one-RHS solves/s, 2D %d^2 Helmholtz" % args.grid, "value": rate,
                          "unit": "solves/s", "higher_is_better": True, "solves": n,
                          "counter_calibration_copy_bytes": cal.numel() * 8}))
        return

    if args.serve_mix:
        mix = serve_mix(args)
        sat = mix["saturated"]
        print(json.dumps({
            "metric": "served solves/s and latency, 8 resident 2D Helmholtz trees 128^2..1024^2 (mixed precision)",
            "value": sat["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "c64", "data": "synthetic",
            "config": {"workload": "8 trees (grid, L, kb) = %s; requests of 1..32 RHS (Gaussian sources + "
                                   "plane-wave g), 2%% refactors"
                                   % [list(x) for x in MIX_TREES], "max_cols": args.nrhs, "lanes": 2,
                       "requests": args.mix_requests, "pass_form": mix["best"]},
            "solves_per_s_by_pass_form": {m: mix[m]["saturated"]["solves_per_s"] for m in ("per_tree", "grouped")},
            "p99_ms_by_pass_form": {m: mix[m]["poisson70"]["lat_p99_ms"] for m in ("per_tree", "grouped")},
            "p50_ms": mix["poisson70"]["lat_p50_ms"], "p99_ms": mix["poisson70"]["lat_p99_ms"], "mix": mix}))
        return

    import torch
    import robinsolver_amd as rsa

    dev = torch.device("cuda:0")
    kb_phys = args.kb
    kb_opt = -args.kb          # Robin a1 = -i kb realised as opts%kb = -kb
    unit = [[0, 1], [0, 1], [0, 1]]
    d = args.dim
    if d == 3:
        args.no_cpu_baseline = True    # the CPU sample and the serving scenarios are 2D
        args.no_serve = True
    log("plan: %dD %d^%d, L=%d, kb=%g, nrhs=%d" % (d, args.grid, d, args.levels, kb_phys, args.nrhs))
    t = rsa.GpuTree(d, args.levels, unit, 1.0 / args.grid, kb_opt, precision=rsa.RS_MIXED)
    # inputs sampled on the GPU at the device coordinates of the points
    xyz, _ = device_points(torch, t)
    pint_dev = xyz[:, t.nb:, :]
    k_dev = potential_dev(torch, pint_dev, kb_phys, args.seed, nmodes=16 if d == 2 else 24, d=d,
                          absorb=0.01 if d == 2 else 0.0).contiguous()
    f = gaussian_sources(torch, pint_dev, args.nrhs, args.seed, sigma=0.02 if d == 2 else 0.05, d=d)
    g = torch.zeros((args.nrhs, t.n_leaves, t.nb), dtype=torch.complex128, device=dev)
    u64 = torch.zeros((args.nrhs, t.n_leaves, t.m), dtype=torch.complex128, device=dev)
    u32 = torch.zeros_like(u64)
    torch.cuda.synchronize()

    # factor (FP64 interface inverses; the mixed tree also converts the solve data)
    log("factor ...")
    t.factor_device(k_dev.data_ptr())
    f_time_mixed = t.timings()[0]
    t.factor_device(k_dev.data_ptr())
    f_time_mixed = min(f_time_mixed, t.timings()[0])
    t64 = rsa.GpuTree(d, args.levels, unit, 1.0 / args.grid, kb_opt, precision=rsa.RS_FP64)
    t64.factor_device(k_dev.data_ptr())
    f_time_64 = t64.timings()[0]
    t64.factor_device(k_dev.data_ptr())
    f_time_64 = min(f_time_64, t64.timings()[0])
    # a refactorization in place (new potential, same tree): the tree keeps its
    # factorization working memory (rs_opts.retain = RS_RETAIN_SCRATCH, as the
    # server's resident trees do), so nothing is allocated or freed
    t64r = rsa.GpuTree(d, args.levels, unit, 1.0 / args.grid, kb_opt, precision=rsa.RS_FP64,
                       retain=rsa.RS_RETAIN_SCRATCH)
    t64r.factor_device(k_dev.data_ptr())
    refactor_64 = []
    for _ in range(3):
        t64r.factor_device(k_dev.data_ptr())
        refactor_64.append(t64r.timings()[0])
    refactor_64 = min(refactor_64)
    del t64r
    log("factor s: fp64 %.4f, mixed %.4f; fp64 refactor in place %.4f" % (f_time_64, f_time_mixed, refactor_64))

    def solve(tree, nrhs, fp, gp, up, prec):
        tree.solve_device(nrhs, fp, gp, up, precision=prec)

    sz_f = t.n_leaves * t.n_int
    sz_g = t.n_leaves * t.nb
    sz_u = t.n_leaves * t.m

    # plain FP64 path: one RHS per call
    nr1 = min(args.r1_solves, args.nrhs)
    for r in range(2):
        solve(t64, 1, f.data_ptr(), g.data_ptr(), u64.data_ptr(), rsa.RS_FP64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(nr1):
        solve(t64, 1, f.data_ptr() + r * sz_f * 16, g.data_ptr() + r * sz_g * 16, u64.data_ptr() + r * sz_u * 16,
              rsa.RS_FP64)
    torch.cuda.synchronize()
    r1_rate = nr1 / (time.perf_counter() - t0)
    log("fp64 R=1: %.1f solves/s" % r1_rate)
    # the same one-RHS solves with HIP events around every launch: algorithmic
    # bytes per solve (leaf blocks twice, inv(M) twice, T_NI and T_IN once per
    # level, the vectors) over the measured time per solve
    t64.profile(True)
    t64.profile_read()
    for r in range(nr1):
        solve(t64, 1, f.data_ptr() + r * sz_f * 16, g.data_ptr() + r * sz_g * 16, u64.data_ptr() + r * sz_u * 16,
              rsa.RS_FP64)
    torch.cuda.synchronize()
    prof1 = t64.profile_read()
    t64.profile(False)
    r1_bytes = sum(v[2] for v in prof1.values()) / nr1
    r1_launches = sum(v[1] for v in prof1.values()) / nr1
    r1_roof = {"bound": "hbm", "achieved": r1_bytes * r1_rate / 1e9, "peak": PEAK["hbm_GBs"], "unit": "GB/s",
               "algorithmic_bytes_per_solve": r1_bytes, "launches_per_solve": r1_launches,
               "kernel_ms_per_solve": {c: v[0] / nr1 for c, v in prof1.items() if v[1]}}
    r1_roof["frac"] = r1_roof["achieved"] / r1_roof["peak"]
    log("fp64 R=1 roofline: %.3f GB/solve, %.0f GB/s, frac %.3f, %.0f launches" %
        (r1_bytes / 1e9, r1_roof["achieved"], r1_roof["frac"], r1_launches))

    # batched FP64
    solve(t64, args.nrhs, f.data_ptr(), g.data_ptr(), u64.data_ptr(), rsa.RS_FP64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nb64 = max(1, args.steps // 2)
    for _ in range(nb64):
        solve(t64, args.nrhs, f.data_ptr(), g.data_ptr(), u64.data_ptr(), rsa.RS_FP64)
    torch.cuda.synchronize()
    b64_rate = nb64 * args.nrhs / (time.perf_counter() - t0)
    log("fp64 batched: %.1f solves/s" % b64_rate)

    # the drop-in's call shape with host buffers (rs_solve on host f and g, then
    # rs_get_solution per right-hand side): the mixed path including the PCIe
    # copies of f, g and u (pageable host memory); reported beside the headline,
    # which times device-resident inputs
    nh = min(32, args.nrhs)
    fh = f[:nh].cpu().numpy()
    gh = g[:nh].cpu().numpy()
    t.solve(fh, gh)
    t0 = time.perf_counter()
    uh = t.solve(fh, gh)
    host_rate = nh / (time.perf_counter() - t0)
    del uh
    log("mixed, host buffers (%d RHS per call): %.1f solves/s" % (nh, host_rate))

    # fewest refinement steps that bring every RHS to rel. l2 <= 1e-8 against
    # the FP64 GPU solution (checked on the whole batch)
    def max_rel(ua, ub):
        return rel_l2_cols(ua, ub)

    refine_err = {}
    n_ref = 2
    for nr in (0, 1, 2):
        t.set_refine(nr)
        solve(t, args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr(), rsa.RS_MIXED)
        torch.cuda.synchronize()
        refine_err[nr] = float(max_rel(u32, u64).max())
        log("n_refine=%d: max rel l2 %.3e" % (nr, refine_err[nr]))
    for nr in (0, 1, 2):
        if refine_err[nr] <= 1e-8:
            n_ref = nr
            break
    t.set_refine(n_ref)

    # the same mixed tree with its factorization in FP64 (merge_fp64_levels -1;
    # the default mixed tree runs its merge products in complex64): its time,
    # and the refinement it then needs (run before the timed steps, whose
    # launches end the kernel trace)
    tc = rsa.GpuTree(d, args.levels, unit, 1.0 / args.grid, kb_opt, precision=rsa.RS_MIXED, merge_fp64_levels=-1)
    tc.factor_device(k_dev.data_ptr())
    tc.factor_device(k_dev.data_ptr())
    f_time_f64m = tc.timings()[0]
    f64m_err = {}
    for nr in (0, 1, 2):
        tc.set_refine(nr)
        solve(tc, args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr(), rsa.RS_MIXED)
        torch.cuda.synchronize()
        f64m_err[nr] = float(max_rel(u32, u64).max())
    tc.close()
    f64m_need = min([nr for nr in f64m_err if f64m_err[nr] <= 1e-8], default=None)
    log("FP64 merges: factor %.4f s, max rel l2 by refinement steps %s" % (f_time_f64m, f64m_err))

    # mixed path, timed steps (replayed as a captured graph, no per-launch events)
    for _ in range(args.warmup):
        solve(t, args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr(), rsa.RS_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve(t, args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr(), rsa.RS_MIXED)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    mixed_rate = args.steps * args.nrhs / elapsed
    log("mixed: %.1f solves/s (%.2f ms/step)" % (mixed_rate, 1e3 * elapsed / args.steps))
    # the same steps again with HIP events around every launch (on the stream the
    # kernels run on), for the per-kernel times and the roofline
    nprof = args.steps
    t.profile(True)
    t.profile_read()
    for _ in range(nprof):
        solve(t, args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr(), rsa.RS_MIXED)
    torch.cuda.synchronize()
    prof = t.profile_read()
    t.profile(False)
    prof_ms = sum(v[0] for v in prof.values()) / nprof
    log("kernel time per step (events): %.2f ms" % prof_ms)

    # accuracy of every mixed RHS against the FP64 GPU solution
    rel = rel_l2_cols(u32, u64)

    # roofline of the dominant kernel category
    cat = max(prof, key=lambda c: prof[c][0])
    ms, nl, byt, fl = prof[cat]
    dur = ms / 1e3 / max(nl, 1)
    bpl, fpl = byt / max(nl, 1), fl / max(nl, 1)
    peak_f = PEAK["fp32_TF"]
    if fpl / max(bpl, 1.0) > peak_f * 1e12 / (PEAK["hbm_GBs"] * 1e9):
        roof = {"bound": "mfma", "achieved": fpl / dur / 1e12, "peak": peak_f, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": bpl / dur / 1e9, "peak": PEAK["hbm_GBs"], "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    roof["kernel"] = cat
    roof["avg_launch_ms"] = dur * 1e3
    roof["launches"] = nl
    roof["algorithmic_bytes_per_launch"] = bpl
    # HBM bytes per launch from the committed hardware-counter summary of this
    # workload (FETCH_SIZE and WRITE_SIZE runs, scaled by the copy calibration)
    tpath = os.path.join(ROOT, "profiles", "traffic_leaf_down_2d1024.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("category") == cat:
            roof["traffic"] = tr["hbm_bytes_per_launch"]
            roof["traffic_source"] = "profiles/traffic_leaf_down_2d1024.json"

    # every category: time per step and achieved rates against its algorithmic counts
    kernels = {}
    for c, (cms, cn, cb, cf) in prof.items():
        if cn == 0:
            continue
        kernels[c] = {"ms_per_step": cms / nprof, "launches_per_step": cn / nprof,
                      "GB_per_s": cb / (cms / 1e3) / 1e9 if cms > 0 else 0.0,
                      "TFLOP_per_s": cf / (cms / 1e3) / 1e12 if cms > 0 else 0.0}

    # FD residual of the global system, FP64 and mixed results
    res64 = t64.residual_norms(args.nrhs, f.data_ptr(), g.data_ptr(), u64.data_ptr())
    res32 = t.residual_norms(args.nrhs, f.data_ptr(), g.data_ptr(), u32.data_ptr())
    log("FD residual: fp64 max %.2e, mixed max %.2e" % (res64.max(), res32.max()))

    ff, fs = t.algorithmic_counts()
    out = {
        "metric": ("solves/s + factor s, 2D 1024^2 Helmholtz; mixed vs FP64 GPU at rel.l2<=1e-8" if d == 2 else
                   "solves/s + factor s, 3D %d^3 Helmholtz; mixed vs FP64 GPU at rel.l2<=1e-8" % args.grid),
        "value": mixed_rate,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "%dD %d^%d Helmholtz, L=%d (%s leaves), kb=%g, seeded %d-mode potential, "
                               "%d Gaussian-source RHS per step"
                               % (d, args.grid, d, args.levels, "x".join(str(x) for x in t.leaf_grid()), kb_phys,
                                  16 if d == 2 else 24, args.nrhs),
                   "dim": d, "grid": args.grid, "levels": args.levels, "kb": kb_phys, "nrhs": args.nrhs,
                   "seed": args.seed},
        "factor_s_fp64": f_time_64,
        "refactor_in_place_s_fp64": refactor_64,
        "factor_s_mixed": f_time_mixed,
        "fp64_r1_solves_per_s": r1_rate,
        "fp64_r1_roofline": r1_roof,
        "fp64_batched_solves_per_s": b64_rate,
        "host_buffers_solves_per_s": host_rate,
        "host_buffers_note": "mixed path, %d RHS per rs_solve call on pageable host arrays, u read back with "
                             "rs_get_solution: includes the PCIe copies of f, g and u" % nh,
        "mixed_over_fp64_r1": mixed_rate / r1_rate,
        "mixed_rel_l2_vs_fp64": {"max": float(rel.max()), "median": float(np.median(rel))},
        "refinement_steps": n_ref,
        "mixed_max_rel_l2_by_refinement_steps": refine_err,
        "fp64_merges": {"merge_fp64_levels": -1, "factor_s": f_time_f64m,
                             "max_rel_l2_by_refinement_steps": f64m_err, "refinement_steps_needed": f64m_need},
        "reference_flops": {"factor_GF": ff / 1e9, "solve_GF_per_rhs": fs / 1e9},
        "roofline": roof,
        "kernel_ms_per_step": {c: v[0] / nprof for c, v in prof.items()},
        "kernel_profile": "HIP events around every launch in %d separate steps after the timed ones" % nprof,
        "kernels": kernels,
        "gemm_tiles_mixed": t.tile_choices(rsa.RS_MIXED).tolist(),
        "fd_residual_fp64": {"max": float(res64.max()), "median": float(np.median(res64))},
        "fd_residual_mixed": {"max": float(res32.max()), "median": float(np.median(res32))},
    }
    if not args.no_serve:
        log("serving ...")
        sv = serve_cfg3(torch, rsa, t, f, g, u32, u64, args, sz_f, sz_g)
        out["served"] = sv
        out["served_solves_per_s"] = sv["saturated"]["solves_per_s"]
        out["served_over_fp64_r1"] = sv["saturated"]["solves_per_s"] / r1_rate
        out["served_p50_ms"] = sv["poisson70"]["lat_p50_ms"]
        out["served_p99_ms"] = sv["poisson70"]["lat_p99_ms"]

    # a plain copy of known size, the calibration point for hardware byte counters
    cal = torch.empty(1 << 25, dtype=torch.float64, device=dev)   # 256 MiB
    cal2 = torch.empty_like(cal)
    cal2.copy_(cal)
    torch.cuda.synchronize()
    out["counter_calibration_copy_bytes"] = cal.numel() * 8

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cb = cpu_baseline(args, kb_opt, k_dev, f, u64, u32)
        out["cpu_baseline"] = cb
        out["mixed_rel_l2_vs_oracle"] = max(cb["mixed_rel_l2_vs_oracle"])
        out["fp64_rel_l2_vs_oracle"] = max(cb["fp64_rel_l2_vs_oracle"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
