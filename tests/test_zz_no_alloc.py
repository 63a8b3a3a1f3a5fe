"""GPU: rs_solve_device_noalloc, the allocation-free variant of rs_solve_device
(its pass workspace in the library's fixed device scratch array), against
rs_solve_device bit for bit: two calls in a row, the second with fewer
right-hand sides (the array keeps the first call's contents), in both
precisions and in 2D and 3D; and a solve too large for the array is refused
with a message before anything is launched."""
import ctypes

import numpy as np
import pytest

import robinsolver_amd as rsa
from robinsolver_amd import workloads as wl

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _torch():
    import torch
    return torch


def _noalloc(t, nrhs, f, g, u, prec):
    L = rsa.lib()
    vp = ctypes.c_void_p
    L.rs_solve_device_noalloc.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
    L.rs_solve_device_noalloc.restype = ctypes.c_int
    return L.rs_solve_device_noalloc(t.h, nrhs, vp(f.data_ptr()), vp(g.data_ptr()), vp(u.data_ptr()), prec, None)


def _k(pts_int, kb):
    x = pts_int
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1) * np.cos(2 * x[..., 1])) + 0.02j * kb ** 2


def _cdev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, np.complex128)).to("cuda:0")


@pytest.mark.parametrize("d,n,L,kb,prec", [(2, 128, 6, 20.0, rsa.RS_MIXED), (2, 128, 6, 20.0, rsa.RS_FP64),
                                           (3, 16, 6, 4.0, rsa.RS_MIXED), (3, 16, 6, 4.0, rsa.RS_FP64)])
def test_noalloc_matches_solve_device(d, n, L, kb, prec, gpu):
    torch = _torch()
    print("case: noalloc", d, n, L, prec, flush=True)
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, -kb, precision=prec, n_refine=1)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], kb))
    rng = np.random.default_rng(n + d)
    nr = 8
    f = _cdev(rng.standard_normal((nr, t.n_leaves, t.n_int)) + 1j * rng.standard_normal((nr, t.n_leaves, t.n_int)))
    g = _cdev(rng.standard_normal((nr, t.n_leaves, t.nb)) + 1j * rng.standard_normal((nr, t.n_leaves, t.nb)))
    shape = (nr, t.n_leaves, t.m)
    # two allocation-free calls in a row, the second with three of the columns
    u1 = torch.full(shape, float("nan"), dtype=torch.complex128, device="cuda:0")
    u2 = torch.full((3,) + shape[1:], float("nan"), dtype=torch.complex128, device="cuda:0")
    f2, g2 = f[5:8].contiguous(), g[5:8].contiguous()
    torch.cuda.synchronize()
    assert _noalloc(t, nr, f, g, u1, prec) == 0, rsa.lib().rs_last_error(t.h)
    assert _noalloc(t, 3, f2, g2, u2, prec) == 0, rsa.lib().rs_last_error(t.h)
    # the existing entry point on the same inputs
    r1 = torch.zeros(shape, dtype=torch.complex128, device="cuda:0")
    r2 = torch.zeros((3,) + shape[1:], dtype=torch.complex128, device="cuda:0")
    t.solve_device(nr, f.data_ptr(), g.data_ptr(), r1.data_ptr(), precision=prec)
    t.solve_device(3, f2.data_ptr(), g2.data_ptr(), r2.data_ptr(), precision=prec)
    torch.cuda.synchronize()
    assert wl.same_bits(u1, r1)
    assert wl.same_bits(u2, r2)
    # and the columns of the smaller call are those of the larger one
    assert wl.same_bits(u2, u1[5:8])


def test_noalloc_refuses_a_solve_beyond_the_scratch(gpu):
    torch = _torch()
    print("case: noalloc too large", flush=True)
    t = rsa.GpuTree(2, 10, UNIT, 1.0 / 512, -20.0, precision=rsa.RS_MIXED, n_refine=1)
    xyz = torch.empty((t.n_leaves, t.m, 3), dtype=torch.float64, device="cuda:0")
    t.points_device(xyz.data_ptr())
    k = (400.0 * (1 + 0.1 * torch.cos(3 * xyz[:, t.nb:, 0]))).to(torch.complex128).contiguous()
    torch.cuda.synchronize()
    t.factor_device(k.data_ptr())
    nr = 64   # about 15 MB of workspace per right-hand side at 512^2: about 1 GB in all
    f = torch.zeros((nr, t.n_leaves, t.n_int), dtype=torch.complex128, device="cuda:0")
    g = torch.zeros((nr, t.n_leaves, t.nb), dtype=torch.complex128, device="cuda:0")
    u = torch.zeros((nr, t.n_leaves, t.m), dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    assert _noalloc(t, nr, f, g, u, rsa.RS_MIXED) == 1
    assert b"bytes of device scratch" in rsa.lib().rs_last_error(t.h)
    # a width that fits runs
    assert _noalloc(t, 4, f, g, u, rsa.RS_MIXED) == 0, rsa.lib().rs_last_error(t.h)
    torch.cuda.synchronize()
