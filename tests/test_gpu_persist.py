"""GPU: factored-tree persistence (rs_tree_save / rs_tree_load).  A tree read
back from its file solves with the bits of the tree that was saved, in every
precision the file holds, and refactors in place afterwards like any tree; a
file whose sections do not match its plan is refused and leaves the handle
unfactored."""
import os
import sys

import numpy as np
import pytest

import robinsolver_amd as rsa

pytestmark = pytest.mark.gpu

UNIT = [[0, 1], [0, 1], [0, 1]]


def _case(tag):
    print("case:", tag, flush=True)
    sys.stdout.flush()


def _k(pts_int, kb, shift=0.0):
    x = pts_int
    return kb ** 2 * (1 + 0.3 * np.sin(3 * x[..., 0] + 1 + shift) * np.cos(2 * x[..., 1])) + 0.05j * kb ** 2


def _rhs(t, nrhs, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nrhs, t.n_leaves, t.n_int)) + 1j * rng.standard_normal((nrhs, t.n_leaves, t.n_int))
    g = rng.standard_normal((nrhs, t.n_leaves, t.nb)) + 1j * rng.standard_normal((nrhs, t.n_leaves, t.nb))
    return f, g


@pytest.mark.parametrize("d,n,L,kb,prec,retain", [
    (2, 64, 4, 10.0, rsa.RS_FP64, 0),
    (2, 64, 4, 10.0, rsa.RS_MIXED, 0),
    (2, 128, 6, 20.0, rsa.RS_MIXED, rsa.RS_RETAIN_FP64),
    (3, 16, 3, 4.0, rsa.RS_MIXED, 0),
])
def test_loaded_tree_solves_with_the_saved_bits(d, n, L, kb, prec, retain, tmp_path, gpu):
    _case("persist %dD %d^%d L=%d prec=%d retain=%d" % (d, n, d, L, prec, retain))
    # (FP64 solves on a mixed tree need its factorization in FP64: merge_fp64_levels -1)
    lv = -1 if retain & rsa.RS_RETAIN_FP64 else 0
    t = rsa.GpuTree(d, L, UNIT, 1.0 / n, kb, precision=prec, retain=retain, merge_fp64_levels=lv)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], kb))
    f, g = _rhs(t, 3, 7)
    precs = [prec] + ([rsa.RS_FP64] if retain & rsa.RS_RETAIN_FP64 else [])
    ref = {p: t.solve(f, g, precision=p) for p in precs}
    path = str(tmp_path / "tree.rst")
    t.save(path)
    hdr = rsa.read_saved_header(path)
    assert (hdr["d"], hdr["L"], hdr["precision"]) == (d, L, prec)
    assert hdr["have32"] == (prec == rsa.RS_MIXED)
    assert hdr["have64"] == (prec == rsa.RS_FP64 or bool(retain & rsa.RS_RETAIN_FP64))
    t2 = rsa.GpuTree.load(path, merge_fp64_levels=lv)
    assert (t2.n_leaves, t2.m, t2.n_int, t2.nb) == (t.n_leaves, t.m, t.n_int, t.nb)
    for p in precs:
        assert np.array_equal(t2.solve(f, g, precision=p), ref[p]), p
    # a refactor of the loaded tree (new potential) matches a fresh factorization
    k2 = _k(pts[:, t.nb:], kb, shift=0.5)
    t2.factor(k2)
    t3 = rsa.GpuTree(d, L, UNIT, 1.0 / n, kb, precision=prec, retain=retain, merge_fp64_levels=lv)
    t3.factor(k2)
    assert np.array_equal(t2.solve(f, g), t3.solve(f, g))


def test_truncated_file_is_refused(tmp_path, gpu):
    _case("persist truncated")
    t = rsa.GpuTree(2, 2, UNIT, 1.0 / 32, 5.0)
    pts, _ = t.all_points()
    t.factor(_k(pts[:, t.nb:], 5.0))
    path = str(tmp_path / "tree.rst")
    t.save(path)
    size = os.path.getsize(path)
    cut = str(tmp_path / "cut.rst")
    with open(path, "rb") as a, open(cut, "wb") as b:
        b.write(a.read(size - 100))
    t2 = rsa.GpuTree(2, 2, UNIT, 1.0 / 32, 5.0)
    with pytest.raises(rsa.RobinError, match="ends early"):
        t2._check(rsa.lib().rs_tree_load(t2.h, cut.encode()))
    f, g = _rhs(t2, 1, 3)
    with pytest.raises(rsa.RobinError):
        t2.solve(f, g)
