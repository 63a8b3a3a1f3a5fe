"""CPU checks of the C ABI library: it loads, exports every entry point the
header declares, and its host-side geometry (the reference's partition, leaf
grids, point coordinates and output order) agrees with the oracle.  No GPU
compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import robinsolver_amd as rsa
from oracle.pyoracle import OracleTree
from tests.conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "robinsolver.h")
UNIT = [[0, 1], [0, 1], [0, 1]]


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rs_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = rsa.lib()
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("d,n,Ln", [(2, 64, 4), (2, 24, 3), (3, 8, 3), (3, 12, 4), (2, 32, 0), (2, 48, 1)])
def test_plan_matches_oracle_geometry(d, n, Ln):
    o = OracleTree(d, Ln, UNIT, 1.0 / n, 5.0)
    t = rsa.GpuTree(d, Ln, UNIT, 1.0 / n, 5.0)
    assert (t.n_leaves, t.m, t.n_int, t.nb) == (o.n_leaves, o.m, o.n_int, o.nb)
    for q in range(t.n_leaves):
        x1, f1 = t.leaf_points(q)
        x2, f2 = o.leaf_points(q)
        assert np.array_equal(f1, f2)
        assert np.array_equal(x1, x2)
        assert np.array_equal(t.leaf_box(q), o.leaf_box(q))


def test_plan_errors_match_reference_messages():
    with pytest.raises(rsa.RobinError, match="Invalid operator domain"):
        rsa.GpuTree(2, 2, [[0, 1], [1, 0], [0, 1]], 0.1, 1.0)
    with pytest.raises(rsa.RobinError, match="Dimension must be 2 or 3"):
        rsa.GpuTree(4, 2, UNIT, 0.1, 1.0)
    with pytest.raises(rsa.RobinError, match="nearly singular"):
        rsa.GpuTree(2, 2, UNIT, 0.1, 1.0, bndry_params=np.ones((2, 2)))


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_factor_without_gpu_fails_loudly():
    t = rsa.GpuTree(2, 2, UNIT, 1.0 / 16, 1.0)
    with pytest.raises(rsa.RobinError, match="no HIP device"):
        t.factor(np.ones((t.n_leaves, t.n_int)))


def test_solve_before_factor_fails():
    t = rsa.GpuTree(2, 2, UNIT, 1.0 / 16, 1.0)
    with pytest.raises(rsa.RobinError, match="factored first"):
        t.solve(np.zeros((t.n_leaves, t.n_int)), np.zeros((t.n_leaves, t.nb)))


def test_opts_defaults_and_layout():
    """rs_opts as the header documents it (the Fortran and Python mirrors bind
    this layout): FP64, skip the root RtR, two refinement steps, autotuning and
    graphs on, an FP64 factorization in a mixed tree."""
    o = rsa.RsOpts()
    rsa.lib().rs_opts_default(ctypes.byref(o))
    assert (o.precision, o.skip_root_rtr, o.n_refine, o.verbose, o.hbm_budget, o.autotune, o.graphs,
            o.merge_fp64_levels) == (0, 1, 2, 0, 0, 1, 1, 0)
    assert ctypes.sizeof(rsa.RsOpts) == 40
