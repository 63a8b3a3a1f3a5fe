"""The Fortran drop-in: the reference's module/type/procedure names compile
with amdflang; the reference's own driver (example.f90, read from the
reference checkout when it is present, never copied) and a 2D driver written
against the same API compile and link unchanged against the MI355X library.
The GPU test runs the 2D driver and checks its leaf solutions against the
oracle on the same problem (rel. l2 <= 1e-10 FP64, <= 1e-8 mixed) and the
O(h^2) convergence of the manufactured solution."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

fb = pytest.importorskip("robinsolver_amd.fortran.build_fortran")
DRIVER = os.path.join(ROOT, "tests", "fortran", "driver_2d.f90")


@pytest.fixture(scope="module")
def driver_exe(tmp_path_factory):
    if not fb.available():
        pytest.skip("amdflang not present")
    lib = os.path.join(fb.OUT, "librobinsolver_f.so")
    if not os.path.exists(lib):
        fb.build()
    exe = str(tmp_path_factory.mktemp("fdrv") / "driver_2d")
    return fb.build_driver(DRIVER, exe)


def test_fortran_driver_compiles_and_links(driver_exe):
    assert os.path.exists(driver_exe)
    out = subprocess.run(["nm", "-D", "--undefined-only", os.path.join(fb.OUT, "librobinsolver_f.so")],
                         capture_output=True, text=True).stdout
    for sym in ("rs_plan", "rs_factor", "rs_solve", "rs_get_solution", "rs_leaf_points"):
        assert sym in out


REF_EXAMPLE = "/root/reference/example.f90"


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE), reason="reference checkout not present")
def test_reference_example_driver_compiles_and_links(tmp_path):
    """example.f90 (the reference's only driver, example.f90:1-89) builds
    unchanged against the drop-in modules and libraries."""
    if not fb.available():
        pytest.skip("amdflang not present")
    if not os.path.exists(os.path.join(fb.OUT, "librobinsolver_f.so")):
        fb.build()
    exe = fb.build_driver(REF_EXAMPLE, str(tmp_path / "example"))
    assert os.path.exists(exe)


def _driver_problem_g(o):
    """The 2D driver's boundary data at the leaf boundary points, sampled as the
    reference samples g_face (tangential coordinate first, fd_module.f90:645-651);
    zero on faces inside the domain (not read)."""
    c = 1.0 / np.sqrt(2.0)
    g = np.zeros((o.n_leaves, o.nb), np.complex128)
    for q in range(o.n_leaves):
        xyz, face = o.leaf_points(q)
        for p in range(o.nb):
            f = int(face[p])
            ax = (f - 1) // 2
            x = xyz[p]
            if not ((f % 2 == 1 and x[ax] == 0.0) or (f % 2 == 0 and x[ax] == 1.0)):
                continue
            t = x[1 - ax]
            if f % 2 == 1:
                g[q, p] = 1j * np.cos(t * c)
            else:
                g[q, p] = 1j * np.cos(c) * np.cos(t * c) - c * np.sin(c) * np.cos(t * c)
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_runs_on_gpu(driver_exe, precision, gpu, tmp_path):
    from oracle.pyoracle import OracleTree
    print("case: fortran driver precision", precision, flush=True)
    errs = []
    for n in (32, 64):
        out = str(tmp_path / ("u_%d_%d.bin" % (n, precision)))
        r = subprocess.run([driver_exe, str(n), "2", str(precision), out], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        line = [l for l in r.stdout.splitlines() if l.startswith("RELL2")][0]
        errs.append(float(line.split()[1]))
        assert "Factored: T" in r.stdout
        # the same problem on the oracle: k = 1, f = 0, the driver's impedance data
        o = OracleTree(2, 2, [[0, 1], [0, 1], [0, 1]], 1.0 / n, 1.0)
        o.factor(np.ones((o.n_leaves, o.n_int), np.complex128))
        u_ref = o.solve(np.zeros((o.n_leaves, o.n_int), np.complex128), _driver_problem_g(o))
        u = np.fromfile(out, dtype=np.complex128).reshape(o.n_leaves, o.m)
        rel = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
        print("  n=%d rel l2 vs oracle %.2e" % (n, rel), flush=True)
        assert rel <= (1e-10 if precision == 0 else 1e-8), rel
    assert errs[1] < 1e-4 and 3.2 < errs[0] / errs[1] < 4.8, errs
