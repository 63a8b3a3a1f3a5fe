"""The Fortran drop-in: the reference's module/type/procedure names compile
with amdflang, and a driver written against the reference API links against
the MI355X library.  The GPU test runs it and checks the manufactured
solution (O(h^2) discretisation error, and agreement with the oracle)."""
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


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_runs_on_gpu(driver_exe, precision, gpu):
    print("case: fortran driver precision", precision, flush=True)
    errs = []
    for n in (32, 64):
        r = subprocess.run([driver_exe, str(n), "2", str(precision)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        line = [l for l in r.stdout.splitlines() if l.startswith("RELL2")][0]
        errs.append(float(line.split()[1]))
        assert "Factored: T" in r.stdout
    assert errs[1] < 1e-4 and 3.2 < errs[0] / errs[1] < 4.8, errs
