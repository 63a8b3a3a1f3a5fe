"""Build the Fortran drop-in (robinsolver_f.f90) with amdflang into
robinsolver_amd/fortran/_build/: module files and librobinsolver_f.so, linked
against ../librobinsolver.so.  Drivers compile with -I<_build> and link
-L<_build> -lrobinsolver_f -L<robinsolver_amd> -lrobinsolver."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def available():
    return os.path.exists(FC) or shutil.which(FC) is not None


def build():
    os.makedirs(OUT, exist_ok=True)
    src = os.path.join(HERE, "robinsolver_f.f90")
    obj = os.path.join(OUT, "robinsolver_f.o")
    subprocess.check_call([FC, "-O2", "-fPIC", "-c", src, "-o", obj, "-module-dir", OUT], cwd=OUT)
    lib = os.path.join(OUT, "librobinsolver_f.so")
    subprocess.check_call([FC, "-shared", "-fPIC", "-o", lib, obj, "-L" + PKG, "-lrobinsolver",
                           "-Wl,-rpath," + PKG])
    return lib


def build_driver(src, exe):
    """Compile and link a driver program against the drop-in."""
    subprocess.check_call([FC, "-O2", "-I" + OUT, src, "-o", exe, "-L" + OUT, "-lrobinsolver_f",
                           "-L" + PKG, "-lrobinsolver", "-Wl,-rpath," + OUT + ":" + PKG])
    return exe


if __name__ == "__main__":
    print(build())
