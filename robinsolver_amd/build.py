"""Build the gfx950 library librobinsolver.so in-tree with hipcc.

Usage: python -m robinsolver_amd.build   (or robinsolver_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librobinsolver.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["k_leaf.hip", "k_gemm.hip", "k_inv.hip", "k_misc.hip", "tree.hip", "plan.cpp", "capi.cpp",
           "server.cpp", "persist.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-result"]
# Packed FP32 instructions stay off (hipcc would otherwise form v_pk_*_f32 from float2 code).
# Matrix-core accumulators live in ordinary VGPRs (amdgpu-mfma-vgpr-form): with the
# accumulation registers as the default, every MFMA result was copied out with
# v_accvgpr_read (12 per layer step of the leaf kernels), and the GEMM kernels held
# both copies (fewer registers with the VGPR form: <double,double,64,64> 276 -> 180,
# two waves per SIMD instead of one).  The MFMA operations themselves are unchanged.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _compile_cmd(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    return cmd, obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        cmd, obj = _compile_cmd(s)
        objs.append(obj)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = "\n".join(l for l in out.decode(errors="replace").splitlines()
                         if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((s, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for s, text in failed:
            sys.stderr.write("---- %s ----\n%s\n" % (s, text))
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = OUT + ".tmp"
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lamdhip64"]
    subprocess.check_call(link)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
