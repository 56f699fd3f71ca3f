"""Builds libglhip.so (HIP kernels for gfx950 + C++ host driver) in-tree.

    python verkko-hem-repo_amd/build.py [--jobs N]

Objects go to verkko-hem-repo_amd/_build/, the library to
verkko-hem-repo_amd/libglhip.so.  hipcc cross-compiles for gfx950 without a
GPU.  Packed-FP32 instructions are disabled for every HIP translation unit.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libglhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
          "-I", SRC]
# fused multiply-adds everywhere except where a kernel asks otherwise with
# "#pragma clang fp contract(off)" (plain -ffp-contract=fast ignores it)
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-ffp-contract=fast-honor-pragmas"]

SOURCES = ["gl_assemble.hip", "gl_linalg.hip", "gl_constraints.hip", "mlp.hip", "mlp.cpp", "glhip_api.cpp", "mesh.cpp", "amr.cpp",
           "matep.cpp", "setup.cpp", "kelly.cpp", "driver.cpp", "prm.cpp", "io.cpp"]


def _compile(src):
    path = os.path.join(SRC, src)
    obj = os.path.join(OUT, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + HIPFLAGS + ["-c", path, "-o", obj]
    else:
        # host-only translation units: the HIP runtime API, no kernels
        cmd = [CXX] + COMMON + ["-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(jobs=4, verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(_compile, SOURCES):
            if rc != 0:
                errors.append("%s failed:\n%s" % (src, log))
            elif verbose and log.strip():
                sys.stderr.write(log)
            objs.append(obj)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # the solver executable (the reference's sol/ driver) over the library
    exe = os.path.join(PKG, "glhip_run")
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), os.path.join(SRC, "run_main.cpp"),
           "-o", exe + ".tmp", "-L", PKG, "-lglhip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("glhip_run link failed:\n" + r.stdout + r.stderr)
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
