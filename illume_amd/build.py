"""Build the HIP library (gfx950) and the CPU oracle in-tree.

    python -m illume_amd.build          # both
    python -m illume_amd.build --lib    # only the HIP library

The HIP library is illume_amd/libillume_amd.so (C ABI of include/illume_c.h); the command line
illume_amd/illume (`illume <save> <scene.json> <spp> <maxdepth>`) links it; the oracle is
oracle/liboracle.so, built with g++ from oracle/oracle.cpp. Object files go to _build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "illume_amd", "csrc")
OUT_LIB = os.path.join(ROOT, "illume_amd", "libillume_amd.so")
OUT_ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
OUT_CLI = os.path.join(ROOT, "illume_amd", "illume")
BUILD_DIR = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off on device and host: no fused multiply-add unless written as fma, so a
# host build of the same formulas gives the same bits. Packed-FP32 instructions are disabled.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
]
SOURCES = ["kernels.hip", "engine.hip", "bvh.cpp", "bvh_cache.cpp", "bvh_capi.cpp", "loader.cpp"]
ORACLE_FLAGS = ["-std=c++17", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
                "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), 4)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = OUT_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs + ["-lpthread", "-lz"])
    os.replace(tmp, OUT_LIB)
    tmp = OUT_CLI + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", os.path.join(CSRC, "illume_cli.cpp"), "-o", tmp,
          "-L" + os.path.dirname(OUT_LIB), "-lillume_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, OUT_CLI)
    return OUT_LIB


def build_oracle(verbose=False):
    tmp = OUT_ORACLE + ".tmp"
    out = _run(["g++"] + ORACLE_FLAGS + [os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", tmp])
    if verbose and out.strip():
        print(out)
    os.replace(tmp, OUT_ORACLE)
    return OUT_ORACLE


def build_all(verbose=False):
    with ThreadPoolExecutor(max_workers=2) as ex:
        a = ex.submit(build_lib, verbose)
        b = ex.submit(build_oracle, verbose)
        return a.result(), b.result()


if __name__ == "__main__":
    if "--lib" in sys.argv:
        print(build_lib(True))
    elif "--oracle" in sys.argv:
        print(build_oracle(True))
    else:
        print(build_all(True))
