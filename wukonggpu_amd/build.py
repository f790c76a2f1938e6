"""Builds the native libraries in-tree.

  wukonggpu_amd/lib/libwkg_gstore.so   host gstore: generator, builder, lookup (g++)
  wukonggpu_amd/lib/libwkg_hip.so      HIP kernels + device C ABI (hipcc, gfx950)
  wukonggpu_amd/lib/libwkg_engine.so   C++ engine above the C ABI (g++)

Run `python -m wukonggpu_amd.build` or call build_all().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

# Packed-FP32 instructions stay off in every HIP build (nothing here is
# floating point, but the flag guards any float code the compiler emits).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gstore():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libwkg_gstore.so")
    _run([CXX, *CXX_FLAGS, "-fopenmp", "-march=x86-64-v2", f"-I{INC}",
          os.path.join(CSRC, "gstore.cpp"), "-o", out])
    return out


HIP_SOURCES = ("wkg_hip.hip", "graph.hip", "batch.hip", "exec.hip", "builder.hip", "split.hip")


def build_hip():
    """Compiles each HIP source to an object in parallel, then links."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB, exist_ok=True)
    obj_dir = os.path.join(LIB, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    out = os.path.join(LIB, "libwkg_hip.so")
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    jobs = []
    for f in HIP_SOURCES:
        obj = os.path.join(obj_dir, f.replace(".hip", ".o"))
        jobs.append(([HIPCC, *cflags, f"-I{INC}", "-c", os.path.join(CSRC, f), "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=min(len(jobs), 4)) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[o for _, o in jobs], "-o", out])
    return out


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libwkg_engine.so")
    _run([CXX, *CXX_FLAGS, "-pthread", f"-I{INC}",
          os.path.join(CSRC, "engine.cpp"), "-o", out,
          f"-L{LIB}", "-lwkg_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle():
    root = os.path.dirname(HERE)
    _run(["make", "-s", "-C", os.path.join(root, "oracle")])


def build_all():
    build_gstore()
    build_hip()
    build_engine()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gstore": build_gstore, "hip": build_hip,
         "engine": build_engine, "oracle": build_oracle}[w]()
