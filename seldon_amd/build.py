"""Build the native parts of seldon_amd in-tree.

* ``seldon_amd/libshk.so``           -- the HIP library (gfx950), from csrc/*.hip
* ``seldon_amd/libshk_workloads.so`` -- seeded input generators (host C)
* C++ programs of the Seldon-compatible API under ``_build/bin`` (examples
  and the C++ parity drivers used by the GPU tests).

hipcc is driven directly (no cmake); every HIP object is compiled for gfx950
only, with FMA contraction off (SpMV row sums must round like Seldon's C++)
and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-I", INCLUDE, "-I", CSRC,
]

HIP_SOURCES = ["shk_context.hip", "shk_spmv.hip", "shk_blas1.hip", "shk_solvers.hip",
               "shk_extras.hip", "shk_batched.hip", "shk_assembly.hip", "shk_sor.hip", "shk_hb.hip",
               "shk_vector_io.hip", "shk_sliced.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_workloads():
    out = os.path.join(PKG, "libshk_workloads.so")
    _run(["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "shk_workloads.c"), "-lm"])
    return out


def build_hip(jobs: int = 8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = os.path.join(PKG, "libshk.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


CPP_PROGRAMS = {
    # name: (source under tests/cpp or include examples, extra flags)
    "seldon_hip_driver": "tests/cpp/seldon_hip_driver.cpp",
}


def build_cpp_programs():
    """C++ programs written against include/seldon_amd (the Seldon-compatible
    API), linked to libshk.so with an rpath so they run in place."""
    bindir = os.path.join(BUILD, "bin")
    os.makedirs(bindir, exist_ok=True)
    outs = []
    for name, src in CPP_PROGRAMS.items():
        path = os.path.join(ROOT, src)
        if not os.path.exists(path):
            continue
        exe = os.path.join(bindir, name)
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", INCLUDE,
              "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-o", exe, path,
              "-L", PKG, "-lshk", "-Wl,-rpath,$ORIGIN/../.."])
        outs.append(exe)
    return outs


def build_oracle():
    """The test-only oracle (oracle/liboracle.so and, where the reference tree
    is present, the reference drivers in oracle/_ref)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])


def build_all():
    build_workloads()
    build_hip()
    build_cpp_programs()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(PKG, "libshk.so"))
