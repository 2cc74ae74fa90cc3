"""Build recipe of the MI355X engine (no cmake, no JIT cache).

Produces, in-tree:
  gsvit_copy_amd/lib/libgsvit_amd.so   C-ABI library (host C + HIP kernels for gfx950)
  gsvit_copy_amd/bin/gsvit_amd         command-line driver (parameter file in, logs out)
  gsvit_copy_amd/lib/libgsv_probe.so   bench.py's streaming-copy probe (tooling, not the engine)
and, for the test suite, the CPU oracle under oracle/_build/ (make).
All code uses the ROCm LLVM OpenMP runtime so that one process can hold the
library and the oracle together.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang")
OMPLIB = os.path.join(ROCM, "llvm", "lib")
ARCH = os.environ.get("GSV_OFFLOAD_ARCH", "gfx950")

C_SOURCES = ["settings.c", "plan.c", "database.c", "containers.c", "output.c", "session.c"]
HIP_SOURCES = ["kernels.hip", "gpu_engine.hip", "adjoint.hip", "batch.hip", "nfff.hip"]
INC = ["-I" + os.path.join(REPO, "include"), "-I" + SRC]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-fopenmp",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
C_FLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (res.returncode, " ".join(cmd), out))
    return out


def build_library(verbose=False):
    objdir = os.path.join(PKG, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in C_SOURCES:
        o = os.path.join(objdir, s + ".o")
        _run([CLANG] + C_FLAGS + INC + ["-c", os.path.join(SRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SOURCES:
        o = os.path.join(objdir, s + ".o")
        _run([HIPCC] + HIP_FLAGS + INC + ["-c", os.path.join(SRC, s), "-o", o])
        objs.append(o)
    lib = os.path.join(PKG, "lib", "libgsvit_amd.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fopenmp", "-o", lib] + objs
         + ["-L" + OMPLIB, "-Wl,-rpath," + OMPLIB, "-lm"])
    bindir = os.path.join(PKG, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "gsvit_amd")
    _run([CLANG] + C_FLAGS + INC + [os.path.join(SRC, "svit.c"), "-o", exe, "-L" + os.path.join(PKG, "lib"),
                                    "-lgsvit_amd", "-Wl,-rpath,$ORIGIN/../lib", "-Wl,-rpath," + OMPLIB,
                                    "-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib"),
                                    "-lamdhip64", "-lm"])
    if verbose:
        print("built", lib, "and", exe)
    return lib


def build_probe(verbose=False):
    """bench.py's measurement tooling (not part of the engine): the streaming
    copy probe of bench_probe.hip as gsvit_copy_amd/lib/libgsv_probe.so"""
    lib = os.path.join(PKG, "lib", "libgsv_probe.so")
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    _run([HIPCC] + HIP_FLAGS + ["-shared", os.path.join(REPO, "bench_probe.hip"), "-o", lib])
    if verbose:
        print("built", lib)
    return lib


def build_oracle(verbose=False):
    _run(["make", "-s", "-B", "-C", os.path.join(REPO, "oracle"), "CC=" + CLANG,
          "LDFLAGS_OMP=-L%s -Wl,-rpath,%s" % (OMPLIB, OMPLIB), "all"])
    if verbose:
        print("built oracle")


def build(verbose=False):
    build_library(verbose)
    build_probe(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
