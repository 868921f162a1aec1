"""Builds the MI355X backend library, libngx.so, in place (next to this file) with hipcc for gfx950.

Every HIP source is compiled with -ffp-contract=off (each float operation is one IEEE operation in source
order, as the reference's CPU backend evaluates it) and without packed FP32 instructions."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libngx.so")
SOURCES = ["ngx_tlas.hip", "ngx_trace.hip", "ngx_refit.hip", "ngx_selftest.hip", "ngx_api.cpp", "ngx_blas.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-I" + os.path.join(ROOT, "include")]


def _compile(src, extra=(), objdir="_obj"):
    obj = os.path.join(CSRC, objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stderr}")
    return obj


def build(verbose=False, extra_flags=(), out=LIB):
    """Compiles every source and links `out`. extra_flags (e.g. -DNGX_TRACE_STATS for the instrumented
    traversal) go to every compile and use their own object directory."""
    objdir = "_obj" if not extra_flags else "_obj_" + "_".join(f.strip("-").replace("=", "_") for f in extra_flags)
    os.makedirs(os.path.join(CSRC, objdir), exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as pool:
        objs = list(pool.map(lambda src: _compile(src, extra_flags, objdir), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libngx.so failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


CPP_TEST = os.path.join(ROOT, "tests", "cpp", "test_batch.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_batch")


def build_cpp_tests(verbose=False):
    """Compiles the C++ front end's test program (tests/cpp/test_batch.cpp) against libngx.so with g++; the
    binary finds the library through its run path, so it runs from any copy of the tree."""
    os.makedirs(os.path.dirname(CPP_TEST_BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "cpp"), CPP_TEST,
           "-L" + HERE, "-lngx", "-Wl,-rpath,$ORIGIN/../../../neraulgraphics_amd", "-o", CPP_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {CPP_TEST} failed:\n{r.stderr}")
    if verbose:
        print("built", CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_tests(verbose=True)
    sys.exit(0)
