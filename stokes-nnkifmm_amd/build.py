"""Build libnnkifmm.so (HIP + C++ host code) for gfx950 with hipcc.

Run as ``python stokes-nnkifmm_amd/build.py`` or through
``__graft_entry__.build()``.  The library lands in ``lib/`` next to this
file and is loaded from there by the Python host layer.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnnkifmm.so")
SELFTEST = os.path.join(LIBDIR, "libnnk_selftest.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["nnk_tree.cpp", "nnk_tree_gpu.hip", "nnk_svd.hip", "nnk_fmm.hip", "nnk_direct.hip", "nnk_reduced.hip",
           "nnk_nearest.hip", "nnk_mobility.hip", "nnk_mixed.hip", "nnk_batch.hip", "nnk_capi.cpp"]

# -packed-fp32-ops: no packed FP32 instructions in device code.
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
          "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    # keep real warnings visible, drop the host pass's unknown-feature notice
    for line in r.stdout.splitlines():
        if "packed-fp32-ops" in line or "not a recognized feature" in line:
            continue
        if line.strip():
            sys.stderr.write(line + "\n")


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
        objs.append(obj)
    # the translation units compile independently: a few at a time
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/llvm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    _run(cmd)
    os.replace(tmp, LIB)
    # test-only companion library (GEMM / internal-array probes, host-only
    # planner checks; its source lives with the tests)
    st = os.path.join(os.path.dirname(HERE), "tests", "native", "nnk_selftest.hip")
    sto = os.path.join(objdir, "nnk_selftest.o")
    _run([HIPCC, "-x", "hip"] + COMMON + DEVICE_FLAGS + ["-I" + CSRC, "-c", st, "-o", sto])
    tmp2 = SELFTEST + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp2, sto, "-L" + LIBDIR, "-lnnkifmm",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp2, SELFTEST)
    build_gateway()
    return LIB


def build_gateway():
    """The MATLAB marshalling layer (matlab/nnk_gateway.c, plain C, gcc) as
    libnnk_gateway.so, and the C program that drives it the way the MEX
    gateway does (tests/native/gateway_test.c)."""
    gcc = os.environ.get("CC", "gcc")
    inc = ["-I" + os.path.join(os.path.dirname(HERE), "include"), "-I" + os.path.join(HERE, "matlab")]
    gw = os.path.join(LIBDIR, "libnnk_gateway.so")
    _run([gcc, "-std=c99", "-O2", "-Wall", "-Wextra", "-fPIC", "-shared"] + inc +
         [os.path.join(HERE, "matlab", "nnk_gateway.c"), "-o", gw + ".tmp", "-L" + LIBDIR, "-lnnkifmm",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(gw + ".tmp", gw)
    exe = os.path.join(LIBDIR, "nnk_gateway_test")
    _run([gcc, "-std=c99", "-O2", "-Wall", "-Wextra"] + inc +
         [os.path.join(os.path.dirname(HERE), "tests", "native", "gateway_test.c"), "-o", exe + ".tmp",
          "-L" + LIBDIR, "-lnnk_gateway", "-lnnkifmm", "-Wl,-rpath,$ORIGIN"])
    os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
