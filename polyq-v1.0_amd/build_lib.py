"""Build libpolyq_amd.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

Usage: python build_lib.py [--jobs N] [--sanitize address,undefined|thread --out DIR]
The library lands in polyq_amd/libpolyq_amd.so next to the Python bindings.
With --sanitize, the host C++ sources are built with the GCC sanitizers
(-fsanitize=..., -O1 -g) and the library and a copy of the Python package land
in DIR/polyq_amd instead; the HIP objects are the ordinary ones (device code is
never sanitized).  Run Python with the matching GCC runtime preloaded
(LD_PRELOAD=libasan.so / libtsan.so; tests/sanitize.sh does this).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "polyq_amd", "libpolyq_amd.so")

ROCM_LLVM = "/opt/rocm/llvm/bin"
HOST_SOURCES = ["pq_compile.cpp", "pq_engine.cpp", "pq_z8.cpp", "pq_seg.cpp", "pq_api.cpp", "pq_engine_abi.cpp", "pq_server.cpp"]
HIP_SOURCES = ["pq_kernels.hip", "pq_valu.hip"]

# x86-64-v3: the library built here also runs on the GPU host's CPU.
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-Wall", "-Wno-unused-function", "-pthread"]
# Packed FP32 instructions stay disabled; no FMA contraction so the two-sum
# accumulators and the reference-style FP64 products round as written.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(jobs=8, verbose=False, sanitize=None, out_dir=None):
    os.makedirs(OBJ, exist_ok=True)
    cmds = []
    objs = []
    host_flags, host_obj, out, pkg = HOST_FLAGS, OBJ, OUT, os.path.join(HERE, "polyq_amd")
    if sanitize:
        if not out_dir:
            raise ValueError("a sanitized build goes to its own --out directory")
        host_obj = os.path.join(out_dir, "obj")
        os.makedirs(host_obj, exist_ok=True)
        pkg = os.path.join(out_dir, "polyq_amd")
        os.makedirs(pkg, exist_ok=True)
        for f in os.listdir(os.path.join(HERE, "polyq_amd")):
            if f.endswith(".py"):
                shutil.copy(os.path.join(HERE, "polyq_amd", f), pkg)
        out = os.path.join(pkg, "libpolyq_amd.so")
        host_flags = [f for f in HOST_FLAGS if f != "-O3"] + ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=" + sanitize]
    # ThreadSanitizer builds use the ROCm clang and its runtime (GCC 11's
    # libtsan does not intercept pthread_cond_clockwait, which
    # condition_variable::wait_for calls, and reports false double locks)
    host_cxx = os.path.join(ROCM_LLVM, "clang++") if sanitize == "thread" else "g++"
    for s in HOST_SOURCES:
        o = os.path.join(host_obj, s + ".o")
        objs.append(o)
        cmds.append([host_cxx] + host_flags + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        objs.append(o)
        cmds.append(["hipcc"] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for text in ex.map(_run, cmds):
            if verbose and text.strip():
                print(text)
    tmp = out + ".tmp"
    runtime = []
    if sanitize and sanitize != "thread":
        # the GCC sanitizer runtimes the instrumented host objects call into
        # (a TSan library takes its runtime from the executable that loads it)
        runtime = ["-l" + {"address": "asan", "undefined": "ubsan"}[k] for k in sanitize.split(",")]
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + runtime)
    os.replace(tmp, out)
    # the CPython fast path of Simulator.amplitudes, linked against the library
    import sysconfig
    fast = os.path.join(pkg, "_pqfast" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run(["gcc", "-O2", "-shared", "-fPIC", "-I", sysconfig.get_paths()["include"], "-I", os.path.join(HERE, "..", "include"),
          os.path.join(CSRC, "pq_pyfast.c"), "-o", fast + ".tmp", "-L", os.path.dirname(out), "-lpolyq_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(fast + ".tmp", fast)
    return out


if __name__ == "__main__":
    j = 8
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    san = sys.argv[sys.argv.index("--sanitize") + 1] if "--sanitize" in sys.argv else None
    od = sys.argv[sys.argv.index("--out") + 1] if "--out" in sys.argv else None
    print(build(jobs=j, verbose=True, sanitize=san, out_dir=od))
