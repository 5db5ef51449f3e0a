"""Build recipe for the native library (hipcc, gfx950 only).

Produces tagmatch_amd/lib/libtagmatch_amd.so from the HIP and C++ sources in
tagmatch_amd/csrc, plus the C++ facade (class tagmatch) and its test driver.
Every call recompiles; objects are built in parallel.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU ops.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", INCLUDE, "-I", os.path.join(INCLUDE, "tagmatch"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE, "-I", os.path.join(INCLUDE, "tagmatch")]

HIP_SOURCES = ["hip_util.hip", "fib.hip", "match.hip", "partition.hip", "cost_model.hip"]
CXX_SOURCES = ["workload.cc", "engine.cc", "server.cc", "io.cc", "expand.cc"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _sources():
    hip = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cxx = [s for s in CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return hip, cxx


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hip, cxx = _sources()
    jobs = []
    for s in hip:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in cxx:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "c++", "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [os.path.join(OBJ, s + ".o") for s in hip + cxx]
    so = os.path.join(LIB, "libtagmatch_amd.so")
    tmp = so + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, so)
    # C++ driver that runs the reference-style API tests (tests/cpp)
    drv_src = os.path.join(ROOT, "tests", "cpp", "api_tests.cc")
    if os.path.exists(drv_src):
        _run([HIPCC, *CXX_FLAGS, drv_src, "-o", os.path.join(LIB, "tm_api_tests"),
              f"-L{LIB}", "-ltagmatch_amd", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    shutil.rmtree(OBJ, ignore_errors=True)
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
