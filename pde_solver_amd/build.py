"""Build libpdehip.so (HIP kernels + C ABI) for gfx950 in-tree with hipcc.

    python -m pde_solver_amd.build            # or build() from Python
Objects go to _build/ at the repository root; the shared library to pde_solver_amd/libpdehip.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpdehip.so")
BUILD = os.path.join(ROOT, "_build")

SOURCES = ["k_stencil.hip", "k_smooth.hip", "k_box.hip", "k_plane.hip", "k_aux.hip", "k_time.hip", "k_krylov.hip", "k_batch.hip",
           "noalloc.hip",
           "runtime.cpp", "serving.cpp", "batch.cpp", "capi.cpp"]

ARCH = os.environ.get("PDEHIP_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in device code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall", "-Wno-unused-function",
         "-Wno-unknown-pragmas"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr[-6000:]}")
    return obj


HOST_SOURCES = ["host/hip_solver.cpp", "host/backend_hip.cpp", "host/latex_expr.cpp"]
HOST_LIB = os.path.join(HERE, "libpdehip_host.so")
SELF_TEST = os.path.join(HERE, "bin", "hip_self_test")


def build_host(verbose: bool = False) -> None:
    """The C++ plug-in layer (hip_solver.h + backend registry) and its self-test driver, with
    g++ against the compatible type headers in host/compat."""
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host", "compat"),
           "-I" + os.path.join(HERE, "host")]
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
    srcs = [os.path.join(HERE, s) for s in HOST_SOURCES]
    link = ["-L" + HERE, "-lpdehip", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(["g++", *flags, *inc, "-shared", *srcs, "-o", HOST_LIB + ".tmp", *link],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host layer build failed\n{r.stderr[-6000:]}")
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.makedirs(os.path.dirname(SELF_TEST), exist_ok=True)
    r = subprocess.run(["g++", *flags, *inc, os.path.join(ROOT, "tests", "cpp", "hip_self_test.cpp"),
                        "-o", SELF_TEST + ".tmp", "-L" + HERE, "-lpdehip_host", "-lpdehip",
                        "-Wl,-rpath,$ORIGIN/..", "-lpthread"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"self-test build failed\n{r.stderr[-6000:]}")
    os.replace(SELF_TEST + ".tmp", SELF_TEST)
    if verbose:
        print(f"built {HOST_LIB} and {SELF_TEST}")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_host(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
