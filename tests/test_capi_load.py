"""CPU-only checks of the C ABI boundary: the library loads, exports every entry point that
include/pdehip.h declares, and its struct layouts agree with the ctypes mirror."""
import ctypes as C
import os
import re
import subprocess

import pytest

from pde_solver_amd import capi

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "pdehip.h")
LIB = os.path.join(ROOT, "pde_solver_amd", "libpdehip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pdehip_[a-z_]+)\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpdehip.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing
    assert lib.pdehip_abi_version() == 2


def test_struct_layouts_match_header():
    """Compile a probe against pdehip.h and compare sizeof/offsetof with the ctypes mirror."""
    probe = r"""
#include <stddef.h>
#include <stdio.h>
#include "pdehip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(pdehip_problem), sizeof(pdehip_options),
         sizeof(pdehip_report), sizeof(pdehip_time_config), sizeof(pdehip_batch_config),
         sizeof(pdehip_server_config), sizeof(pdehip_server_stats), sizeof(pdehip_face_bc));
  printf("%zu %zu %zu %zu\n", offsetof(pdehip_problem, bc), offsetof(pdehip_problem, nonlinear),
         offsetof(pdehip_problem, initial_velocity), offsetof(pdehip_options, cancel));
  return 0;
}
"""
    tmp = os.path.join(ROOT, "_build")
    os.makedirs(tmp, exist_ok=True)
    src = os.path.join(tmp, "layout_probe.c")
    exe = os.path.join(tmp, "layout_probe")
    with open(src, "w") as f:
        f.write(probe)
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out[:8]]
    offs = [int(v) for v in out[8:]]
    assert sizes == [C.sizeof(capi.Problem), C.sizeof(capi.Options), C.sizeof(capi.Report),
                     C.sizeof(capi.TimeConfig), C.sizeof(capi.BatchConfig),
                     C.sizeof(capi.ServerConfig), C.sizeof(capi.ServerStats), C.sizeof(capi.FaceBC)]
    assert offs == [capi.Problem.bc.offset, capi.Problem.nonlinear.offset,
                    capi.Problem.initial_velocity.offset, capi.Options.cancel.offset]
