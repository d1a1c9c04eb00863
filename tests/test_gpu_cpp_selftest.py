"""The C++ plug-in (hip_solver.h + backend registry with the HIP row) end to end: the
reference's self-test cases through SolveWithBackend(..., BackendKind::HIP)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "..", "pde_solver_amd", "bin", "hip_self_test")


def test_cpp_self_test_through_solve_with_backend(tmp_path):
    assert os.path.exists(BIN), "build the host layer first (python -m pde_solver_amd.build)"
    # the 129^2 workload's .vtk output: PDEHIP_VTK_OUT when set (to keep it), else a temporary file
    vtk = os.environ.get("PDEHIP_VTK_OUT") or str(tmp_path / "cfg1_poisson_129_cg_hip.vtk")
    r = subprocess.run([BIN, vtk], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[FAIL]" not in r.stdout
