"""The integration patch (pde_solver_amd/integration/pde_solver_hip.patch) against the reference
tree: it applies cleanly to copies of the reference files, the patched backend registry files
compile together with the HIP plug-in built against the reference's own headers, link with
libpdehip.so and the reference CPU solver, and the HIP row, tokens and gates behave as
declared.  Runs on the CPU only, and only where the reference sources are present."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
REF = "/root/reference"
PATCH = os.path.join(ROOT, "pde_solver_amd", "integration", "pde_solver_hip.patch")
LIB = os.path.join(ROOT, "pde_solver_amd", "libpdehip.so")
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")

FILES = ["include/backend.h", "src/backend_detect.cpp", "src/solver_tokens.cpp",
         "src/backend_capability_matrix.cpp", "src/backend_solve.cpp", "src/main.cpp",
         "gui_gl/systems/backend_providers.h", "gui_gl/systems/backend_providers.cpp",
         "gui_gl/systems/backend_capabilities.cpp"]
# Compiled here.  main.cpp needs nlohmann/json and the GUI provider needs ImGui, neither of
# which is available offline: those hunks are applied (checked to apply) but not compiled.
COMPILED = ["src/backend_detect.cpp", "src/solver_tokens.cpp", "src/backend_capability_matrix.cpp",
            "src/backend_solve.cpp", "gui_gl/systems/backend_capabilities.cpp"]

pytestmark = pytest.mark.skipif(
    not (os.path.exists(os.path.join(REF, "src", "backend_solve.cpp")) and os.path.exists(LIB) and
         os.path.isdir(REF_OBJ) and shutil.which("patch") and shutil.which("g++")),
    reason="needs the reference sources, the built library and the reference CPU objects")


def test_patch_applies_compiles_links_and_runs(tmp_path):
    tree = tmp_path / "ref"
    # the whole include directory: quoted includes resolve next to the including header first
    shutil.copytree(os.path.join(REF, "include"), tree / "include")
    for f in FILES:
        d = tree / f
        d.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy(os.path.join(REF, f), d)
    r = subprocess.run(["patch", "-p1", "--no-backup-if-mismatch", "-i", PATCH], cwd=tree,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "offset" not in r.stdout and "fuzz" not in r.stdout, r.stdout
    inc = [f"-I{tree / 'include'}", f"-I{REF}/src", f"-I{ROOT}/include",
           f"-I{ROOT}/pde_solver_amd/host", f"-I{REF}/gui_gl/systems"]
    flags = ["g++", "-std=c++17", "-O0", "-DUSE_HIP", "-include", "algorithm", "-w"]
    objs = []
    srcs = [str(tree / f) for f in COMPILED] + [
        os.path.join(ROOT, "pde_solver_amd", "host", "hip_solver.cpp"),
        os.path.join(ROOT, "pde_solver_amd", "host", "latex_expr.cpp"),
        os.path.join(ROOT, "tests", "cpp", "integration_driver.cpp")]
    for i, src in enumerate(srcs):
        obj = str(tmp_path / f"o{i}.o")
        r = subprocess.run(flags + inc + ["-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, f"{src}\n{r.stderr[-4000:]}"
        objs.append(obj)
    ref_objs = []
    for dp, _, fs in os.walk(REF_OBJ):
        ref_objs += [os.path.join(dp, f) for f in fs if f.endswith(".o")]
    exe = str(tmp_path / "driver")
    r = subprocess.run(["g++", "-o", exe, *objs, *ref_objs, "-fopenmp", f"-L{os.path.dirname(LIB)}",
                        "-lpdehip", f"-Wl,-rpath,{os.path.dirname(LIB)}", "-lpthread"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
