"""The PCSHELL adapter on the GPU: a host PETSc-style CG (mock PETSc) calls
PC_setup / PC_apply_vec of premixfem_amd/petsc/pmx_pcshell.c exactly as
src/Expe_scal.c:100-119 registers them, and the result must match the
GPU-resident PCG (pmx_solve) on the same system: iterations within 1 and
solutions within 1e-8 relative."""
from __future__ import annotations

import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "mock_petsc" / "_build" / "driver_gpu"


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["-mesh_x", "16", "-mesh_y", "16", "-mesh_z", "16", "-sd", "2", "-en_lv1", "4",
     "-pin", "15,0,0"],
    ["-mesh_x", "20", "-mesh_y", "12", "-mesh_z", "10", "-sdx", "3", "-sdy", "2", "-sdz", "2",
     "-en_lv1", "3", "-pin", "19,0,0", "-eg_op", "2"],
])
def test_host_cg_through_pcshell_matches_gpu_pcg(args):
    assert EXE.exists(), f"{EXE} is not built: run __graft_entry__.build()"
    print("pcshell case", " ".join(args), flush=True)
    r = subprocess.run([str(EXE), *args, "-ksp_rtol", "1e-8", "-compare_gpu_pcg"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    res = next(l for l in r.stdout.splitlines() if l.startswith("RESULT"))
    f = dict(kv.split("=") for kv in res.split()[1:])
    assert f["reason"] == "2" and f["gpu_reason"] == "2"
    assert abs(int(f["its"]) - int(f["gpu_its"])) <= 1
    assert float(f["rel_diff"]) <= 1e-8
