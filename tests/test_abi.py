"""The C ABI without a GPU: the library loads and exports every entry point
include/premixfem_gpu.h declares, argument checking returns the PETSc error
codes of src/PreMixFEM_3D.c:1650-1741, and the PCSHELL adapter
(premixfem_amd/petsc) drives the reference call sequence of
src/Expe_scal.c:66-160 through a mock PETSc."""
from __future__ import annotations

import ctypes as C
import pathlib
import re
import subprocess

import numpy as np
import pytest

from premixfem_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "premixfem_gpu.h"
MOCK = ROOT / "tests" / "mock_petsc"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pmx_\w+)\s*\(", text, re.M)))


def test_header_declarations_match_binding():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def _init(dom, mesh, options):
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.pmx_init(C.byref(h), np.asarray(dom, np.float64), np.asarray(mesh, np.int32),
                    options.encode())
    return rc, L.pmx_last_error().decode()


@pytest.mark.parametrize("dom,mesh,options,code", [
    ((0.0, 1.0, 1.0), (8, 8, 8), "", 62),            # :1653 dom check
    ((1.0, 1.0, 1.0), (8, 0, 8), "", 62),            # :1654 mesh check
    ((1.0, 1.0, 1.0), (8, 8, 8), "-sd 0", 62),       # :1672 sub_domains >= 1
    ((1.0, 1.0, 1.0), (8, 8, 8), "-en_lv1 0", 62),   # :1676
    ((1.0, 1.0, 1.0), (8, 8, 8), "-en_lv2 0", 62),   # :1681
    ((1.0, 1.0, 1.0), (8, 8, 8), "-eg_op 3", 62),    # :1685
    ((1.0, 1.0, 1.0), (8, 8, 8), "-sd 9", 62),       # more subdomains than cells
    ((1.0, 1.0, 1.0), (8, 8, 8), "-pc_precision 16", 62),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-use_full_Cholesky_lv1 -pc_precision 32", 56),  # FP64 only
    ((1.0, 1.0, 1.0), (8, 8, 8), "-pin 8,0,0", 63),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-eig_ortho 2", 63),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-no_such_option 1", 62),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-procs_virtual 2,2", 62),         # px,py,pz or one count
    ((1.0, 1.0, 1.0), (8, 8, 8), "-procs_virtual 2,x,1", 62),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-procs_virtual 0,1,1", 62),
    ((1.0, 1.0, 1.0), (8, 8, 8), "-procs_virtual 3,1,1 -sd 3", 62),  # boxes of 2 cells
    # the band Cholesky coarse factor has no process-box form
    ((1.0, 1.0, 1.0), (8, 8, 8), "-procs_virtual 2 -coarse_band 1 -coarse_planes 0", 56),
])
def test_init_argument_errors(dom, mesh, options, code):
    rc, msg = _init(dom, mesh, options)
    assert rc == code, (rc, msg)
    assert msg


def test_init_null_context():
    L = _lib.lib()
    rc = L.pmx_init(None, np.ones(3), np.full(3, 8, np.int32), b"")
    assert rc == 85


def _make(target):
    r = subprocess.run(["make", "-s", "-C", str(MOCK), target], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_pcshell_adapter_over_stub_core():
    _make("stub")
    exe = MOCK / "_build" / "driver_stub"
    r = subprocess.run([str(exe), "-mesh_x", "12", "-mesh_y", "10", "-mesh_z", "8", "-sd", "3",
                        "-en_lv1", "5", "-si_lv2", "2", "-eb_lv1", "0.5", "-use_2level",
                        "-pin", "11,0,0", "-pc_precision", "32", "-ksp_rtol", "1e-8"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    opts = next(l for l in r.stdout.splitlines() if l.startswith("STUB options:"))
    for frag in ("-sd 3", "-en_lv1 5", "-si_lv2 2", "-eb_lv1 0.5", "-use_2level",
                 "-pin 11,0,0", "-pc_precision 32", "-en_lv2 4", "-si_lv1 1"):
        assert frag in opts, (frag, opts)
    res = next(l for l in r.stdout.splitlines() if l.startswith("RESULT"))
    fields = dict(kv.split("=") for kv in res.split()[1:])
    assert fields["reason"] == "2"
    assert int(fields["its"]) > 0
    applies = next(l for l in r.stdout.splitlines() if l.startswith("STUB applies"))
    assert int(applies.split()[-1]) == int(fields["its"])


def test_pcshell_adapter_propagates_error_codes():
    _make("stub")
    exe = MOCK / "_build" / "driver_stub"
    r = subprocess.run([str(exe), "-en_lv1", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 62
    assert "max_eigen_num_lv1=0" in r.stderr


def test_pcshell_setup_reads_kappa_again():
    """The reference's _PC_setup reads s_ctx->kappa at every setup
    (src/PreMixFEM_3D.c:1398-1402): after kappa is doubled (and the operator
    with it), the second KSPSolve uses the new kappa, so the solution halves."""
    _make("stub")
    exe = MOCK / "_build" / "driver_stub"
    r = subprocess.run([str(exe), "-mesh_x", "12", "-mesh_y", "10", "-mesh_z", "8",
                        "-ksp_rtol", "1e-10", "-kappa_rescale"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    res = next(l for l in r.stdout.splitlines() if l.startswith("RESULT2"))
    fields = dict(kv.split("=") for kv in res.split()[1:])
    assert fields["reason"] == "2"
    assert float(fields["ratio"]) == pytest.approx(2.0, rel=1e-12)


def test_pcshell_setup_refuses_an_operator_not_built_from_kappa():
    """PC_apply_vec's residuals use the PC's operator (:1852); the GPU
    preconditioner is built from kappa, so an operator that differs from
    PC_create_A's is refused with PETSC_ERR_ARG_WRONG."""
    _make("stub")
    exe = MOCK / "_build" / "driver_stub"
    r = subprocess.run([str(exe), "-mesh_x", "12", "-mesh_y", "10", "-mesh_z", "8",
                        "-operator_mismatch"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 62, (r.returncode, r.stderr)
    assert "differs from the one built from s_ctx->kappa" in r.stderr
