"""ctypes binding of ``libpremixfem_gpu.so`` (declared in include/premixfem_gpu.h).

The library is built in-tree by ``premixfem_amd.build``; importing this module
never falls back to anything else: a missing library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libpremixfem_gpu.so"

PMX_SUCCESS = 0
ERRORS = {
    55: "PETSC_ERR_MEM", 56: "PETSC_ERR_SUP", 62: "PETSC_ERR_ARG_WRONG",
    63: "PETSC_ERR_ARG_OUTOFRANGE", 73: "PETSC_ERR_ORDER", 76: "PETSC_ERR_LIB",
    83: "PETSC_ERR_USER", 85: "PETSC_ERR_ARG_NULL", 91: "PETSC_ERR_NOT_CONVERGED",
}
REASONS = {2: "CONVERGED_RTOL", 3: "CONVERGED_ATOL", -3: "DIVERGED_ITS",
           -4: "DIVERGED_DTOL", -8: "DIVERGED_INDEFINITE_PC",
           -9: "DIVERGED_NANORINF", -10: "DIVERGED_INDEFINITE_MAT", 0: "ITERATING"}

# Every entry point include/premixfem_gpu.h declares.
EXPORTS = ("pmx_init", "pmx_print_info", "pmx_set_kappa", "pmx_create_A", "pmx_setup",
           "pmx_setup_from_basis", "pmx_apply", "pmx_solve", "pmx_solve_device",
           "pmx_get_kernel_times", "pmx_get_stats", "pmx_print_stat", "pmx_get_basis",
           "pmx_get_coarse_matrix", "pmx_final", "pmx_last_error", "pmx_cnn_num_params",
           "pmx_cnn_create", "pmx_cnn_get_params", "pmx_cnn_set_params", "pmx_cnn_predict",
           "pmx_cnn_loss_grad", "pmx_cnn_adam_step", "pmx_cnn_destroy", "pmx_cnn_inputs",
           "pmx_cnn_targets", "pmx_set_warm_start", "pmx_set_kappa_device",
           "pmx_set_well_term", "pmx_set_well_term_device", "pmx_velocity",
           "pmx_velocity_device", "pmx_spe10_create", "pmx_spe10_assemble",
           "pmx_spe10_get_system", "pmx_spe10_solve", "pmx_spe10_transport",
           "pmx_spe10_saturation_range", "pmx_spe10_get_state", "pmx_spe10_set_state",
           "pmx_spe10_run", "pmx_spe10_destroy", "pmx_setup_batch", "pmx_spe10_set_step_hook",
           "pmx_solve_device_noalloc")
TIMED_KERNELS = ("k_lv1_first", "k_lv1_last", "k_resid_restrict", "k_prolong", "k_cg_p_spmv",
                 "k_cg_update", "k_gemv", "coarse_small")


class PMXError(RuntimeError):
    """A non-zero return code of the C ABI (values are PETSc error codes)."""

    def __init__(self, code: int, msg: str):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class Stats(C.Structure):
    _fields_ = [
        ("M", C.c_int), ("N", C.c_int), ("P", C.c_int), ("sd", C.c_int * 3),
        ("nsub", C.c_int), ("size_coarse", C.c_int), ("size_mat_lv3", C.c_int),
        ("eigen_num_lv1_range", C.c_int * 2), ("eigen_min_lv1_range", C.c_double * 2),
        ("eigen_max_lv1_range", C.c_double * 2), ("eigen_num_lv2", C.c_int),
        ("eigen_min_lv2", C.c_double), ("eigen_max_lv2", C.c_double),
        ("t_stages", C.c_double * 10), ("eig_iters_max", C.c_int),
        ("eig_iters_mean", C.c_double), ("eig_fallbacks", C.c_int), ("lv3_iters", C.c_int),
        ("pc_precision", C.c_int), ("its", C.c_int), ("rnorm", C.c_double),
        ("rnorm0", C.c_double), ("reason", C.c_int), ("t_solve", C.c_double),
        ("eig_stagnated", C.c_int), ("eig_block", C.c_int), ("coarse_mode", C.c_int),
        ("coarse_stages", C.c_int), ("lv3_floor", C.c_double),
        ("coarse_factor_bytes", C.c_double), ("coarse_solve_bytes", C.c_double),
        ("t_eig_tune", C.c_double), ("eig_shape", C.c_int),
        ("t_eig_kernel", C.c_double),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class Spe10Log(C.Structure):
    """One row of the IMPES log (pmx_spe10_log)."""
    _fields_ = [
        ("step", C.c_int), ("simu_time", C.c_double), ("stage", C.c_int),
        ("DS_max", C.c_double), ("S_w_steps", C.c_int), ("its", C.c_int), ("reason", C.c_int),
        ("S_w_min", C.c_double), ("S_w_max", C.c_double), ("delta_t", C.c_double),
        ("t_assemble", C.c_double), ("t_setup", C.c_double), ("t_solve", C.c_double),
        ("t_transport", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Per-step hook of pmx_spe10_run: int (*)(void *user, const pmx_spe10_log *row).
STEP_HOOK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(Spe10Log))

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load the library once; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run premixfem_amd.build.build() first")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.pmx_init.argtypes = [C.POINTER(vp), _dp, _ip, C.c_char_p]
    L.pmx_print_info.argtypes = [vp]
    L.pmx_set_kappa.argtypes = [vp, _dp, _dp, _dp]
    L.pmx_create_A.argtypes = [vp, _dp, _dp]
    L.pmx_setup.argtypes = [vp]
    L.pmx_setup_from_basis.argtypes = [vp, _dp, C.c_void_p, C.c_void_p]
    L.pmx_setup_batch.argtypes = [C.POINTER(vp), C.c_int]
    L.pmx_apply.argtypes = [vp, _dp, _dp]
    L.pmx_solve.argtypes = [vp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                            C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.pmx_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.pmx_print_stat.argtypes = [vp]
    L.pmx_get_basis.argtypes = [vp, C.c_int, _dp, C.c_void_p]
    L.pmx_get_coarse_matrix.argtypes = [vp, _dp]
    L.pmx_final.argtypes = [vp]
    L.pmx_solve_device.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.pmx_solve_device_noalloc.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                           C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.pmx_get_kernel_times.argtypes = [vp, _dp, _ip, C.c_int, C.c_int]
    L.pmx_cnn_create.argtypes = [C.POINTER(vp), C.c_int, C.c_void_p, C.c_ulonglong]
    L.pmx_cnn_get_params.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.pmx_cnn_set_params.argtypes = [vp, C.c_void_p]
    L.pmx_cnn_predict.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
    L.pmx_cnn_loss_grad.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int,
                                    C.POINTER(C.c_double)]
    L.pmx_cnn_adam_step.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    L.pmx_cnn_destroy.argtypes = [vp]
    L.pmx_cnn_inputs.argtypes = [vp, vp]
    L.pmx_cnn_targets.argtypes = [vp, C.c_int, vp]
    L.pmx_set_warm_start.argtypes = [vp, vp]
    L.pmx_set_kappa_device.argtypes = [vp, vp, vp, vp]
    L.pmx_set_well_term.argtypes = [vp, C.c_void_p]
    L.pmx_set_well_term_device.argtypes = [vp, vp]
    L.pmx_velocity.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.pmx_velocity_device.argtypes = [vp, vp, vp, vp, vp]
    L.pmx_spe10_create.argtypes = [C.POINTER(vp), _ip, _dp, _dp, _dp, _dp, _dp]
    L.pmx_spe10_assemble.argtypes = [vp, vp]
    L.pmx_spe10_get_system.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
    L.pmx_spe10_solve.argtypes = [vp, vp, C.c_double, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
    L.pmx_spe10_transport.argtypes = [vp, C.c_int, C.c_double, C.POINTER(C.c_double)]
    L.pmx_spe10_saturation_range.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pmx_spe10_get_state.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.pmx_spe10_set_state.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.pmx_spe10_run.argtypes = [vp, vp, C.c_int, _dp, _dp, _ip, C.c_double, C.c_int,
                                C.c_void_p, C.POINTER(C.c_int)]
    L.pmx_spe10_destroy.argtypes = [vp]
    L.pmx_spe10_set_step_hook.argtypes = [vp, C.c_void_p, C.c_void_p]
    L.pmx_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "pmx_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    if rc != PMX_SUCCESS:
        raise PMXError(rc, lib().pmx_last_error().decode())
