"""Python host mirror of the reference interface ``src/PreMixFEM_3D.h``.

Same function names and argument meaning as the reference's nine entry points
(``src/PreMixFEM_3D.h:47-67``), backed by the HIP library through its C ABI.
Errors raise :class:`premixfem_amd._lib.PMXError` carrying the PETSc error
code the reference would have returned.

Typical sequence (the one of ``src/Expe_scal.c:66-160``)::

    ctx = PCCtx()
    PC_init(ctx, dom, mesh, "-sd 4 -en_lv1 4")
    ctx.kappa = kappa                        # (3, P, N, M)
    diag, off = PC_create_A(ctx)
    pc = PCShell(ctx)
    PC_setup(pc)
    x, its, rnorm, reason = KSPSolve(pc, b, rtol=1e-8)
    PC_print_stat(ctx)
    PC_final(ctx)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PMXError, check, lib  # noqa: F401

STAGE_NAMES = ("su_lv1", "su_lv2_1", "su_lv2_2", "su_lv3_1", "su_lv3_2",
               "av_lv1", "av_lv2", "av_lv3", "su", "av")


class PCCtx:
    """Mirror of ``PCCtx`` (src/PreMixFEM_3D.h:32-45)."""

    def __init__(self):
        self.handle = None
        self.kappa = None          # (3, P, N, M) float64, natural order
        self.M = self.N = self.P = 0
        self.L = self.W = self.H = 0.0
        self.H_x = self.H_y = self.H_z = 0.0
        self.options = ""

    @property
    def ncell(self):
        return self.M * self.N * self.P

    def _require(self):
        if self.handle is None:
            raise PMXError(73, "PC_init has not been called")
        return self.handle


class PCShell:
    """The PETSc PCSHELL the reference registers (src/Expe_scal.c:104-112)."""

    def __init__(self, ctx: PCCtx, name="3levels-MG-via-GMsFEM-with-velocity-elimination"):
        self.ctx = ctx
        self.name = name


def PC_init(s_ctx: PCCtx, dom, mesh, options: str = "") -> None:
    """PC_init (src/PreMixFEM_3D.c:1650-1741)."""
    dom = np.ascontiguousarray(dom, dtype=np.float64)
    mesh_a = np.ascontiguousarray(mesh, dtype=np.int32)
    h = C.c_void_p()
    check(lib().pmx_init(C.byref(h), dom, mesh_a, options.encode()))
    s_ctx.handle = h
    s_ctx.L, s_ctx.W, s_ctx.H = (float(v) for v in dom)
    s_ctx.M, s_ctx.N, s_ctx.P = (int(v) for v in mesh_a)
    s_ctx.H_x, s_ctx.H_y, s_ctx.H_z = (s_ctx.L / s_ctx.M, s_ctx.W / s_ctx.N, s_ctx.H / s_ctx.P)
    s_ctx.options = options


def PC_print_info(s_ctx: PCCtx) -> None:
    """PC_print_info (src/PreMixFEM_3D.c:1743-1755)."""
    check(lib().pmx_print_info(s_ctx._require()))


def _push_kappa(s_ctx: PCCtx):
    if s_ctx.kappa is None:
        raise PMXError(73, "s_ctx.kappa has not been filled")
    k = np.asarray(s_ctx.kappa, dtype=np.float64)
    if k.shape != (3, s_ctx.P, s_ctx.N, s_ctx.M):
        raise PMXError(62, f"kappa must have shape (3, {s_ctx.P}, {s_ctx.N}, {s_ctx.M})")
    check(lib().pmx_set_kappa(s_ctx._require(), *(np.ascontiguousarray(k[a]).ravel()
                                                   for a in range(3))))


def PC_create_A(s_ctx: PCCtx):
    """PC_create_A (src/PreMixFEM_3D.c:1757-1824): returns (diag, off) with
    off[c, d] the coupling of cell c to its -x,+x,-y,+y,-z,+z neighbour."""
    _push_kappa(s_ctx)
    diag = np.zeros(s_ctx.ncell)
    off = np.zeros(6 * s_ctx.ncell)
    check(lib().pmx_create_A(s_ctx._require(), diag, off))
    return diag, off.reshape(s_ctx.ncell, 6)


def PC_setup(pc: PCShell) -> None:
    """PC_setup (src/PreMixFEM_3D.c:1826-1838)."""
    _push_kappa(pc.ctx)
    check(lib().pmx_setup(pc.ctx._require()))


def PC_setup_batch(pcs) -> None:
    """PC_setup of several independent problems at once: each context is set
    up as PC_setup does (src/PreMixFEM_3D.c:1826-1838), with the subdomain
    eigenproblems of all of them solved by one batched LOBPCG launch."""
    pcs = list(pcs)
    for pc in pcs:
        _push_kappa(pc.ctx)
    arr = (C.c_void_p * len(pcs))(*[pc.ctx._require() for pc in pcs])
    check(lib().pmx_setup_batch(arr, len(pcs)))


def PC_setup_from_basis(pc: PCShell, phi, eigen_num=None, eigvals=None) -> None:
    """PC_setup with the level-2 eigenvectors given (shape (en_lv1, ncell))."""
    _push_kappa(pc.ctx)
    phi = np.ascontiguousarray(phi, dtype=np.float64).ravel()
    en = None if eigen_num is None else np.ascontiguousarray(eigen_num, dtype=np.int32)
    ev = None if eigvals is None else np.ascontiguousarray(eigvals, dtype=np.float64)
    check(lib().pmx_setup_from_basis(pc.ctx._require(), phi,
                                     None if en is None else en.ctypes.data,
                                     None if ev is None else ev.ctypes.data))


def PC_apply_vec(pc: PCShell, x, y=None):
    """PC_apply_vec (src/PreMixFEM_3D.c:1840-1936): y = B x."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    if y is None:
        y = np.zeros_like(x)
    check(lib().pmx_apply(pc.ctx._require(), x, y))
    return y


def KSPSolve(pc: PCShell, b, rtol=1e-8, x0=None, max_it=10000):
    """-ksp_type cg with KSP_NORM_UNPRECONDITIONED on the GPU
    (src/Expe_scal.c:101-119).  Returns (x, its, rnorm, reason)."""
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).ravel().copy()
    its = C.c_int(0)
    rn = C.c_double(0.0)
    rc = lib().pmx_solve(pc.ctx._require(), b, x, int(x0 is not None), float(rtol), int(max_it),
                         C.byref(its), C.byref(rn))
    st = PC_get_stats(pc.ctx)
    if rc not in (0, 91):
        check(rc)
    return x, its.value, rn.value, _lib.REASONS.get(st["reason"], st["reason"])


def PC_get_stats(s_ctx: PCCtx) -> dict:
    st = _lib.Stats()
    check(lib().pmx_get_stats(s_ctx._require(), C.byref(st)))
    d = st.as_dict()
    d["t_stages"] = dict(zip(STAGE_NAMES, d["t_stages"]))
    return d


def PC_get_range(value):
    """PC_get_range (src/PreMixFEM_3D.c:1938-1952): min/max over ranks; one
    GPU is one rank, so both equal the local value."""
    return [value, value]


def PC_print_stat(s_ctx: PCCtx) -> None:
    """PC_print_stat (src/PreMixFEM_3D.c:1954-1999)."""
    check(lib().pmx_print_stat(s_ctx._require()))


def PC_get_basis(s_ctx: PCCtx, level: int = 2):
    """ms_bases_c (level 2: (en_lv1, ncell)) or ms_bases_cc (level 3)."""
    st = PC_get_stats(s_ctx)
    k = int(s_ctx.options_en_lv1()) if level == 2 else st["size_mat_lv3"]
    out = np.zeros(k * s_ctx.ncell)
    nev = st["nsub"] * k if level == 2 else max(k, 1)
    ev = np.zeros(max(nev, 16))
    check(lib().pmx_get_basis(s_ctx._require(), level, out, ev.ctypes.data))
    return out.reshape(k, s_ctx.ncell), ev[:nev]


def PC_get_coarse_matrix(s_ctx: PCCtx):
    st = PC_get_stats(s_ctx)
    n = st["size_coarse"]
    out = np.zeros(n * n)
    check(lib().pmx_get_coarse_matrix(s_ctx._require(), out))
    return out.reshape(n, n)


def PC_final_default(s_ctx: PCCtx) -> None:
    """PC_final_default (src/PreMixFEM_3D.c:2028-2046)."""
    if s_ctx.handle is not None:
        check(lib().pmx_final(s_ctx.handle))
        s_ctx.handle = None


def PC_final(s_ctx: PCCtx) -> None:
    """PC_final (src/PreMixFEM_3D.c:2001-2026)."""
    PC_final_default(s_ctx)


def _en_lv1(self):
    toks = self.options.split()
    for i, t in enumerate(toks):
        if t == "-en_lv1" and i + 1 < len(toks):
            return int(toks[i + 1])
    return 4


PCCtx.options_en_lv1 = _en_lv1
