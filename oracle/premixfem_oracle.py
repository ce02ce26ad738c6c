"""FP64 CPU oracle for PreMixFEM's three-level GMsFEM preconditioner.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``premixfem_amd``) never calls it.

It restates, in numpy/scipy, the live part of ``src/PreMixFEM_3D.c:1-2046``
and the outer PETSc CG that ``src/Expe_scal.c:100-119`` drives:

* ``partition``            -- ``_PC_setup_lv1_partition`` (:47-78)
* ``transmissibilities``   -- the TPFA face coefficient of ``PC_create_A`` (:1777-1816)
* ``assemble_A``           -- ``PC_create_A`` (:1757-1824) (+ optional pin)
* ``dic_factor/dic_solve`` -- level-1 block ICC(0) of ``_PC_setup_lv1_J`` (:80-222),
                              applied by ``_PC_apply_vec_lv1`` (:1429-1479)
* ``ExactBlocks``          -- the same blocks factored exactly
                              (``-use_full_Cholesky_lv1``, :205-207)
* ``mass_diag``            -- the eigen mass matrices of ``_PC_setup_lv2_eigen`` (:265-343)
* ``ghep``                 -- the per-subdomain generalized eigenproblems (:345-390)
* ``Oracle.setup``         -- ``_PC_setup`` (:1392-1427): lv2 Galerkin (:737-916),
                              lv3 eigen (:918-1175), lv3 coarse-coarse (:1177-1390)
* ``Oracle.apply``         -- ``PC_apply_vec`` (:1840-1936)
* ``Layout(procs=...)``    -- process boxes: DMDA ownership (:1711), subdomains per
                              box (:58-72); ``Oracle._lv3_boxes`` and
                              ``energy_gram(inner=True)`` -- the per-process
                              level 2 (:806-888) and level 3 (:918-1390)
* ``pcg``                  -- PETSc ``KSPSolve_CG`` with ``KSP_NORM_UNPRECONDITIONED``
                              (``src/Expe_scal.c:114``) and the default convergence test.

The eigensolver boundary (SLEPc Krylov-Schur + shift-invert, ``:354-365``)
is restated with ARPACK implicitly-restarted Lanczos in shift-invert mode
(``scipy.sparse.linalg.eigsh``), or a dense ``scipy.linalg.eigh`` on small
blocks; both return the same eigenpairs up to sign, which is normalised here.
The coarse Cholesky / LU boundary (MUMPS :900, SuperLU_dist :1371) is a dense
LAPACK Cholesky / LU.

Arrays use the reference's natural layout ``arr[ez][ey][ex]`` = numpy shape
(P, N, M); flattened index ``(ez*N + ey)*M + ex``.
"""
from __future__ import annotations

import time
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

NIL = 1.0e-12
EIG_OP_MOD, EIG_OP_MEAN, EIG_OP_DIAG = 0, 1, 2
STAGES = ("su_lv1", "su_lv2_1", "su_lv2_2", "su_lv3_1", "su_lv3_2",
          "av_lv1", "av_lv2", "av_lv3", "su", "av")


# --------------------------------------------------------------------------
# Partition  (src/PreMixFEM_3D.c:47-78, :23-45)
# --------------------------------------------------------------------------
def partition(n: int, sd: int):
    """len_i = n/sd + (n%sd > i); start_i = start_{i-1} + len_{i-1} (:62-72)."""
    lens = [n // sd + (1 if n % sd > i else 0) for i in range(sd)]
    starts = [0] * sd
    for i in range(1, sd):
        starts[i] = starts[i - 1] + lens[i - 1]
    return starts, lens


def sub_id(sx, sy, sz, sd):
    """x-fastest subdomain id (:23-27) for per-direction counts sd=(sdx,sdy,sdz)."""
    return sx + sd[0] * (sy + sd[1] * sz)


def dof_idx(eigen_num):
    """Prefix sum of per-subdomain eigen counts (:40-45)."""
    out = np.zeros(len(eigen_num) + 1, dtype=np.int64)
    out[1:] = np.cumsum(eigen_num)
    return out


def dmda_ownership(n: int, p: int):
    """PETSc's default DMDA ownership of n cells over p processes along one
    axis (DMSetUp_DA_3D: l_j = n/p + (n%p > j)), as (starts, lens)."""
    return partition(n, p)


class Layout:
    """Grid geometry, partition and per-cell local coordinates.

    ``procs`` = (px, py, pz) process boxes (the DMDA decomposition of
    ``src/PreMixFEM_3D.c:1711``): each box is cut into ``sd`` subdomains per
    axis by ``_PC_setup_lv1_partition`` (:47-78) from its own corners, so the
    global subdomain grid has px*sdx x py*sdy x pz*sdz entries.  ``self.sd``
    holds those global counts; ``self.sd_box`` the per-box ones."""

    def __init__(self, dom, mesh, sd, procs=(1, 1, 1)):
        self.dom = tuple(float(v) for v in dom)
        self.mesh = tuple(int(v) for v in mesh)
        if isinstance(sd, int):
            sd = (sd, sd, sd)
        # The reference resets sd == 1 to 2 (:1718-1721).
        sd_box = tuple(2 if s == 1 else int(s) for s in sd)
        self.procs = tuple(int(p) for p in procs)
        self.sd_box = sd_box
        self.sd = tuple(p * s for p, s in zip(self.procs, sd_box))
        M, N, P = self.mesh
        self.Hx, self.Hy, self.Hz = (self.dom[0] / M, self.dom[1] / N, self.dom[2] / P)
        self.meas_elem = self.Hx * self.Hy * self.Hz
        self.meas_face_yz = self.Hy * self.Hz
        self.meas_face_zx = self.Hz * self.Hx
        self.meas_face_xy = self.Hx * self.Hy
        self.parts = []
        self.box_parts = []
        for n, p, s in zip(self.mesh, self.procs, sd_box):
            bst, bln = dmda_ownership(n, p)
            if min(bln) < s:
                raise ValueError("a process box has fewer cells than subdomains")
            starts, lens = [], []
            for b0, bl in zip(bst, bln):
                st, ln = partition(bl, s)
                starts += [b0 + v for v in st]
                lens += ln
            self.parts.append((starts, lens))
            self.box_parts.append((bst, bln))
        self.nsub = self.sd[0] * self.sd[1] * self.sd[2]
        self.nbox = self.procs[0] * self.procs[1] * self.procs[2]
        # Per-axis subdomain index and local coordinate of each cell.
        self.ax_sub, self.ax_loc = [], []
        for (starts, lens), n in zip(self.parts, self.mesh):
            s_of = np.zeros(n, dtype=np.int64)
            l_of = np.zeros(n, dtype=np.int64)
            for i, (st, ln) in enumerate(zip(starts, lens)):
                s_of[st:st + ln] = i
                l_of[st:st + ln] = np.arange(ln)
            self.ax_sub.append(s_of)
            self.ax_loc.append(l_of)

    def corners(self, s):
        """(startx, nx, starty, ny, startz, nz) of subdomain s (:29-38)."""
        sdx, sdy, _ = self.sd
        sx, sy, sz = s % sdx, (s // sdx) % sdy, s // (sdx * sdy)
        (stx, lnx), (sty, lny), (stz, lnz) = self.parts
        return stx[sx], lnx[sx], sty[sy], lny[sy], stz[sz], lnz[sz]

    def cell_sub(self):
        """Subdomain id of every cell, shape (P, N, M)."""
        Z, Y, X = np.meshgrid(self.ax_sub[2], self.ax_sub[1], self.ax_sub[0], indexing="ij")
        return X + self.sd[0] * (Y + self.sd[1] * Z)

    def sub_box(self, s):
        """Process box (x-fastest rank) holding subdomain s."""
        sdx, sdy, _ = self.sd
        sx, sy, sz = s % sdx, (s // sdx) % sdy, s // (sdx * sdy)
        bx, by, bz = sx // self.sd_box[0], sy // self.sd_box[1], sz // self.sd_box[2]
        return bx + self.procs[0] * (by + self.procs[1] * bz)

    def axis_box(self, a):
        """Process-box index of every coordinate along axis a."""
        return self.ax_sub[a] // self.sd_box[a]


# --------------------------------------------------------------------------
# Operator  (src/PreMixFEM_3D.c:1757-1824)
# --------------------------------------------------------------------------
def transmissibilities(lay: Layout, kappa):
    """Face coefficients T = (|F|^2/|E|) * 2/(1/k_L + 1/k_R) (:1785-1786).

    Returns (Tx, Ty, Tz) with shapes (P,N,M-1), (P,N-1,M), (P-1,N,M); Tx[z,y,x]
    is the face between cells x and x+1."""
    kx, ky, kz = kappa
    cx = lay.meas_face_yz * lay.meas_face_yz / lay.meas_elem
    cy = lay.meas_face_zx * lay.meas_face_zx / lay.meas_elem
    cz = lay.meas_face_xy * lay.meas_face_xy / lay.meas_elem
    Tx = cx * (2.0 / (1.0 / kx[:, :, :-1] + 1.0 / kx[:, :, 1:]))
    Ty = cy * (2.0 / (1.0 / ky[:, :-1, :] + 1.0 / ky[:, 1:, :]))
    Tz = cz * (2.0 / (1.0 / kz[:-1, :, :] + 1.0 / kz[1:, :, :]))
    return Tx, Ty, Tz


def cell_faces(lay: Layout, T):
    """Per-cell face coefficients (lower and upper in x, y, z), 0 on the
    domain boundary (no-flow).  Each is shape (P, N, M)."""
    Tx, Ty, Tz = T
    M, N, P = lay.mesh
    z = np.zeros((P, N, M))
    xl, xh, yl, yh, zl, zh = (z.copy() for _ in range(6))
    xl[:, :, 1:] = Tx
    xh[:, :, :-1] = Tx
    yl[:, 1:, :] = Ty
    yh[:, :-1, :] = Ty
    zl[1:, :, :] = Tz
    zh[:-1, :, :] = Tz
    return xl, xh, yl, yh, zl, zh


def pin_tau(lay: Layout, T, pin_cell):
    """Pin strength: the largest face coefficient of the pinned cell."""
    x, y, z = pin_cell
    f = cell_faces(lay, T)
    return max(float(a[z, y, x]) for a in f)


def assemble_A(lay: Layout, T, pin_cell=None):
    """7-point TPFA matrix in natural order (+ tau at the pin cell)."""
    M, N, P = lay.mesh
    n = M * N * P
    xl, xh, yl, yh, zl, zh = cell_faces(lay, T)
    diag = (xl + xh + yl + yh + zl + zh).ravel()
    if pin_cell is not None:
        x, y, z = pin_cell
        diag[(z * N + y) * M + x] += pin_tau(lay, T, pin_cell)
    idx = np.arange(n).reshape(P, N, M)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [diag]
    for Tf, a, b in ((T[0], idx[:, :, :-1], idx[:, :, 1:]),
                     (T[1], idx[:, :-1, :], idx[:, 1:, :]),
                     (T[2], idx[:-1, :, :], idx[1:, :, :])):
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
        vals += [-Tf.ravel(), -Tf.ravel()]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sum_duplicates()
    return A


def stencil_apply(lay: Layout, T, x, pin=None):
    """y = A x with the face arrays (matrix-free form of ``assemble_A``)."""
    Tx, Ty, Tz = T
    M, N, P = lay.mesh
    x = np.asarray(x, dtype=np.float64).reshape(P, N, M)
    y = np.zeros_like(x)
    d = Tx * (x[:, :, :-1] - x[:, :, 1:])
    y[:, :, :-1] += d
    y[:, :, 1:] -= d
    d = Ty * (x[:, :-1, :] - x[:, 1:, :])
    y[:, :-1, :] += d
    y[:, 1:, :] -= d
    d = Tz * (x[:-1, :, :] - x[1:, :, :])
    y[:-1, :, :] += d
    y[1:, :, :] -= d
    if pin is not None:
        (px, py, pz), tau = pin
        y[pz, py, px] += tau * x[pz, py, px]
    return y.ravel()


# --------------------------------------------------------------------------
# Level 1: block ICC(0)  (src/PreMixFEM_3D.c:80-222, :1429-1479)
# --------------------------------------------------------------------------
def lv1_block_diag(lay: Layout, T, pin=None):
    """Diagonal of the level-1 blocks A[Omega_i, Omega_i] in the accumulation
    order of ``_PC_setup_lv1_J`` (:93-197): the cell's own iteration adds its
    lower x face, its upper x face if it is the last cell of the row, then the
    same for y and z; the iterations of its upper neighbours inside the
    subdomain then add the remaining upper faces (x, then y, then z).  The
    pin term (if any) is added last."""
    M, N, P = lay.mesh
    xl, xh, yl, yh, zl, zh = cell_faces(lay, T)
    lx, ly, lz = lay.ax_loc
    nx_of = np.array(lay.parts[0][1])[lay.ax_sub[0]]
    ny_of = np.array(lay.parts[1][1])[lay.ax_sub[1]]
    nz_of = np.array(lay.parts[2][1])[lay.ax_sub[2]]
    LZ, LY, LX = np.meshgrid(lz, ly, lx, indexing="ij")
    NZ, NY, NX = np.meshgrid(nz_of, ny_of, nx_of, indexing="ij")
    lastx, lasty, lastz = LX == NX - 1, LY == NY - 1, LZ == NZ - 1
    d = xl.copy()
    d = d + np.where(lastx, xh, 0.0)
    d = d + yl
    d = d + np.where(lasty, yh, 0.0)
    d = d + zl
    d = d + np.where(lastz, zh, 0.0)
    d = d + np.where(~lastx, xh, 0.0)
    d = d + np.where(~lasty, yh, 0.0)
    d = d + np.where(~lastz, zh, 0.0)
    d = d.ravel()
    if pin is not None:
        (x, y, z), tau = pin
        d[(z * N + y) * M + x] += tau
    return d


class DIC:
    """Block ICC(0) of the principal submatrices A[Omega_i, Omega_i].

    The 7-point graph is bipartite, so ICC(0) in natural ordering has no
    off-diagonal updates: L_ij = a_ij / d_j and
    d_i = a_ii - sum_{j<i, j~i} a_ij^2 / d_j, with j restricted to the same
    subdomain.  The preconditioner is (D+L) D^{-1} (D+L^T).  Cells are
    processed by local hyperplanes lx+ly+lz = w, all subdomains at once."""

    def __init__(self, lay: Layout, T, A_diag):
        M, N, P = lay.mesh
        xl, xh, yl, yh, zl, zh = cell_faces(lay, T)
        lx, ly, lz = lay.ax_loc
        LZ, LY, LX = np.meshgrid(lz, ly, lx, indexing="ij")
        # Couplings inside the subdomain only.
        self.cxl = np.where(LX > 0, xl, 0.0).ravel()
        self.cyl = np.where(LY > 0, yl, 0.0).ravel()
        self.czl = np.where(LZ > 0, zl, 0.0).ravel()
        nx_of = np.array(lay.parts[0][1])[lay.ax_sub[0]]
        ny_of = np.array(lay.parts[1][1])[lay.ax_sub[1]]
        nz_of = np.array(lay.parts[2][1])[lay.ax_sub[2]]
        NZ, NY, NX = np.meshgrid(nz_of, ny_of, nx_of, indexing="ij")
        self.cxh = np.where(LX < NX - 1, xh, 0.0).ravel()
        self.cyh = np.where(LY < NY - 1, yh, 0.0).ravel()
        self.czh = np.where(LZ < NZ - 1, zh, 0.0).ravel()
        self.sx, self.sy, self.sz = 1, M, M * N
        wave = (LX + LY + LZ).ravel()
        order = np.argsort(wave, kind="stable")
        counts = np.bincount(wave)
        self.waves = np.split(order, np.cumsum(counts)[:-1])
        n = M * N * P
        d = np.zeros(n)
        a = np.asarray(A_diag, dtype=np.float64)
        for ids in self.waves:
            v = a[ids].copy()
            for c, s in ((self.czl, self.sz), (self.cyl, self.sy), (self.cxl, self.sx)):
                cc = c[ids]
                m = cc != 0.0
                v[m] -= cc[m] * cc[m] / d[ids[m] - s]
            d[ids] = v
        self.d = d
        self.dinv = 1.0 / d

    def solve(self, r):
        """y = [(D+L) D^{-1} (D+L^T)]^{-1} r, block by block."""
        r = np.asarray(r, dtype=np.float64).ravel()
        w = np.zeros_like(r)
        for ids in self.waves:
            v = r[ids].copy()
            for c, s in ((self.cxl, self.sx), (self.cyl, self.sy), (self.czl, self.sz)):
                cc = c[ids]
                m = cc != 0.0
                v[m] += cc[m] * w[ids[m] - s]
            w[ids] = v * self.dinv[ids]
        y = np.zeros_like(r)
        for ids in reversed(self.waves):
            v = np.zeros(len(ids))
            for c, s in ((self.cxh, self.sx), (self.cyh, self.sy), (self.czh, self.sz)):
                cc = c[ids]
                m = cc != 0.0
                v[m] += cc[m] * y[ids[m] + s]
            y[ids] = w[ids] + v * self.dinv[ids]
        return y


class ExactBlocks:
    """Level-1 blocks A[Omega_i, Omega_i] solved exactly
    (``-use_full_Cholesky_lv1``: PCCHOLESKY with CHOLMOD on every block,
    src/PreMixFEM_3D.c:205-207).  The blocks are those of :class:`DIC`: the
    block diagonal ``A_diag`` and the couplings inside each subdomain; the
    block-diagonal matrix is factored once by a sparse direct solver."""

    def __init__(self, lay: Layout, T, A_diag):
        dic = DIC(lay, T, A_diag)
        n = len(A_diag)
        idx = np.arange(n)
        rows, cols, vals = [idx], [idx], [np.asarray(A_diag, dtype=np.float64)]
        for c, s in ((dic.cxl, dic.sx), (dic.cyl, dic.sy), (dic.czl, dic.sz)):
            m = c != 0.0
            i = idx[m]
            rows += [i, i - s]
            cols += [i - s, i]
            vals += [-c[m], -c[m]]
        self.B = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(n, n))
        self._solve = spla.factorized(self.B)

    def solve(self, r):
        return self._solve(np.asarray(r, dtype=np.float64).ravel())


# --------------------------------------------------------------------------
# Level 2: local eigenproblems  (src/PreMixFEM_3D.c:224-421)
# --------------------------------------------------------------------------
def mass_diag(lay: Layout, kappa, T, eg_op=EIG_OP_MOD, wd=None):
    """Per-cell diagonal of the eigen mass matrix M_i (natural order).

    MOD (:265-343): 2*kbar/H^2 face weights scaled by |E|.  In x a cell gets
    its upper face once, its upper face again when the upper neighbour is in
    the same subdomain (the face term is added to cell ex-1 at :266), and its
    lower face only when it is the subdomain's first cell.  In y and z every
    cell gets both of its faces.  Domain-boundary faces never count.
    MEAN: kx+ky+kz.  DIAG: diag of the Neumann block A_i^N, including the
    well term ``wd`` when there is one (``src/spe10.c:990-1012``)."""
    M, N, P = lay.mesh
    kx, ky, kz = kappa
    if eg_op == EIG_OP_MEAN:
        return (kx + ky + kz).ravel().copy()
    lx, ly, lz = lay.ax_loc
    nx_of = np.array(lay.parts[0][1])[lay.ax_sub[0]]
    ny_of = np.array(lay.parts[1][1])[lay.ax_sub[1]]
    nz_of = np.array(lay.parts[2][1])[lay.ax_sub[2]]
    if eg_op == EIG_OP_DIAG:
        xl, xh, yl, yh, zl, zh = cell_faces(lay, T)
        LZ, LY, LX = np.meshgrid(lz, ly, lx, indexing="ij")
        NZ, NY, NX = np.meshgrid(nz_of, ny_of, nx_of, indexing="ij")
        m = (np.where(LX > 0, xl, 0.0) + np.where(LX < NX - 1, xh, 0.0)
             + np.where(LY > 0, yl, 0.0) + np.where(LY < NY - 1, yh, 0.0)
             + np.where(LZ > 0, zl, 0.0) + np.where(LZ < NZ - 1, zh, 0.0))
        m = m.ravel()
        if wd is not None:
            m = m + wd
        return m
    if eg_op != EIG_OP_MOD:
        raise ValueError("eg_op")
    ax = 2.0 * (2.0 / (1.0 / kx[:, :, :-1] + 1.0 / kx[:, :, 1:])) / lay.Hx / lay.Hx
    ay = 2.0 * (2.0 / (1.0 / ky[:, :-1, :] + 1.0 / ky[:, 1:, :])) / lay.Hy / lay.Hy
    az = 2.0 * (2.0 / (1.0 / kz[:-1, :, :] + 1.0 / kz[1:, :, :])) / lay.Hz / lay.Hz
    z = np.zeros((P, N, M))
    lo_x, hi_x, lo_y, hi_y, lo_z, hi_z = (z.copy() for _ in range(6))
    hi_x[:, :, :-1] = ax
    lo_x[:, :, 1:] = ax
    hi_y[:, :-1, :] = ay
    lo_y[:, 1:, :] = ay
    hi_z[:-1, :, :] = az
    lo_z[1:, :, :] = az
    first = (lx == 0)[None, None, :]
    notlast = np.zeros(M, dtype=bool)
    notlast[:-1] = lx[1:] > 0
    # Accumulation order of the reference loop for one cell: its own
    # iteration adds [lower x if first], upper x, lower y, upper y, lower z,
    # upper z; the next cell's iteration then adds upper x again.
    m = np.where(first, lo_x, 0.0)
    m = m + hi_x
    m = m + lo_y
    m = m + hi_y
    m = m + lo_z
    m = m + hi_z
    m = m + np.where(notlast[None, None, :], hi_x, 0.0)
    return (m * lay.meas_elem).ravel()


def local_neumann(lay: Layout, T, s, wd=None):
    """A_i^N of subdomain s (interior faces only, :254-264), local order,
    plus the subdomain's well term ``wd`` if given (``src/spe10.c:1005-1012``)."""
    stx, nx, sty, ny, stz, nz = lay.corners(s)
    Tx, Ty, Tz = T
    tx = Tx[stz:stz + nz, sty:sty + ny, stx:stx + nx - 1]
    ty = Ty[stz:stz + nz, sty:sty + ny - 1, stx:stx + nx]
    tz = Tz[stz:stz + nz - 1, sty:sty + ny, stx:stx + nx]
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for t, a, b in ((tx, idx[:, :, :-1], idx[:, :, 1:]),
                    (ty, idx[:, :-1, :], idx[:, 1:, :]),
                    (tz, idx[:-1, :, :], idx[1:, :, :])):
        a, b, t = a.ravel(), b.ravel(), t.ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [-t, -t]
        np.add.at(diag, a, t)
        np.add.at(diag, b, t)
    if wd is not None:
        M, N, _ = lay.mesh
        diag = diag + np.asarray(wd).reshape(-1, N, M)[stz:stz + nz, sty:sty + ny,
                                                     stx:stx + nx].ravel()
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def sign_normalize(V, weights=None):
    """Fix the sign of each column: positive inner product with a fixed
    non-symmetric weight vector (eigenvectors are defined up to sign)."""
    n = V.shape[0]
    if weights is None:
        weights = sign_weights(n)
    s = np.sign(weights @ V)
    s[s == 0] = 1.0
    return V * s


def sign_weights(n):
    """Deterministic weights 1 + frac(0.618.. * (i+1)) used for sign fixing."""
    i = np.arange(n, dtype=np.float64)
    return 1.0 + np.mod(0.6180339887498949 * (i + 1.0), 1.0)


def ghep(A_N, m, k, dense_max=1200):
    """Smallest k eigenpairs of A_N phi = lam diag(m) phi, M-normalised.

    Dense ``eigh`` for small blocks, shift-invert Lanczos at -NIL otherwise
    (the SLEPc setup of :354-365)."""
    n = A_N.shape[0]
    if n <= dense_max or k >= n - 1:
        lam, V = sla.eigh(A_N.toarray(), np.diag(m), subset_by_index=[0, k - 1])
    else:
        lam, V = spla.eigsh(A_N.tocsc(), k=k, M=sp.diags(m).tocsc(), sigma=-NIL, which="LM",
                            tol=1e-12)
        o = np.argsort(lam)
        lam, V = lam[o], V[:, o]
    nrm = np.sqrt(np.einsum("ij,i,ij->j", V, m, V))
    V = V / nrm
    return lam, sign_normalize(V)


# --------------------------------------------------------------------------
# Level 2 of the spe10 fork: ICC(0) of the coarse matrix  (src/spe10.c:509-516,
# :1072-1213)
# --------------------------------------------------------------------------
DENSE_COARSE_MAX = 6000


class BlockICC:
    """PETSc ICC(0) (natural ordering, no fill) of the Galerkin matrix A_c.

    A_c's pattern is block 7-point over the subdomains: every subdomain's
    self block and its blocks with the six face neighbours are stored whole
    (``MatSetValues`` of every entry, :1083-1205).  Scalar IC(0) on that
    pattern equals the block incomplete factorisation below: between a
    subdomain I and a lower neighbour J the only common lower neighbour is J
    itself, so L_IJ = A_IJ L_JJ^{-T} exactly, while the dropped fill is that
    between I and the neighbours of J; inside a block nothing is dropped:
        D_I = A_II - sum_{J in {I-x, I-y, I-z}} A_IJ D_J^{-1} A_JI,
        M = (D + L) D^{-1} (D + L^T),  L = block strictly lower part of A_c.
    A non-positive pivot (which PETSc would answer with its positive-definite
    shift) raises ``np.linalg.LinAlgError``."""

    def __init__(self, lay: Layout, A_c, dof):
        self.lay = lay
        self.dof = np.asarray(dof)
        A = A_c.tocsr()
        sdx, sdy, sdz = lay.sd
        ns = lay.nsub
        self.nbr_lo, self.nbr_hi = [], []
        for s in range(ns):
            sx, sy, sz = s % sdx, (s // sdx) % sdy, s // (sdx * sdy)
            lo = [t for t, ok in ((s - 1, sx > 0), (s - sdx, sy > 0), (s - sdx * sdy, sz > 0)) if ok]
            hi = [t for t, ok in ((s + 1, sx < sdx - 1), (s + sdx, sy < sdy - 1),
                                  (s + sdx * sdy, sz < sdz - 1)) if ok]
            self.nbr_lo.append(lo)
            self.nbr_hi.append(hi)
        self.blk = {}

        def block(i, j):
            return A[self.dof[i]:self.dof[i + 1], self.dof[j]:self.dof[j + 1]].toarray()

        self.Dinv = []
        for s in range(ns):
            D = block(s, s)
            for t in self.nbr_lo[s]:
                Ast = block(s, t)
                self.blk[(s, t)] = Ast
                D = D - Ast @ self.Dinv[t] @ Ast.T
            np.linalg.cholesky(D)   # raises on a non-positive pivot
            self.Dinv.append(np.linalg.inv(D))
        for s in range(ns):
            for t in self.nbr_hi[s]:
                self.blk[(s, t)] = block(s, t)

    def solve(self, r):
        r = np.asarray(r, dtype=np.float64)
        ns, d = self.lay.nsub, self.dof
        w = np.zeros_like(r)
        for s in range(ns):
            v = r[d[s]:d[s + 1]].copy()
            for t in self.nbr_lo[s]:
                v -= self.blk[(s, t)] @ w[d[t]:d[t + 1]]
            w[d[s]:d[s + 1]] = self.Dinv[s] @ v
        y = np.zeros_like(r)
        for s in reversed(range(ns)):
            v = np.zeros(d[s + 1] - d[s])
            for t in self.nbr_hi[s]:
                v += self.blk[(s, t)] @ y[d[t]:d[t + 1]]
            y[d[s]:d[s + 1]] = w[d[s]:d[s + 1]] - self.Dinv[s] @ v
        return y


# --------------------------------------------------------------------------
# The preconditioner
# --------------------------------------------------------------------------
class Oracle:
    """``PCCtx`` + ``_PC_setup`` + ``PC_apply_vec`` on one process."""

    GRAM_CHUNK = 1 << 23   # faces per chunk of the energy-form Galerkin sums

    def __init__(self, dom, mesh, kappa, sd=2, en_lv1=4, en_lv2=4, eg_op=EIG_OP_MOD,
                 pin_cell=None, si_lv1=1, si_lv2=1, eb_lv1=-1.0, eb_lv2=-1.0,
                 use_W_cycle=False, no_shift_A_cc=False, use_2level=False, wd=None,
                 lv2_icc=False, use_full_Cholesky_lv1=False, procs=(1, 1, 1)):
        if min(dom) <= 0.0:
            raise ValueError("dom must be positive (:1653)")
        if min(mesh) <= 0:
            raise ValueError("mesh must be positive (:1654)")
        if en_lv1 < 1 or en_lv2 < 1:
            raise ValueError("eigen counts must be >= 1 (:1681,:1685)")
        if eg_op not in (EIG_OP_MOD, EIG_OP_MEAN, EIG_OP_DIAG):
            raise ValueError("eg_op (:1690)")
        self.lay = Layout(dom, mesh, sd, procs)
        self.kappa = np.asarray(kappa, dtype=np.float64)
        self.en_lv1, self.en_lv2, self.eg_op = int(en_lv1), int(en_lv2), int(eg_op)
        self.pin_cell = pin_cell
        self.si_lv1, self.si_lv2 = max(1, si_lv1), max(1, si_lv2)
        self.eb_lv1 = 1.0 / NIL if eb_lv1 < 0 else eb_lv1
        self.eb_lv2 = 1.0 / NIL if eb_lv2 < 0 else eb_lv2
        self.use_W_cycle, self.no_shift_A_cc, self.use_2level = use_W_cycle, no_shift_A_cc, use_2level
        self.use_full_Cholesky_lv1 = use_full_Cholesky_lv1
        self.t = dict.fromkeys(STAGES, 0.0)
        self.T = transmissibilities(self.lay, self.kappa)
        self.pin = None
        if pin_cell is not None:
            self.pin = (tuple(pin_cell), pin_tau(self.lay, self.T, pin_cell))
        self.A = assemble_A(self.lay, self.T, pin_cell)
        # Well term of the spe10 fork (src/spe10.c:780-786): a per-cell
        # diagonal addition to the operator, the level-1 blocks (:921-927),
        # the eigen stiffness and DIAG mass (:1005-1012) and, through A, the
        # coarse matrices.  lv2_icc: the level-2 factor is PETSc ICC(0) of
        # A_c instead of an exact factorisation (:516).
        self.wd = None if wd is None else np.asarray(wd, dtype=np.float64).ravel()
        if self.wd is not None:
            self.A = (self.A + sp.diags(self.wd)).tocsr()
        self.lv2_icc = bool(lv2_icc)

    def energy_gram(self, Phi, inner=False):
        """Phi^T A Phi in the reference's energy form: every face contributes
        T (phi_L - phi_R)(phi_L - phi_R)^T, computed from the differences of
        the basis values (:737-890, :1177-1360), never from A Phi.  For a face
        between two subdomains the difference row holds phi_L and -phi_R in
        different columns, which gives the reference's T phi phi^T diagonal
        and -T phi phi'^T neighbour blocks.  The pin and well terms add
        d phi phi^T.  Avoids the cancellation of A Phi near the null space.

        inner=True drops the faces between process boxes altogether: the
        box-local Neumann matrix A_i_inner of ``_PC_setup_lv3_eigen``
        (``ex != proc_startx`` ..., :939-1073)."""
        M, N, P = self.lay.mesh
        n = M * N * P
        idx = np.arange(n).reshape(P, N, M)
        bx, by, bz = (self.lay.axis_box(a) for a in range(3))
        keep = (np.broadcast_to(bx[:-1] == bx[1:], (P, N, M - 1)),
                np.broadcast_to((by[:-1] == by[1:])[:, None], (P, N - 1, M)),
                np.broadcast_to((bz[:-1] == bz[1:])[:, None, None], (P - 1, N, M)))
        fa, fb, tf = [], [], []
        for Tf, a, b, kf in ((self.T[0], idx[:, :, :-1], idx[:, :, 1:], keep[0]),
                             (self.T[1], idx[:, :-1, :], idx[:, 1:, :], keep[1]),
                             (self.T[2], idx[:-1, :, :], idx[1:, :, :], keep[2])):
            a, b, Tf = a.ravel(), b.ravel(), Tf.ravel()
            if inner:
                kf = kf.ravel()
                a, b, Tf = a[kf], b[kf], Tf[kf]
            fa.append(a)
            fb.append(b)
            tf.append(Tf)
        fa, fb, tf = np.concatenate(fa), np.concatenate(fb), np.concatenate(tf)
        # Faces in chunks (one chunk unless the grid is large): the sum over
        # faces is split, the per-face terms are unchanged.
        out = None
        for c0 in range(0, fa.size, self.GRAM_CHUNK):
            a, b = fa[c0:c0 + self.GRAM_CHUNK], fb[c0:c0 + self.GRAM_CHUNK]
            f = np.arange(a.size)
            inc = sp.csr_matrix((np.concatenate([np.ones(a.size), -np.ones(a.size)]),
                                 (np.concatenate([f, f]), np.concatenate([a, b]))),
                                shape=(a.size, n))
            G = (inc @ Phi).tocsr()
            part = G.T @ sp.diags(tf[c0:c0 + self.GRAM_CHUNK]) @ G
            out = part if out is None else out + part
        d = np.zeros(n)
        if self.pin is not None:
            (px, py, pz), tau = self.pin
            d[(pz * N + py) * M + px] += tau
        if self.wd is not None:
            d = d + self.wd
        if d.any():
            out = out + Phi.T @ sp.diags(d) @ Phi
        return sp.csr_matrix(out)

    def matvec(self, x):
        y = stencil_apply(self.lay, self.T, x, self.pin)
        if self.wd is not None:
            y = y + self.wd * np.asarray(x, dtype=np.float64).ravel()
        return y

    # ---- setup ----------------------------------------------------------
    def setup(self, ghep_fn=None):
        t0 = time.perf_counter()
        lay = self.lay
        # Level 1.
        ts = time.perf_counter()
        a_diag = lv1_block_diag(lay, self.T, self.pin)
        if self.wd is not None:
            a_diag = a_diag + self.wd
        self.dic = (ExactBlocks if self.use_full_Cholesky_lv1 else DIC)(lay, self.T, a_diag)
        self.t["su_lv1"] += time.perf_counter() - ts
        # Level 2 eigen.
        ts = time.perf_counter()
        self.mdiag = mass_diag(lay, self.kappa, self.T, self.eg_op, self.wd)
        ghep_fn = ghep_fn or ghep
        M, N, P = lay.mesh
        gidx = np.arange(M * N * P).reshape(P, N, M)
        self.sub_cells, self.eig_lv1, self.phi = [], [], []
        self.eigen_num_lv1 = np.zeros(lay.nsub, dtype=np.int64)
        self.eigen_min_lv1 = np.zeros(lay.nsub)
        self.eigen_max_lv1 = np.zeros(lay.nsub)
        for s in range(lay.nsub):
            stx, nx, sty, ny, stz, nz = lay.corners(s)
            cells = gidx[stz:stz + nz, sty:sty + ny, stx:stx + nx].ravel()
            AN = local_neumann(lay, self.T, s, self.wd)
            lam, V = ghep_fn(AN, self.mdiag[cells], self.en_lv1)
            # Adaptive count against eb_lv1 (:368-384).
            num = self.en_lv1
            for j in range(self.en_lv1):
                if lam[j] >= self.eb_lv1:
                    num = j + 1
                    break
            self.eigen_num_lv1[s] = num
            self.eigen_min_lv1[s] = lam[0]
            self.eigen_max_lv1[s] = lam[num - 1]
            self.sub_cells.append(cells)
            self.eig_lv1.append(lam)
            self.phi.append(V[:, :num])
        self.t["su_lv2_1"] += time.perf_counter() - ts
        # Level 2 Galerkin.
        ts = time.perf_counter()
        self.dof_idx = dof_idx(self.eigen_num_lv1)
        nc = int(self.dof_idx[-1])
        rows, cols, vals = [], [], []
        for s in range(lay.nsub):
            V = self.phi[s]
            for j in range(V.shape[1]):
                rows.append(self.sub_cells[s])
                cols.append(np.full(len(self.sub_cells[s]), self.dof_idx[s] + j))
                vals.append(V[:, j])
        self.Phi = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                 shape=(M * N * P, nc))
        A_c = self.energy_gram(self.Phi).tocsr()
        self.dof_box = np.repeat([lay.sub_box(s) for s in range(lay.nsub)], self.eigen_num_lv1)
        if lay.nbox > 1 and not self.use_2level:
            # _PC_setup_lv2_J (:806-888) couples only the subdomains of one
            # process: the blocks between boxes are dropped, the faces
            # between boxes stay on the diagonal blocks.  So does the spe10
            # fork's ICC(0) level 2 (spe10.c:1074-1205, one SBAIJ matrix per
            # process).  The 2-level variant's A_c keeps every coupling
            # (_PC_setup_lv2_c, :423-735: the neighbours' dof offsets are
            # exchanged, :448-479, and one matrix spans all processes).
            A_c = A_c.tocoo()
            same = self.dof_box[A_c.row] == self.dof_box[A_c.col]
            A_c = sp.csr_matrix((A_c.data[same], (A_c.row[same], A_c.col[same])), shape=A_c.shape)
        self.A_c_sp = (0.5 * (A_c + A_c.T)).tocsr()
        self.A_c = self.A_c_sp.toarray() if nc <= DENSE_COARSE_MAX else None
        if self.lv2_icc:
            self.bicc = BlockICC(lay, self.A_c_sp, self.dof_idx)
            self.t["su_lv2_2"] += time.perf_counter() - ts
            if self.use_2level:
                self.t["su"] += time.perf_counter() - t0
                return self
        elif self.use_2level:
            # _PC_setup_lv2_c (:423-735): same Galerkin matrix, NIL shift, LU.
            Ac = self.A_c + (0.0 if self.no_shift_A_cc else NIL) * np.eye(nc)
            self.lu_c = sla.lu_factor(Ac)
            self.t["su_lv2_2"] += time.perf_counter() - ts
            self.t["su"] += time.perf_counter() - t0
            return self
        if not self.lv2_icc:
            # Exact level-2 factor (MUMPS Cholesky, :894-906): dense LAPACK
            # Cholesky, or a sparse LU (SuperLU) of the block 7-point A_c
            # when a dense matrix would be too large.
            self.chol_c = self.lu_c_sp = None
            self.z_c = None
            if self.pin is None and self.wd is None and lay.nbox == 1:
                # Pure Neumann (the reference drivers' system, B6): A_c is
                # singular with null vector z = Phi^T M 1 (the coarse
                # coordinates of the constant).  Its pseudo-inverse on the
                # consistent residuals: (A_c + g z z^T)^{-1} (t - z z^T t),
                # g = trace / nc.  What MUMPS does with the singular
                # factorisation (:894-906) is parity unpinned.
                z = np.asarray(self.Phi.T @ self.mdiag).ravel()
                self.z_c = z / np.linalg.norm(z)
                if self.A_c is None:
                    raise ValueError("the unpinned coarse pseudo-inverse needs a dense A_c")
                g = np.trace(self.A_c) / nc
                self.chol_c = sla.cho_factor(self.A_c + g * np.outer(self.z_c, self.z_c),
                                             lower=True)
            elif self.A_c is None:
                self.lu_c_sp = spla.splu(self.A_c_sp.tocsc())
            else:
                self.chol_c = sla.cho_factor(self.A_c, lower=True)
            self.t["su_lv2_2"] += time.perf_counter() - ts
        # Level 3 eigen (:918-1175): smallest en_lv2 of the coarse matrix.
        ts = time.perf_counter()
        if lay.nbox > 1:
            self._lv3_boxes(nc)
            self.t["su_lv3_1"] += time.perf_counter() - ts
            ts = time.perf_counter()
            self.A_cc = self.energy_gram(sp.csr_matrix(self.Phi_cc)).toarray()
            if not self.no_shift_A_cc:
                self.A_cc = self.A_cc + NIL * np.eye(self.A_cc.shape[0])
            self.lu_cc = sla.lu_factor(self.A_cc)
            self.t["su_lv3_2"] += time.perf_counter() - ts
            self.t["su"] += time.perf_counter() - t0
            return self
        k2 = min(self.en_lv2, nc)
        if self.A_c is not None:
            lam2, V2 = sla.eigh(self.A_c, subset_by_index=[0, k2 - 1])
        else:
            lam2, V2 = spla.eigsh(self.A_c_sp.tocsc(), k=k2, sigma=-NIL, which="LM", tol=1e-12)
            o = np.argsort(lam2)
            lam2, V2 = lam2[o], V2[:, o]
        V2 = sign_normalize(V2)
        num2 = k2
        for j in range(k2):
            if lam2[j] >= self.eb_lv2:
                num2 = j + 1
                break
        self.eigen_num_lv2 = num2
        self.eigen_min_lv2, self.eigen_max_lv2 = lam2[0], lam2[num2 - 1]
        self.eig_lv2 = lam2
        self.V2 = V2[:, :num2]
        self.Phi_cc = np.asarray(self.Phi @ self.V2)
        self.t["su_lv3_1"] += time.perf_counter() - ts
        # Level 3 coarse-coarse (:1177-1390).
        ts = time.perf_counter()
        self.A_cc = self.energy_gram(sp.csr_matrix(self.Phi_cc)).toarray()
        if not self.no_shift_A_cc:
            self.A_cc = self.A_cc + NIL * np.eye(num2)
        self.lu_cc = sla.lu_factor(self.A_cc)
        self.t["su_lv3_2"] += time.perf_counter() - ts
        self.t["su"] += time.perf_counter() - t0
        return self

    def _lv3_boxes(self, nc):
        """Level 3 with process boxes (:918-1175): every box b takes the
        smallest en_lv2 eigenpairs of its box-local Neumann matrix A_i_inner
        (faces between boxes dropped), counted against eb_lv2 per box
        (:1100-1115); its Phi_cc vectors Phi_b v_j live on box b only.  The
        columns are ordered by box (x-fastest rank), as the ranks own the
        rows of A_cc (:1196-1199)."""
        A_in = self.energy_gram(self.Phi, inner=True).toarray()
        self.A_inner = 0.5 * (A_in + A_in.T)
        cols, self.eig_lv2_box, self.eigen_num_lv2_box = [], [], []
        for b in range(self.lay.nbox):
            d = np.flatnonzero(self.dof_box == b)
            k2 = min(self.en_lv2, d.size)
            lam, V = sla.eigh(self.A_inner[np.ix_(d, d)], subset_by_index=[0, k2 - 1])
            num = k2
            for j in range(k2):
                if lam[j] >= self.eb_lv2:
                    num = j + 1
                    break
            full = np.zeros((nc, num))
            full[d, :] = V[:, :num]
            cols.append(full)
            self.eig_lv2_box.append(lam)
            self.eigen_num_lv2_box.append(num)
        self.V2 = sign_normalize(np.hstack(cols))
        self.eigen_num_lv2 = self.V2.shape[1]
        self.eigen_min_lv2 = min(lam[0] for lam in self.eig_lv2_box)
        self.eigen_max_lv2 = max(lam[n - 1] for lam, n in zip(self.eig_lv2_box,
                                                              self.eigen_num_lv2_box))
        self.eig_lv2 = np.concatenate(self.eig_lv2_box)
        self.Phi_cc = np.asarray(self.Phi @ self.V2)

    # ---- apply ----------------------------------------------------------
    def apply_lv1(self, r):
        ts = time.perf_counter()
        y = self.dic.solve(r)
        self.t["av_lv1"] += time.perf_counter() - ts
        return y

    def apply_lv2(self, r):
        ts = time.perf_counter()
        c = self.Phi.T @ r
        if self.lv2_icc:
            cs = self.bicc.solve(c)
        elif self.use_2level:
            cs = sla.lu_solve(self.lu_c, c)
        elif self.chol_c is None:
            cs = self.lu_c_sp.solve(c)
        elif getattr(self, "z_c", None) is not None:
            cs = sla.cho_solve(self.chol_c, c - self.z_c * (self.z_c @ c))
        else:
            cs = sla.cho_solve(self.chol_c, c)
        y = self.Phi @ cs
        self.t["av_lv2"] += time.perf_counter() - ts
        return y

    def apply_lv3(self, r):
        ts = time.perf_counter()
        c = self.Phi_cc.T @ r
        cs = sla.lu_solve(self.lu_cc, c)
        y = self.Phi_cc @ cs
        self.t["av_lv3"] += time.perf_counter() - ts
        return y

    def apply(self, x):
        """PC_apply_vec (:1840-1936): y = 0, then residual-corrected stages."""
        t0 = time.perf_counter()
        x = np.asarray(x, dtype=np.float64).ravel()
        y = np.zeros_like(x)

        def stage(fn):
            nonlocal y
            r = x - self.matvec(y)
            y = y + fn(r)

        for _ in range(self.si_lv1):
            stage(self.apply_lv1)
        if self.use_2level:
            stage(self.apply_lv2)
        elif self.use_W_cycle:
            for _ in range(self.si_lv2):
                stage(self.apply_lv2)
            stage(self.apply_lv3)
            for _ in range(self.si_lv2):
                stage(self.apply_lv2)
            stage(self.apply_lv3)
            for _ in range(self.si_lv2):
                stage(self.apply_lv2)
        else:
            for _ in range(self.si_lv2):
                stage(self.apply_lv2)
            stage(self.apply_lv3)
            for _ in range(self.si_lv2):
                stage(self.apply_lv2)
        for _ in range(self.si_lv1):
            stage(self.apply_lv1)
        self.t["av"] += time.perf_counter() - t0
        return y


# --------------------------------------------------------------------------
# Outer solver: PETSc KSPCG, unpreconditioned norm
# --------------------------------------------------------------------------
def pcg(matvec, b, apply_pc, rtol=1e-8, abstol=1e-50, dtol=1e5, max_it=10000, x0=None):
    """PETSc ``KSPSolve_CG`` (standard variant) with ``KSP_NORM_UNPRECONDITIONED``
    and ``KSPConvergedDefault``: converged when ||r|| <= max(rtol*||r0||, abstol).

    Returns (x, its, history, reason)."""
    b = np.asarray(b, dtype=np.float64).ravel()
    if x0 is None:
        x = np.zeros_like(b)
        r = b.copy()
    else:
        x = np.asarray(x0, dtype=np.float64).ravel().copy()
        r = b - matvec(x)
    dp = np.linalg.norm(r)
    hist = [dp]
    rnorm0 = dp
    if x0 is not None:
        # KSPConvergedDefault with a nonzero initial guess measures against
        # ||b|| (||r0|| when b = 0).
        bn = np.linalg.norm(b)
        rnorm0 = bn if bn > 0.0 else dp
    ttol = max(rtol * rnorm0, abstol)
    if dp <= ttol:
        return x, 0, hist, "CONVERGED_RTOL" if dp > abstol else "CONVERGED_ATOL"
    z = apply_pc(r)
    beta = float(z @ r)
    p = None
    betaold = beta
    dpi = 0.0
    i = 0
    while True:
        its = i + 1
        if beta == 0.0:
            return x, its, hist, "CONVERGED_ATOL"
        if i == 0:
            p = z.copy()
        else:
            p = z + (beta / betaold) * p
        dpiold = dpi
        w = matvec(p)
        dpi = float(p @ w)
        betaold = beta
        if dpi == 0.0 or (i > 0 and np.sign(dpi) * np.sign(dpiold) < 0.0):
            return x, its, hist, "DIVERGED_INDEFINITE_MAT"
        a = beta / dpi
        x = x + a * p
        r = r - a * w
        dp = np.linalg.norm(r)
        hist.append(dp)
        if dp <= ttol:
            return x, its, hist, "CONVERGED_RTOL"
        if dp >= dtol * rnorm0:
            return x, its, hist, "DIVERGED_DTOL"
        if its >= max_it:
            return x, its, hist, "DIVERGED_ITS"
        z = apply_pc(r)
        beta = float(z @ r)
        i += 1


def face_fluxes(lay: Layout, T, p):
    """RT0 face fluxes F = T (p_L - p_R) (velocity recovery from the pressure)."""
    M, N, P = lay.mesh
    p = p.reshape(P, N, M)
    Tx, Ty, Tz = T
    return (Tx * (p[:, :, :-1] - p[:, :, 1:]),
            Ty * (p[:, :-1, :] - p[:, 1:, :]),
            Tz * (p[:-1, :, :] - p[1:, :, :]))


def solve_problem(prob, ghep_fn=None):
    """Set up the oracle for a ``premixfem_amd.problems.Problem`` and solve."""
    orc = Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                 en_lv2=prob.en_lv2, eg_op=prob.eg_op, pin_cell=prob.pin_cell)
    orc.setup(ghep_fn)
    x, its, hist, reason = pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    return orc, x, its, hist, reason


def create_A_diagonals(lay: Layout, T, pin_cell=None):
    """PC_create_A's matrix as seven diagonals, natural order: diag[c] and
    off[c, d] for the -x,+x,-y,+y,-z,+z neighbours.  The diagonal follows the
    accumulation order of the reference's face loop (:1777-1816): a cell's
    lower x, y, z faces (its own iteration), then its upper x, y, z faces
    (the iterations of the upper neighbours); the pin term last."""
    M, N, P = lay.mesh
    xl, xh, yl, yh, zl, zh = cell_faces(lay, T)
    d = xl.copy()
    d = d + yl
    d = d + zl
    d = d + xh
    d = d + yh
    d = d + zh
    d = d.ravel()
    if pin_cell is not None:
        x, y, z = pin_cell
        d[(z * N + y) * M + x] += pin_tau(lay, T, pin_cell)
    off = -np.stack([a.ravel() for a in (xl, xh, yl, yh, zl, zh)], axis=1)
    return d, off


def basis_columns(orc: "Oracle"):
    """The level-2 basis as en_lv1 natural-order vectors (vector j holds
    eigenvector j of every subdomain, zero where a subdomain has fewer), the
    per-subdomain counts and the eigenvalues [nsub][en_lv1]."""
    n = orc.A.shape[0]
    k = orc.en_lv1
    phi = np.zeros((k, n))
    ev = np.zeros((orc.lay.nsub, k))
    for s in range(orc.lay.nsub):
        V = orc.phi[s]
        phi[:V.shape[1], orc.sub_cells[s]] = V.T
        ev[s, :len(orc.eig_lv1[s])] = orc.eig_lv1[s][:k]
    return phi, orc.eigen_num_lv1.astype(np.int32), ev
