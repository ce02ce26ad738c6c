"""Pins of the oracle's process-box semantics (CPU only).

With P process boxes (the DMDA decomposition of src/PreMixFEM_3D.c:1711)
levels 2 and 3 change: A_c keeps only the couplings inside a box
(_PC_setup_lv2_J, :806-888), each box takes the smallest en_lv2 eigenpairs
of its box-local Neumann coarse matrix A_i_inner (:918-1175), and A_cc
couples the box vectors through the global operator (:1177-1390).  Each
check below is an independent construction (dense products with the
operator assembled on the box alone, brute-force eigensolves), chosen so a
kept cross-box block, a box-boundary face left in A_i_inner or a vector
spread over two boxes fails.
"""
import numpy as np
import pytest

from oracle import premixfem_oracle as O
from premixfem_amd import problems

MESH = (13, 10, 9)
PROCS = (2, 1, 2)


def _orc(procs=PROCS, **kw):
    k = problems.grf_kappa(MESH, 1.5, 0.2, 17)
    args = dict(sd=2, en_lv1=3, en_lv2=3, pin_cell=(MESH[0] - 1, 0, 0))
    args.update(kw)
    return O.Oracle((1.0, 1.0, 1.0), MESH, k, procs=procs, **args).setup()


@pytest.fixture(scope="module")
def orc():
    return _orc()


def test_box_partition_kat():
    lay = O.Layout((1, 1, 1), MESH, 2, PROCS)
    # x: boxes of 7 and 6 cells (13/2 + (13%2 > j)), each cut in 2 by
    # len_i = n/sd + (n%sd > i).
    assert lay.parts[0] == ([0, 4, 7, 10], [4, 3, 3, 3])
    assert lay.parts[1] == ([0, 5], [5, 5])
    assert lay.parts[2] == ([0, 3, 5, 7], [3, 2, 2, 2])
    assert lay.sd == (4, 2, 4) and lay.nbox == 4
    assert [lay.sub_box(s) for s in (0, 1, 2, 3, 8, 16, 31)] == [0, 0, 1, 1, 0, 2, 3]


def _box_cells(lay, b):
    px, py, _ = lay.procs
    bx, by, bz = b % px, (b // px) % py, b // (px * py)
    (sx, lx), (sy, ly), (sz, lz) = lay.box_parts
    return (sx[bx], lx[bx]), (sy[by], ly[by]), (sz[bz], lz[bz])


def _box_index(orc, b):
    M, N, P = orc.lay.mesh
    (x0, nx), (y0, ny), (z0, nz) = _box_cells(orc.lay, b)
    idx = np.arange(M * N * P).reshape(P, N, M)
    return idx[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx].ravel(), (x0, nx, y0, ny, z0, nz)


def _box_neumann(orc, b):
    """The operator of box b alone (no-flow on all its faces), assembled by
    assemble_A on the box's own layout, plus the pin if it lies there."""
    cells, (x0, nx, y0, ny, z0, nz) = _box_index(orc, b)
    Tx, Ty, Tz = orc.T
    Tb = (Tx[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx - 1],
          Ty[z0:z0 + nz, y0:y0 + ny - 1, x0:x0 + nx],
          Tz[z0:z0 + nz - 1, y0:y0 + ny, x0:x0 + nx])
    lay_b = O.Layout((nx, ny, nz), (nx, ny, nz), 2)
    A = O.assemble_A(lay_b, Tb).toarray()
    (px, py, pz), tau = orc.pin
    if x0 <= px < x0 + nx and y0 <= py < y0 + ny and z0 <= pz < z0 + nz:
        i = ((pz - z0) * ny + (py - y0)) * nx + (px - x0)
        A[i, i] += tau
    return cells, A


def test_level2_keeps_only_box_couplings(orc):
    Phi = orc.Phi.toarray()
    A = orc.A.toarray()
    db = orc.dof_box
    for b in range(orc.lay.nbox):
        cells, _ = _box_index(orc, b)
        d = np.flatnonzero(db == b)
        ref = Phi[np.ix_(cells, d)].T @ A[np.ix_(cells, cells)] @ Phi[np.ix_(cells, d)]
        assert np.allclose(orc.A_c[np.ix_(d, d)], ref, rtol=1e-12, atol=1e-14)
        o = np.flatnonzero(db != b)
        assert np.all(orc.A_c[np.ix_(d, o)] == 0.0)
    # Without boxes the same subdomains couple across the box faces.
    full = Phi.T @ A @ Phi
    d0, d1 = np.flatnonzero(db == 0), np.flatnonzero(db == 1)
    assert np.abs(full[np.ix_(d0, d1)]).max() > 1e-3 * np.abs(full).max()
    # Every box block is nonsingular through its faces to the other boxes.
    assert np.linalg.eigvalsh(orc.A_c).min() > 1e-8 * np.abs(orc.A_c).max()


def test_level3_matrix_is_the_box_neumann_galerkin(orc):
    Phi = orc.Phi.toarray()
    for b in range(orc.lay.nbox):
        cells, AN = _box_neumann(orc, b)
        d = np.flatnonzero(orc.dof_box == b)
        Pb = Phi[np.ix_(cells, d)]
        assert np.allclose(orc.A_inner[np.ix_(d, d)], Pb.T @ AN @ Pb, rtol=1e-12,
                           atol=1e-13 * np.abs(AN).max())
        assert np.all(orc.A_inner[np.ix_(d, np.flatnonzero(orc.dof_box != b))] == 0.0)


def test_level3_box_eigenpairs(orc):
    V = orc.V2
    col = 0
    assert len(orc.eigen_num_lv2_box) == orc.lay.nbox
    (px, py, pz), _ = orc.pin
    for b, (lam, num) in enumerate(zip(orc.eig_lv2_box, orc.eigen_num_lv2_box)):
        d = np.flatnonzero(orc.dof_box == b)
        Ab = orc.A_inner[np.ix_(d, d)]
        brute = np.linalg.eigvalsh(Ab)
        assert np.allclose(lam, brute[:len(lam)], rtol=1e-9, atol=1e-12 * brute[-1])
        Vb = V[:, col:col + num]
        # Supported on box b only, orthonormal, eigenvectors of its block.
        assert np.all(Vb[np.flatnonzero(orc.dof_box != b)] == 0.0)
        assert np.allclose(Vb.T @ Vb, np.eye(num), atol=1e-12)
        assert np.allclose(Ab @ Vb[d], Vb[d] * lam[:num], atol=1e-10 * brute[-1])
        # Boxes without the pin are pure Neumann: lambda_0 = 0 with the box
        # constant as eigenvector.
        cells, (x0, nx, y0, ny, z0, nz) = _box_index(orc, b)
        if not (x0 <= px < x0 + nx and y0 <= py < y0 + ny and z0 <= pz < z0 + nz):
            assert abs(lam[0]) <= 1e-12 * brute[-1]
            u = orc.Phi.toarray()[cells][:, d] @ Vb[d, 0]
            assert np.allclose(u, u[0], rtol=1e-8)
        col += num
    assert col == orc.eigen_num_lv2 == V.shape[1]


def test_level3_coarse_coarse_matrix(orc):
    Phi_cc = orc.Phi.toarray() @ orc.V2
    assert np.allclose(orc.Phi_cc, Phi_cc, atol=1e-14)
    A = orc.A.toarray()
    ref = Phi_cc.T @ A @ Phi_cc + O.NIL * np.eye(orc.eigen_num_lv2)
    assert np.allclose(orc.A_cc, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # Boxes that share a face couple in A_cc.
    n0 = orc.eigen_num_lv2_box[0]
    assert np.abs(orc.A_cc[:n0, n0:2 * n0]).max() > 0.0


def test_adaptive_count_per_box():
    """eb_lv2 is tested per box (:1100-1115): an unpinned box stops after its
    zero eigenvalue (count 2), the pinned box (x = M-1 side, box 1) at its
    first eigenvalue (~3e-5 >= 1e-9, count 1)."""
    o = _orc(eb_lv2=1e-9)
    assert o.eigen_num_lv2_box == [2, 1, 2, 2]
    assert o.eigen_num_lv2 == 7 and o.V2.shape[1] == 7


def test_box_preconditioner_symmetric_and_converges(orc):
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal((2, orc.A.shape[0]))
    a, b = orc.apply(u) @ v, u @ orc.apply(v)
    assert abs(a - b) <= 1e-10 * abs(a)
    assert orc.apply(u) @ u > 0
    rhs = problems.corner_rhs(MESH, (1, 1, 1)).ravel()
    x, its, hist, reason = O.pcg(orc.matvec, rhs, orc.apply, rtol=1e-10)
    assert reason == "CONVERGED_RTOL"
    assert np.linalg.norm(rhs - orc.matvec(x)) <= 1e-9 * np.linalg.norm(rhs)


def test_one_box_is_the_one_process_path():
    a = _orc(procs=(1, 1, 1))
    b = O.Oracle((1.0, 1.0, 1.0), MESH, problems.grf_kappa(MESH, 1.5, 0.2, 17), sd=2, en_lv1=3,
                 en_lv2=3, pin_cell=(MESH[0] - 1, 0, 0)).setup()
    assert np.array_equal(a.A_c, b.A_c)
    assert np.array_equal(a.V2, b.V2)


def test_two_level_keeps_every_coupling():
    """The 2-level variant on P processes (_PC_setup_lv2_c, :423-735) builds
    one coarse matrix over all processes, the neighbours' dof offsets
    exchanged (:448-479): A_c = Phi^T A Phi with the couplings between
    boxes, plus the NIL shift (:690-692), against a dense product."""
    o = _orc(use_2level=True)
    Phi = o.Phi.toarray()
    full = Phi.T @ o.A.toarray() @ Phi
    assert np.allclose(o.A_c, full, rtol=1e-12, atol=1e-14 * np.abs(full).max())
    db = o.dof_box
    d0, d1 = np.flatnonzero(db == 0), np.flatnonzero(db == 1)
    assert np.abs(o.A_c[np.ix_(d0, d1)]).max() > 1e-3 * np.abs(full).max()
    # the apply is the exact coarse solve between two level-1 stages
    rng = np.random.default_rng(5)
    r = rng.standard_normal(o.A.shape[0])
    y1 = o.apply_lv1(r)
    t = Phi.T @ (r - o.matvec(y1))
    y2 = y1 + Phi @ np.linalg.solve(full + O.NIL * np.eye(full.shape[0]), t)
    y = y2 + o.apply_lv1(r - o.matvec(y2))
    assert np.allclose(o.apply(r), y, rtol=1e-9, atol=1e-12 * np.abs(y).max())


def test_block_icc_on_boxes_is_the_boxes_own_ic0():
    """The spe10 fork's level 2 on P processes is one SBAIJ matrix per process
    (spe10.c:1074-1205) with PETSc ICC(0): per box, the textbook scalar
    IC(0) of that box's block of A_c on its block pattern (from the box
    alone), and nothing between boxes."""
    import math

    o = _orc(lv2_icc=True)
    A = o.A_c_sp.toarray()
    d = o.dof_idx
    lay = o.lay
    rng = np.random.default_rng(7)
    r = rng.standard_normal(A.shape[0])
    y = o.bicc.solve(r)
    for b in range(lay.nbox):
        subs = [s for s in range(lay.nsub) if lay.sub_box(s) == b]
        idx = np.concatenate([np.arange(d[s], d[s + 1]) for s in subs])
        pos = {s: i for i, s in enumerate(subs)}
        # block pattern: self and face neighbours inside the box
        sdx, sdy, _ = lay.sd
        pat = np.zeros((idx.size, idx.size), dtype=bool)
        off = np.concatenate([[0], np.cumsum([d[s + 1] - d[s] for s in subs])])
        for s in subs:
            for t in (s, s - 1, s + 1, s - sdx, s + sdx, s - sdx * sdy, s + sdx * sdy):
                if t in pos and (abs(t - s) != 1 or t // sdx == s // sdx) and \
                        (abs(t - s) != sdx or t // (sdx * sdy) == s // (sdx * sdy)):
                    i, j = pos[s], pos[t]
                    pat[off[i]:off[i + 1], off[j]:off[j + 1]] = True
        a = A[np.ix_(idx, idx)]
        n = idx.size
        L = np.zeros((n, n))
        for i in range(n):
            for j in range(i + 1):
                if pat[i, j]:
                    s = a[i, j] - L[i, :j] @ L[j, :j]
                    L[i, j] = math.sqrt(s) if i == j else s / L[j, j]
        np.testing.assert_allclose(L @ (L.T @ y[idx]), r[idx], rtol=1e-9,
                                   atol=1e-9 * np.abs(r).max())
    # the preconditioner symmetric and positive on the boxes
    u, v = rng.standard_normal((2, o.A.shape[0]))
    assert abs(u @ o.apply(v) - v @ o.apply(u)) <= 1e-10 * abs(u @ o.apply(v))
    assert u @ o.apply(u) > 0
