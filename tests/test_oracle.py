"""Pins of the FP64 CPU oracle (CPU only).

The reference holds no tests, golden vectors or fixtures, so each oracle
function is checked against what the mathematics fixes, chosen so that a
dropped term, a wrong sign or index, or a transposed operand fails:
closed forms, exact invariants, textbook reductions and brute force
(dense linear algebra) on small grids.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import premixfem_oracle as O
from premixfem_amd import problems


# ---------------------------------------------------------------- partition
def test_partition_kat():
    # n = 10, sd = 4 -> lens 3,3,2,2 (remainder to the first ones), starts by prefix.
    assert O.partition(10, 4) == ([0, 3, 6, 8], [3, 3, 2, 2])
    assert O.partition(16, 2) == ([0, 8], [8, 8])
    assert O.partition(7, 3) == ([0, 3, 5], [3, 2, 2])
    # sd = 1 is reset to 2 (src/PreMixFEM_3D.c:1718-1721)
    assert O.Layout((1, 1, 1), (8, 8, 8), 1).sd == (2, 2, 2)


def test_subdomain_order_x_fastest():
    lay = O.Layout((1, 1, 1), (8, 6, 4), (2, 3, 2))
    sub = lay.cell_sub()
    assert sub[0, 0, 0] == 0 and sub[0, 0, 7] == 1 and sub[0, 5, 0] == 4 and sub[3, 0, 0] == 6


# ---------------------------------------------------------------- operator
def test_transmissibility_closed_form():
    mesh = (4, 2, 2)
    lay = O.Layout((1.0, 0.5, 0.5), mesh, 2)
    k = np.ones((3, 2, 2, 4))
    k[0, :, :, 1] = 3.0
    Tx, Ty, Tz = O.transmissibilities(lay, k)
    H = 0.25
    # T = H * harmonic mean for a cubic cell of side H
    assert np.isclose(Tx[0, 0, 0], H * 2.0 / (1.0 + 1.0 / 3.0))
    assert np.isclose(Tx[0, 0, 2], H * 1.0)
    assert np.isclose(Ty[0, 0, 1], H * 1.0)


def test_operator_textbook_laplacian():
    # constant kappa: A = H * (7-point Neumann Laplacian), built independently by kron
    n = 5
    lay = O.Layout((1, 1, 1), (n, n, n), 2)
    T = O.transmissibilities(lay, np.ones((3, n, n, n)))
    A = O.assemble_A(lay, T).toarray()
    d = sp.diags([-np.ones(n - 1), np.r_[1, 2 * np.ones(n - 2), 1], -np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity(n)
    L = sp.kron(I, sp.kron(I, d)) + sp.kron(I, sp.kron(d, I)) + sp.kron(d, sp.kron(I, I))
    assert np.allclose(A, L.toarray() / n)


def test_operator_invariants():
    p = problems.config("tiny")
    lay = O.Layout(p.dom, p.mesh, p.sd)
    T = O.transmissibilities(lay, p.kappa)
    A = O.assemble_A(lay, T)
    assert abs(A - A.T).max() == 0.0
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-12 * abs(A).max()   # pure Neumann
    Ap = O.assemble_A(lay, T, p.pin_cell).toarray()
    assert np.linalg.eigvalsh(Ap).min() > 0.0                               # pinned: SPD
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    assert np.allclose(O.stencil_apply(lay, T, x, (p.pin_cell, O.pin_tau(lay, T, p.pin_cell))),
                       Ap @ x)
    d, off = O.create_A_diagonals(lay, T, p.pin_cell)
    assert np.allclose(d, np.diag(Ap))


# ---------------------------------------------------------------- level 1
def _dense_blocks(p):
    lay = O.Layout(p.dom, p.mesh, p.sd)
    T = O.transmissibilities(lay, p.kappa)
    pin = (p.pin_cell, O.pin_tau(lay, T, p.pin_cell))
    A = O.assemble_A(lay, T, p.pin_cell).toarray()
    sub = lay.cell_sub().ravel()
    return lay, T, pin, A, sub


def test_dic_is_icc0():
    """ICC(0) defining property: L D L^T equals A on the pattern of each
    principal block, and the solve inverts exactly that product."""
    p = problems.config("tiny")
    lay, T, pin, A, sub = _dense_blocks(p)
    dic = O.DIC(lay, T, O.lv1_block_diag(lay, T, pin))
    assert np.allclose(O.lv1_block_diag(lay, T, pin), np.diag(A))
    r = np.random.default_rng(1).standard_normal(A.shape[0])
    y = dic.solve(r)
    for s in range(lay.nsub):
        idx = np.where(sub == s)[0]
        Ai = A[np.ix_(idx, idx)]
        D = np.diag(dic.d[idx])
        Lst = np.tril(Ai, -1)
        Mi = (D + Lst) @ np.linalg.inv(D) @ (D + Lst).T
        pattern = Ai != 0
        assert np.allclose(Mi[pattern], Ai[pattern], rtol=1e-12, atol=0)
        assert np.allclose(y[idx], np.linalg.solve(Mi, r[idx]), rtol=1e-10)


def test_dic_exact_on_lines():
    # M x 1 x 1 blocks: ICC(0) is the exact (Thomas) factorisation
    mesh = (12, 2, 2)
    k = problems.grf_kappa(mesh, 1.0, 0.3, 5)
    lay = O.Layout((1, 1, 1), mesh, (2, 2, 2))
    T = O.transmissibilities(lay, k)
    A = O.assemble_A(lay, T).toarray()
    dic = O.DIC(lay, T, O.lv1_block_diag(lay, T))
    r = np.random.default_rng(2).standard_normal(A.shape[0])
    y = dic.solve(r)
    sub = lay.cell_sub().ravel()
    for s in range(lay.nsub):
        idx = np.where(sub == s)[0]
        assert np.allclose(y[idx], np.linalg.solve(A[np.ix_(idx, idx)], r[idx]), rtol=1e-12)


# ---------------------------------------------------------------- level 2
def test_mod_mass_x_quirk_kat():
    """Constant kappa, 8 cells in x in 2 subdomains: the x weights per cell are
    [2, 2, 2, 1 | 3, 2, 2, 0] * w (upper face once, again when the upper
    neighbour is in the subdomain, lower face for the first cell only; no
    domain-boundary faces); y and z count both existing faces."""
    mesh = (8, 3, 3)
    lay = O.Layout((1.0, 3 / 8, 3 / 8), mesh, (2, 2, 2))
    k = np.ones((3, 3, 3, 8))
    T = O.transmissibilities(lay, k)
    m = O.mass_diag(lay, k, T).reshape(3, 3, 8)
    H = 1 / 8
    w = 2.0 / H / H * H ** 3
    xw = np.array([2, 2, 2, 1, 3, 2, 2, 0], dtype=float)
    # interior in y and z (ey = ez = 1): 2 faces in y, 2 in z
    assert np.allclose(m[1, 1, :], (xw + 4) * w)
    # corner row (ey = ez = 0): one face in y, one in z
    assert np.allclose(m[0, 0, :], (xw + 2) * w)


def test_mean_and_diag_mass():
    p = problems.config("tiny")
    lay, T, pin, A, sub = _dense_blocks(p)
    assert np.allclose(O.mass_diag(lay, p.kappa, T, O.EIG_OP_MEAN), p.kappa.sum(0).ravel())
    md = O.mass_diag(lay, p.kappa, T, O.EIG_OP_DIAG)
    for s in range(lay.nsub):
        idx = np.where(sub == s)[0]
        assert np.allclose(md[idx], O.local_neumann(lay, T, s).diagonal())


def test_ghep_invariants_and_bruteforce():
    p = problems.config("small")
    lay, T, pin, A, sub = _dense_blocks(p)
    m_all = O.mass_diag(lay, p.kappa, T)
    for s in (0, 5):
        AN = O.local_neumann(lay, T, s)
        # Neumann block: interior faces only -> A^N 1 = 0
        assert np.abs(AN @ np.ones(AN.shape[0])).max() < 1e-12 * abs(AN).max()
        idx = np.where(sub == s)[0]
        m = m_all[idx]
        lam, V = O.ghep(AN, m, 5, dense_max=0)   # shift-invert Lanczos
        lam_d, V_d = O.ghep(AN, m, 5)            # dense brute force
        assert np.allclose(lam[1:], lam_d[1:], rtol=1e-9)
        assert abs(lam[0]) < 1e-12
        # M-orthonormal, constant first vector, same spans
        assert np.allclose(V.T @ (m[:, None] * V), np.eye(5), atol=1e-9)
        assert np.allclose(V[:, 0], 1 / np.sqrt(m.sum()), rtol=1e-8)
        ang = sla.subspace_angles(np.sqrt(m)[:, None] * V, np.sqrt(m)[:, None] * V_d)
        assert ang.max() < 1e-7
        # residual of the generalized problem
        R = AN @ V - (m[:, None] * V) * lam
        assert np.abs(R).max() < 1e-9 * abs(AN).max()


def test_galerkin_and_level3():
    p = problems.config("tiny")
    orc = O.Oracle(p.dom, p.mesh, p.kappa, sd=p.sd, en_lv1=p.en_lv1, en_lv2=p.en_lv2,
                   pin_cell=p.pin_cell).setup()
    Phi = orc.Phi.toarray()
    A = orc.A.toarray()
    assert np.allclose(orc.A_c, Phi.T @ A @ Phi, rtol=1e-12, atol=1e-15)
    assert np.array_equal(orc.A_c, orc.A_c.T)
    # block-7-point structure: blocks of non-adjacent subdomains vanish
    k = p.en_lv1
    assert np.all(orc.A_c[0:k, 7 * k:8 * k] == 0.0)
    # level 3: Phi_cc = Phi V; A_cc = V^T A_c V + NIL I
    assert np.allclose(orc.Phi_cc, Phi @ orc.V2)
    assert np.allclose(orc.A_cc, orc.V2.T @ orc.A_c @ orc.V2 + O.NIL * np.eye(orc.eigen_num_lv2))
    lam_all = np.linalg.eigvalsh(orc.A_c)
    assert np.allclose(orc.eig_lv2, lam_all[:p.en_lv2], rtol=1e-10)


def test_level2_projection_identity():
    """After a level-2 stage y += Phi A_c^-1 Phi^T r, the new residual is
    Phi-orthogonal (exact Galerkin solve on one process)."""
    p = problems.config("tiny")
    orc = O.Oracle(p.dom, p.mesh, p.kappa, sd=p.sd, en_lv1=p.en_lv1, en_lv2=p.en_lv2,
                   pin_cell=p.pin_cell).setup()
    x = np.random.default_rng(4).standard_normal(orc.A.shape[0])
    y = orc.apply_lv2(x)
    r = x - orc.matvec(y)
    assert np.abs(orc.Phi.T @ r).max() < 1e-9 * np.abs(orc.Phi.T @ x).max()


def test_vcycle_symmetric():
    p = problems.config("tiny")
    orc = O.Oracle(p.dom, p.mesh, p.kappa, sd=p.sd, en_lv1=p.en_lv1, en_lv2=p.en_lv2,
                   pin_cell=p.pin_cell).setup()
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, orc.A.shape[0]))
    a, b = orc.apply(u) @ v, u @ orc.apply(v)
    assert abs(a - b) <= 1e-10 * abs(a)
    assert orc.apply(u) @ u > 0


def test_complete_coarse_space_one_iteration():
    """With en_lv1 = cells per subdomain the coarse space is everything, the
    level-2 solve is exact and PCG converges in one iteration."""
    mesh = (4, 4, 4)
    k = problems.grf_kappa(mesh, 1.0, 0.3, 9)
    b = problems.corner_rhs(mesh, (1, 1, 1))
    orc = O.Oracle((1, 1, 1), mesh, k, sd=2, en_lv1=8, en_lv2=2, pin_cell=(3, 0, 0)).setup()
    x, its, hist, reason = O.pcg(orc.matvec, b.ravel(), orc.apply, rtol=1e-8)
    assert its == 1 and reason == "CONVERGED_RTOL"


def test_pcg_matches_scipy_cg():
    """The PETSc-semantics PCG (unpreconditioned norm, ||r|| <= rtol ||b||)
    against SciPy's independent CG with the same preconditioner."""
    p = problems.config("tiny")
    orc = O.Oracle(p.dom, p.mesh, p.kappa, sd=p.sd, en_lv1=p.en_lv1, en_lv2=p.en_lv2,
                   pin_cell=p.pin_cell).setup()
    b = p.rhs.ravel()
    x, its, hist, reason = O.pcg(orc.matvec, b, orc.apply, rtol=1e-10)
    n = b.size
    Aop = spla.LinearOperator((n, n), matvec=orc.matvec)
    Mop = spla.LinearOperator((n, n), matvec=orc.apply)
    count = [0]
    x2, info = spla.cg(Aop, b, rtol=1e-10, atol=0.0, M=Mop,
                       callback=lambda xk: count.__setitem__(0, count[0] + 1))
    assert info == 0
    assert abs(its - count[0]) <= 1
    assert np.linalg.norm(x - x2) <= 1e-8 * np.linalg.norm(x)
    assert hist[-1] <= 1e-10 * hist[0]


def test_reference_driver_problem_runs():
    """The Expe_scal configuration (cross channels, XxY normalised RHS, pure
    Neumann, 2-level variant with the NIL-shifted coarse LU)."""
    mesh = (16, 16, 16)
    k = problems.cross_kappa(mesh, 2)
    b = problems.xxy_rhs(mesh, (1, 1, 1))
    assert abs(b.sum()) < 1e-12 and np.isclose(np.linalg.norm(b), 1.0)
    orc = O.Oracle((1, 1, 1), mesh, k, sd=2, en_lv1=4, en_lv2=4, use_2level=True).setup()
    x, its, hist, reason = O.pcg(orc.matvec, b.ravel(), orc.apply, rtol=1e-8)
    assert reason == "CONVERGED_RTOL" and its < 60


def test_bad_arguments():
    with pytest.raises(ValueError):
        O.Oracle((1, -1, 1), (4, 4, 4), np.ones((3, 4, 4, 4)))
    with pytest.raises(ValueError):
        O.Oracle((1, 1, 1), (4, 4, 4), np.ones((3, 4, 4, 4)), en_lv1=0)
    with pytest.raises(ValueError):
        O.Oracle((1, 1, 1), (4, 4, 4), np.ones((3, 4, 4, 4)), eg_op=7)


def test_exact_level1_blocks_are_the_principal_submatrices():
    """-use_full_Cholesky_lv1 (src/PreMixFEM_3D.c:205-207): the oracle's
    exact level-1 matrix is A with every coupling between different
    subdomains removed (diagonal kept), and its solve is exact."""
    prob = problems.config("small")
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell, use_full_Cholesky_lv1=True)
    orc.setup()
    B = orc.dic.B.toarray()
    A = orc.A.toarray()
    lay = orc.lay
    sub = lay.cell_sub().ravel()
    same = sub[:, None] == sub[None, :]
    off = same & ~np.eye(B.shape[0], dtype=bool)
    np.testing.assert_array_equal(B[off], A[off])
    # the diagonal in the reference's accumulation order (lv1_block_diag)
    np.testing.assert_allclose(np.diag(B), np.diag(A), rtol=1e-14)
    assert not B[~same].any()
    r = np.random.default_rng(0).standard_normal(B.shape[0])
    y = orc.dic.solve(r)
    assert np.linalg.norm(B @ y - r) <= 1e-12 * np.linalg.norm(r)
    # exact blocks precondition at least as well as ICC(0)
    ref = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell).setup()
    b = prob.rhs.ravel()
    _, its_exact, _, reason = O.pcg(orc.matvec, b, orc.apply, rtol=1e-8)
    _, its_icc, _, _ = O.pcg(ref.matvec, b, ref.apply, rtol=1e-8)
    assert reason == "CONVERGED_RTOL" and its_exact <= its_icc


def test_unpinned_coarse_solve_is_the_pseudo_inverse():
    """The reference drivers' pure-Neumann system (B6): A_c is singular with
    the coarse constant as null vector; the oracle's level-2 solve is its
    pseudo-inverse on consistent vectors (checked against numpy's pinv)."""
    mesh = (12, 12, 12)
    k = problems.cross_kappa(mesh, 2)
    orc = O.Oracle((1, 1, 1), mesh, k, sd=2, en_lv1=4, en_lv2=4).setup()
    A_c = orc.A_c
    z = orc.z_c
    assert np.linalg.norm(A_c @ z) <= 1e-10 * np.linalg.norm(A_c)
    t = np.random.default_rng(1).standard_normal(A_c.shape[0])
    t -= z * (z @ t)
    want = np.linalg.pinv(A_c, rcond=1e-12, hermitian=True) @ t
    got = sla.cho_solve(orc.chol_c, t - z * (z @ t))
    assert np.linalg.norm(got - want) <= 1e-8 * np.linalg.norm(want)
    assert abs(z @ got) <= 1e-10 * np.linalg.norm(got)
