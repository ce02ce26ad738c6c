"""CPU checks of the oracle for the spe10 fork: the block ICC(0) level-2
factor, the well term, the pressure system, the upwind transport and the
PETSc initial-guess convergence rule (src/spe10.c)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import premixfem_oracle as O
from oracle import spe10_oracle as S
from premixfem_amd import problems as PB


def ic0_scalar(A, pat=None):
    """Textbook scalar IC(0) (no fill) on a sparsity pattern of SPD A (the
    stored entries of A if pat is None): for j <= i in the pattern,
    L_ij = (a_ij - sum_{k<j} L_ik L_jk) / L_jj, L_ii = sqrt(a_ii - sum L_ik^2).
    Returns dense L."""
    a = A.toarray() if sp.issparse(A) else np.asarray(A)
    n = a.shape[0]
    if pat is None:
        pat = (a != 0) | np.eye(n, dtype=bool)
    L = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            if not pat[i, j]:
                continue
            s = a[i, j] - L[i, :j] @ L[j, :j]
            L[i, j] = math.sqrt(s) if i == j else s / L[j, j]
    return L


def small_case(mesh=(8, 10, 6), sd=(2, 2, 2), seed=3001, S_w=None, k=4):
    prm = S.Spe10Params()
    lay = S.layout(mesh, sd, prm.H)
    kap, poro = PB.reservoir_fields(mesh, seed)
    M, N, P = mesh
    Sw = np.full((P, N, M), prm.S_wi) if S_w is None else S_w
    kappa, T, wd, rhs = S.linear_system(prm, lay, list(kap), Sw)
    return prm, lay, list(kap), S.modified_poro(poro), Sw, kappa, T, wd, rhs


def test_block_icc_equals_scalar_ic0_on_pattern():
    """Block ICC(0) of the oracle == scalar IC(0) on A_c's block pattern
    (explicit zeros kept, as MatSetValues inserts every block entry)."""
    prm, lay, perm, poro, Sw, kappa, T, wd, rhs = small_case()
    orc = O.Oracle(lay.dom, lay.mesh, kappa, sd=lay.sd, en_lv1=4, en_lv2=2,
                   eg_op=O.EIG_OP_DIAG, wd=wd, lv2_icc=True, no_shift_A_cc=True).setup()
    A = orc.A_c_sp.toarray()
    n = A.shape[0]
    d = orc.dof_idx
    # Pattern: dense blocks between each subdomain and itself / face neighbours.
    pat = np.zeros((n, n), dtype=bool)
    sdx, sdy, sdz = lay.sd
    for s in range(lay.nsub):
        for t in [s] + orc.bicc.nbr_lo[s] + orc.bicc.nbr_hi[s]:
            pat[d[s]:d[s + 1], d[t]:d[t + 1]] = True
    L = ic0_scalar(np.where(pat, A, 0.0), pat)
    Mref = L @ L.T
    rng = np.random.default_rng(0)
    r = rng.standard_normal(n)
    y = orc.bicc.solve(r)
    np.testing.assert_allclose(Mref @ y, r, rtol=1e-10, atol=1e-10 * np.abs(r).max())
    # and the factor is not exact: some fill was dropped
    assert np.abs(Mref - A).max() > 1e-8 * np.abs(A).max()


def test_scalar_ic0_reference_is_exact_on_tridiagonal():
    """The scalar IC(0) used above is the exact Cholesky for a tridiagonal
    SPD matrix (no fill exists): pins the checker itself."""
    n = 12
    A = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    L = ic0_scalar(A)
    np.testing.assert_allclose(L @ L.T, A.toarray(), rtol=1e-13, atol=1e-13)


def test_well_index_and_mobility_known_values():
    # Peaceman-type WI for 20 x 10 ft cells, rw = 1 ft (src/spe10.c:399-405),
    # evaluated step by step.
    r1 = math.sqrt(9 * 400 + 100) / 2
    r2 = math.sqrt(400 + 9 * 100) / 2
    t = (10 / 20 * math.log(r1) + 20 / 10 * math.log(r2)) * 2 / math.pi - 1
    assert S.well_index(20.0, 10.0, 1.0) == pytest.approx((1 / 400 + 1 / 100) / t, rel=1e-15)
    prm = S.Spe10Params()
    lw, lo, lt = S.mobility(prm, np.array([prm.S_wc, 1.0 - prm.S_or, 0.5]))
    assert lw[0] == 0.0 and lo[1] == 0.0
    assert lo[0] == pytest.approx(1 / prm.mu_o, rel=1e-15)
    assert lw[1] == pytest.approx(1 / prm.mu_w, rel=1e-15)
    # S* = 0.5: krw = kro = 1/4
    assert lt[2] == pytest.approx(0.25 / prm.mu_w + 0.25 / prm.mu_o, rel=1e-15)


def test_pressure_system_row_sums_and_rhs():
    prm, lay, perm, poro, Sw, kappa, T, wd, rhs = small_case()
    M, N, P = lay.mesh
    A = O.assemble_A(lay, T) + sp.diags(wd)
    one = np.ones(M * N * P)
    np.testing.assert_allclose(A @ one, wd, atol=1e-12 * np.abs(A).max())
    prod = S.producer_mask(lay.mesh).ravel()
    inj = S.injector_mask(lay.mesh).ravel()
    assert prod.sum() == 4 * P and inj.sum() == 4 * P
    assert np.all(wd[~prod] == 0.0) and np.all(wd[prod] > 0.0)
    np.testing.assert_allclose(rhs[prod], wd[prod] * prm.bottomhole_pressure, rtol=1e-15)
    np.testing.assert_allclose(rhs[inj], prm.injection_rate / (4 * P), rtol=1e-15)
    # total injection equals the reference's rate * B_w
    assert rhs[inj].sum() == pytest.approx(5.0e3 * 5.61458 * 1.01, rel=1e-13)


def test_transport_conserves_water():
    """Interior upwind fluxes cancel: sum poro rate |E| = injection + producer
    terms, for any pressure and saturation."""
    rng = np.random.default_rng(5)
    mesh = (8, 10, 6)
    M, N, P = mesh
    Sw = 0.2 + 0.4 * rng.random((P, N, M))
    prm, lay, perm, poro, _, kappa, T, wd, rhs = small_case(mesh, S_w=Sw)
    p = 4000.0 + 100.0 * rng.standard_normal(M * N * P)
    r = S.saturation_rate(prm, lay, perm, kappa, poro, p, Sw)
    total = float((r * poro).sum() * lay.meas_elem)
    lw = S.mobility(prm, Sw)[0]
    prod = S.producer_mask(mesh)
    src = prm.injection_rate + float((lay.meas_elem * perm[0] * lw * prm.well_index
                                      * (prm.bottomhole_pressure - p.reshape(P, N, M)))[prod].sum())
    assert total == pytest.approx(src, rel=1e-10, abs=1e-9 * prm.injection_rate)


def test_transport_step_bounds_and_dt():
    prm, lay, perm, poro, Sw, kappa, T, wd, rhs = small_case()
    orc = O.Oracle(lay.dom, lay.mesh, kappa, sd=lay.sd, en_lv1=4, en_lv2=2,
                   eg_op=O.EIG_OP_DIAG, wd=wd, lv2_icc=True, no_shift_A_cc=True).setup()
    x, its, hist, reason = O.pcg(orc.matvec, rhs, orc.apply, rtol=1e-8,
                                 x0=np.zeros_like(rhs))
    assert reason == "CONVERGED_RTOL"
    S1, dt = S.transport(prm, lay, perm, kappa, poro, x, Sw, 5, 0.001)
    assert dt > 0.0
    assert S1.min() >= S.S_W_BOUND_MIN - 1e-12 and S1.max() <= S.S_W_BOUND_MAX
    # every substep raises the largest rate cell by exactly DS_max
    assert S1.max() <= Sw.max() + 5 * 0.001 + 1e-12


def test_spe10_preconditioner_symmetric_and_pcg_converges():
    prm, lay, perm, poro, Sw, kappa, T, wd, rhs = small_case(mesh=(12, 12, 8), sd=(2, 3, 2))
    orc = O.Oracle(lay.dom, lay.mesh, kappa, sd=lay.sd, en_lv1=4, en_lv2=3,
                   eg_op=O.EIG_OP_DIAG, wd=wd, lv2_icc=True, no_shift_A_cc=True).setup()
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal((2, rhs.size))
    bu, bv = orc.apply(u), orc.apply(v)
    assert float(bu @ v) == pytest.approx(float(u @ bv), rel=1e-9)
    assert float(u @ bu) > 0.0
    x, its, hist, reason = O.pcg(orc.matvec, rhs, orc.apply, rtol=1e-8)
    assert reason == "CONVERGED_RTOL" and its < 60
    assert np.linalg.norm(orc.A @ x - rhs) <= 1e-8 * np.linalg.norm(rhs) * 1.0001


def test_pcg_nonzero_guess_uses_rhs_norm():
    """KSPConvergedDefault with a nonzero initial guess tests against ||b||:
    a guess within rtol ||b|| of the solution converges at iteration 0."""
    prob = PB.config("tiny")
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell).setup()
    b = prob.rhs.ravel()
    x, its, _, _ = O.pcg(orc.matvec, b, orc.apply, rtol=1e-12)
    x0 = x + 1e-11 * np.linalg.norm(x) * np.random.default_rng(2).standard_normal(x.size)
    r0 = np.linalg.norm(b - orc.matvec(x0))
    assert r0 < 1e-8 * np.linalg.norm(b)
    _, its0, _, reason = O.pcg(orc.matvec, b, orc.apply, rtol=1e-8, x0=x0)
    assert its0 == 0 and reason == "CONVERGED_RTOL"


def test_face_fluxes_divergence_is_operator():
    """-div F = A p (no wells): the velocity recovery is consistent with the
    operator; for constant kappa and p linear in x the flux is uniform."""
    prob = PB.config("tiny")
    lay = O.Layout(prob.dom, prob.mesh, prob.sd)
    T = O.transmissibilities(lay, prob.kappa)
    rng = np.random.default_rng(3)
    M, N, P = prob.mesh
    p = rng.standard_normal(M * N * P)
    Fx, Fy, Fz = O.face_fluxes(lay, T, p)
    out = np.zeros((P, N, M))
    out[:, :, :-1] += Fx
    out[:, :, 1:] -= Fx
    out[:, :-1, :] += Fy
    out[:, 1:, :] -= Fy
    out[:-1, :, :] += Fz
    out[1:, :, :] -= Fz
    np.testing.assert_allclose(out.ravel(), O.stencil_apply(lay, T, p), rtol=1e-12, atol=1e-12)
    k1 = np.ones((3, P, N, M))
    T1 = O.transmissibilities(lay, k1)
    xs = np.tile(np.arange(M, dtype=float), N * P)
    Fx1, Fy1, Fz1 = O.face_fluxes(lay, T1, xs)
    np.testing.assert_allclose(Fx1, -lay.meas_face_yz ** 2 / lay.meas_elem, rtol=1e-14)
    assert np.abs(Fy1).max() == 0.0 and np.abs(Fz1).max() == 0.0


def _upwind_case(sign):
    """A 6 x 6 x 5 reservoir whose pressure varies along z only, so the x and
    y faces carry no flux; column (1, 1) has no well."""
    mesh = (6, 6, 5)
    prm = S.Spe10Params()
    lay = S.layout(mesh, (2, 2, 1), prm.H)
    M, N, P = mesh
    rng = np.random.default_rng(4)
    perm = [np.exp(rng.standard_normal((P, N, M))) * 1e-13 for _ in range(3)]
    Sw = 0.2 + 0.6 * rng.random((P, N, M))
    kappa = S.kappa_from(prm, perm, Sw)
    poro = 0.1 + 0.2 * rng.random((P, N, M))
    p = np.broadcast_to((sign * (P - np.arange(P)) * 1.0e3)[:, None, None], (P, N, M)).copy()
    return prm, lay, perm, kappa, poro, p, Sw


@pytest.mark.parametrize("sign", [1.0, -1.0], ids=["flow+z", "flow-z"])
def test_saturation_rate_takes_the_upwind_fractional_flow(sign):
    """get_next_S_w_multisteps (src/spe10.c:1985-2032) weights each face flux
    by the fractional flow of the cell the flow comes from (the cell with
    the higher pressure).  Cell (1, 1, z) of a column whose pressure falls
    (flow +z) or rises (flow -z) along z, computed by hand per face: a
    downwind weighting gives a different rate."""
    prm, lay, perm, kappa, poro, p, Sw = _upwind_case(sign)
    rate = S.saturation_rate(prm, lay, perm, kappa, poro, p, Sw)
    lw, _, lt = S.mobility(prm, Sw)
    fw = lw / lt
    hz2 = lay.Hz * lay.Hz
    x = y = 1
    for z in (1, 2, 3):
        up, down = 0.0, 0.0
        for nz in (z - 1, z + 1):
            kb = 2.0 / (1.0 / kappa[2][nz, y, x] + 1.0 / kappa[2][z, y, x])
            flux_in = kb * (p[nz, y, x] - p[z, y, x]) / hz2
            src = nz if p[nz, y, x] > p[z, y, x] else z
            dst = z if src == nz else nz
            up += fw[src, y, x] * flux_in
            down += fw[dst, y, x] * flux_in
        want = up / poro[z, y, x]
        assert rate[z, y, x] == pytest.approx(want, rel=1e-12)
        assert abs(rate[z, y, x] - down / poro[z, y, x]) > 1e-6 * abs(want)
