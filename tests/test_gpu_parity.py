"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Tolerances: face coefficients / operator diagonals bit-exact (same
arithmetic, no contraction possible); coarse matrices and one V-cycle apply
1e-11 relative (FMA contraction and reduction order differ); PCG iteration
counts within +-1 and the FP64 pressure within 1e-8 relative of the oracle's
(both stop at rtol 1e-8; the difference is rounding propagated through the
iteration).  Eigenvalues from the batched LOBPCG agree with shift-invert
Lanczos to 1e-6 relative.
"""
import numpy as np
import pytest

from oracle import premixfem_oracle as O
from premixfem_amd import pc as P
from premixfem_amd import problems, runner

pytestmark = pytest.mark.gpu


def _uneven():
    mesh = (10, 9, 7)
    dom = (1.0, 0.9, 0.7)
    kap = problems.grf_kappa(mesh, 1.5, 0.2, 31)
    return problems.Problem("uneven", dom, mesh, kap, problems.corner_rhs(mesh, dom),
                            sd=(3, 2, 2), en_lv1=3, en_lv2=3, pin_cell=(9, 0, 0))


CASES = {
    "tiny": lambda: problems.config("tiny"),
    "small": lambda: problems.config("small"),
    "uneven": _uneven,
}


def _oracle(prob):
    return O.solve_problem(prob)


@pytest.mark.parametrize("name", list(CASES))
def test_create_A_bitexact(name):
    prob = CASES[name]()
    print("case", name, prob.mesh, prob.sd, flush=True)
    lay = O.Layout(prob.dom, prob.mesh, prob.sd)
    T = O.transmissibilities(lay, prob.kappa)
    d_ref, off_ref = O.create_A_diagonals(lay, T, prob.pin_cell)
    ctx = P.PCCtx()
    P.PC_init(ctx, prob.dom, prob.mesh, runner.options_for(prob))
    try:
        ctx.kappa = prob.kappa
        d, off = P.PC_create_A(ctx)
    finally:
        P.PC_final(ctx)
    assert np.array_equal(off, off_ref)
    assert np.array_equal(d, d_ref)


@pytest.mark.parametrize("name", list(CASES))
def test_apply_and_pcg_with_oracle_basis(name):
    prob = CASES[name]()
    print("case", name, flush=True)
    orc, x_ref, its_ref, _, _ = _oracle(prob)
    phi, num, ev = O.basis_columns(orc)
    res = runner.run(prob, basis=(phi, num, ev), keep_ctx=True)
    ctx, shell = res["ctx"], res["shell"]
    try:
        Ac = P.PC_get_coarse_matrix(ctx)
        scale = np.abs(orc.A_c).max()
        assert np.abs(Ac - orc.A_c).max() <= 1e-12 * scale
        st = res["stats"]
        assert st["size_mat_lv3"] == orc.eigen_num_lv2
        assert np.allclose(st["eigen_min_lv2"], orc.eig_lv2[0], rtol=1e-7, atol=1e-14 * scale)
        rng = np.random.default_rng(3)
        x = rng.standard_normal(prob.ncell)
        y = P.PC_apply_vec(shell, x)
        y_ref = orc.apply(x)
        assert np.linalg.norm(y - y_ref) <= 1e-10 * np.linalg.norm(y_ref)
    finally:
        P.PC_final(ctx)
    assert res["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 1
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-8, err


@pytest.mark.parametrize("kernel", [1, 0], ids=["mfma", "valu"])
@pytest.mark.parametrize("name", list(CASES))
def test_full_setup_lobpcg(name, kernel):
    prob = CASES[name]()
    print("case", name, "eig_kernel", kernel, flush=True)
    orc, x_ref, its_ref, _, _ = _oracle(prob)
    res = runner.run(prob, keep_ctx=True, extra=f"-eig_kernel {kernel}")
    ctx = res["ctx"]
    try:
        phi, ev = P.PC_get_basis(ctx, 2)
        ev = ev.reshape(orc.lay.nsub, prob.en_lv1)
        for s in range(orc.lay.nsub):
            lam_ref = orc.eig_lv1[s][1:]
            assert np.allclose(ev[s, 1:], lam_ref, rtol=1e-6), (s, ev[s], lam_ref)
        st = res["stats"]
        print("lobpcg iterations max", st["eig_iters_max"], "mean", st["eig_iters_mean"])
    finally:
        P.PC_final(ctx)
    assert res["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 1, (res["its"], its_ref)
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-7, err


def test_fp32_preconditioner():
    prob = problems.config("small")
    orc, x_ref, its_ref, _, _ = _oracle(prob)
    res = runner.run(prob, precision=32)
    assert res["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 2, (res["its"], its_ref)
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-6, err
