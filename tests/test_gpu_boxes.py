"""GPU parity of the virtual process boxes (-procs_virtual px,py,pz): the
P-process preconditioner of the reference on one GPU.  Level 2 keeps only
the couplings inside a box (_PC_setup_lv2_J, src/PreMixFEM_3D.c:806-888),
level 3 takes en_lv2 eigenvectors of every box's Neumann coarse matrix
(:918-1175) and couples them through A (:1177-1390).  Against the FP64
oracle: A_c within 1e-12, the per-box level-3 eigenvalues within 1e-9,
y = B x within 1e-10, PCG iterations within 1 and the pressure within 1e-8
(FP32 preconditioner: 2 iterations, 1e-6)."""
import numpy as np
import pytest

from oracle import premixfem_oracle as O
from premixfem_amd import pc as P
from premixfem_amd import problems, runner

pytestmark = pytest.mark.gpu


def _prob(mesh, en_lv1, en_lv2, seed):
    k = problems.grf_kappa(mesh, 1.5, 0.2, seed)
    return problems.Problem(f"boxes{mesh}", (1.0, 1.0, 1.0), mesh, k,
                            problems.corner_rhs(mesh, (1.0, 1.0, 1.0)), sd=(2, 2, 2),
                            en_lv1=en_lv1, en_lv2=en_lv2, pin_cell=(mesh[0] - 1, 0, 0))


CASES = [
    # name, mesh, en_lv1, en_lv2, procs, oracle kwargs, extra options
    ("2x1x2", (13, 10, 9), 3, 3, (2, 1, 2), {}, ""),
    ("2x1x2-stagewise", (13, 10, 9), 3, 3, (2, 1, 2), {}, "-coarse_composite 0"),
    ("2x1x2-W", (13, 10, 9), 3, 3, (2, 1, 2), dict(use_W_cycle=True), "-use_W_cycle"),
    ("2x1x2-eb_lv2", (13, 10, 9), 3, 3, (2, 1, 2), dict(eb_lv2=1e-9), "-eb_lv2 1e-9"),
    ("1x3x1-noshift", (12, 15, 8), 4, 2, (1, 3, 1), dict(no_shift_A_cc=True), "-no_shift_A_cc"),
    ("2x1x2-si2", (13, 10, 9), 3, 3, (2, 1, 2), dict(si_lv1=2, si_lv2=2), "-si_lv1 2 -si_lv2 2"),
    ("2x2x1-uneven", (11, 13, 7), 3, 2, (2, 2, 1), {}, ""),
    # 8 boxes x 4 vectors: a level-3 space above 16
    ("2x2x2", (16, 16, 16), 4, 4, (2, 2, 2), {}, ""),
    ("2x2x2-stagewise", (16, 16, 16), 4, 4, (2, 2, 2), {}, "-coarse_composite 0"),
]


def _opts(procs, extra):
    return ("-procs_virtual {},{},{} ".format(*procs) + extra).strip()


def _oracle(prob, procs, kw):
    return O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                    en_lv2=prob.en_lv2, pin_cell=prob.pin_cell, procs=procs, **kw).setup()


@pytest.mark.parametrize("name,mesh,k1,k2,procs,kw,extra", CASES, ids=[c[0] for c in CASES])
def test_boxes_match_oracle(name, mesh, k1, k2, procs, kw, extra):
    prob = _prob(mesh, k1, k2, 17)
    opts = _opts(procs, extra)
    print("boxes", name, mesh, opts, flush=True)
    orc = _oracle(prob, procs, kw)
    x_ref, its_ref, _, reason_ref = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    phi, num, ev = O.basis_columns(orc)
    res = runner.run(prob, basis=(phi, num, ev), extra=opts, keep_ctx=True)
    ctx, shell = res["ctx"], res["shell"]
    try:
        Ac = P.PC_get_coarse_matrix(ctx)
        _, ev3 = P.PC_get_basis(ctx, 3)
        st = P.PC_get_stats(ctx)
        x = np.random.default_rng(3).standard_normal(prob.ncell)
        y = P.PC_apply_vec(shell, x)
    finally:
        P.PC_final(ctx)
    y_ref = orc.apply(x)
    e_ac = np.abs(Ac - orc.A_c).max() / np.abs(orc.A_c).max()
    lam_ref = np.concatenate([lam[:n] for lam, n in zip(orc.eig_lv2_box, orc.eigen_num_lv2_box)])
    e_lam = np.abs(ev3 - lam_ref).max() / np.abs(lam_ref).max()
    e_apply = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    print(f"{name}: A_c {e_ac:.1e}, lv3 eig {e_lam:.1e} (K={st['size_mat_lv3']}), apply "
          f"{e_apply:.1e}, its {res['its']} vs {its_ref}, x {err:.1e}", flush=True)
    assert st["size_mat_lv3"] == orc.eigen_num_lv2 == len(ev3)
    assert e_ac <= 1e-12
    assert e_lam <= 1e-9
    assert e_apply <= 1e-10
    assert res["reason"] == "CONVERGED_RTOL" and reason_ref == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 1
    assert err <= 1e-8


VARIANTS = [
    # name, mesh, en_lv1, en_lv2, procs, oracle kwargs, extra options
    ("2x1x2-2level", (13, 10, 9), 3, 3, (2, 1, 2), dict(use_2level=True), "-use_2level"),
    ("2x2x1-2level-noshift", (11, 13, 7), 3, 2, (2, 2, 1),
     dict(use_2level=True, no_shift_A_cc=True), "-use_2level -no_shift_A_cc"),
    ("2x1x2-icc", (13, 10, 9), 3, 3, (2, 1, 2), dict(lv2_icc=True), "-lv2_icc"),
    ("2x2x2-icc-noshift", (16, 16, 16), 4, 4, (2, 2, 2), dict(lv2_icc=True, no_shift_A_cc=True),
     "-lv2_icc -no_shift_A_cc"),
    ("2x1x2-planes", (13, 10, 9), 3, 3, (2, 1, 2), {}, "-coarse_band 1"),
    ("2x2x2-planes-W", (16, 16, 16), 4, 4, (2, 2, 2), dict(use_W_cycle=True),
     "-coarse_band 1 -use_W_cycle"),
]


@pytest.mark.parametrize("name,mesh,k1,k2,procs,kw,extra", VARIANTS,
                         ids=[c[0] for c in VARIANTS])
def test_box_variants_match_oracle(name, mesh, k1, k2, procs, kw, extra):
    """Process boxes with the 2-level variant (one coarse matrix over all
    processes with the couplings between them, _PC_setup_lv2_c :423-735),
    the spe10 fork's ICC(0) level 2 (one SBAIJ factor per process,
    spe10.c:1074-1205) and the exact coarse solve by z-planes of the
    box-local A_c: y = B x within 1e-10, level-3 eigenvalues within 1e-9
    (none for 2-level), the 2-level A_c within 1e-12, PCG iterations within
    1 and the pressure within 1e-8 of the oracle (given its basis)."""
    prob = _prob(mesh, k1, k2, 17)
    opts = _opts(procs, extra)
    print("boxes", name, mesh, opts, flush=True)
    orc = _oracle(prob, procs, kw)
    x_ref, its_ref, _, reason_ref = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    phi, num, ev = O.basis_columns(orc)
    res = runner.run(prob, basis=(phi, num, ev), extra=opts, keep_ctx=True)
    ctx, shell = res["ctx"], res["shell"]
    two_level = kw.get("use_2level", False)
    try:
        Ac = P.PC_get_coarse_matrix(ctx)
        ev3 = None if two_level else P.PC_get_basis(ctx, 3)[1]
        x = np.random.default_rng(3).standard_normal(prob.ncell)
        y = P.PC_apply_vec(shell, x)
    finally:
        P.PC_final(ctx)
    y_ref = orc.apply(x)
    e_apply = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    e_ac = np.abs(Ac - orc.A_c).max() / np.abs(orc.A_c).max()
    msg = (f"{name}: A_c {e_ac:.1e}, apply {e_apply:.1e}, its {res['its']} vs {its_ref}, "
           f"x {err:.1e}")
    assert e_ac <= 1e-12
    if not two_level:
        lam_ref = np.concatenate([lam[:n] for lam, n in zip(orc.eig_lv2_box,
                                                            orc.eigen_num_lv2_box)])
        e_lam = np.abs(ev3 - lam_ref).max() / np.abs(lam_ref).max()
        msg += f", lv3 eig {e_lam:.1e}"
        assert len(ev3) == orc.eigen_num_lv2
        assert e_lam <= 1e-9
    print(msg, flush=True)
    assert e_apply <= 1e-10
    assert res["reason"] == "CONVERGED_RTOL" and reason_ref == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 1
    assert err <= 1e-8


def test_box_band_factor_refused():
    """The band Cholesky coarse factor has no process-box form: refused with
    PETSC_ERR_SUP, the plane factor is the exact solve offered instead."""
    prob = _prob((13, 10, 9), 3, 3, 17)
    with pytest.raises(P.PMXError) as e:
        runner.run(prob, extra=_opts((2, 1, 2), "-coarse_band 1 -coarse_planes 0"))
    assert e.value.code == 56


@pytest.mark.parametrize("precision", [64, 32])
def test_boxes_full_gpu_setup(precision):
    """The whole setup on the GPU (its own LOBPCG basis) against the oracle's
    PCG; the FP32 preconditioner within 2 iterations and 1e-6."""
    procs = (2, 1, 2)
    prob = _prob((13, 10, 9), 3, 3, 23)
    orc = _oracle(prob, procs, {})
    x_ref, its_ref, _, _ = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    res = runner.run(prob, precision=precision, extra=_opts(procs, ""))
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    print(f"full setup p{precision}: its {res['its']} vs {its_ref}, x {err:.1e}", flush=True)
    assert res["reason"] == "CONVERGED_RTOL"
    if precision == 64:
        assert abs(res["its"] - its_ref) <= 1 and err <= 1e-8
    else:
        assert abs(res["its"] - its_ref) <= 2 and err <= 1e-6


def test_one_box_is_the_default_path():
    prob = _prob((13, 10, 9), 3, 3, 17)
    a = runner.run(prob, extra="-procs_virtual 1,1,1")
    b = runner.run(prob)
    assert a["its"] == b["its"]
    assert np.array_equal(a["x"], b["x"])


def test_boxes_in_a_batched_setup():
    """Two problems with process boxes set up by one pmx_setup_batch call end
    up as their own setups (iterations equal, solution within 1e-12), and
    the FP32 learned warm start leaves the box result unchanged."""
    from premixfem_amd.cnn import WarmStartNet

    probs = [_prob((13, 10, 9), 3, 3, 31), _prob((12, 15, 8), 3, 3, 32)]
    opts = _opts((2, 1, 2), "")
    outs, _ = runner.run_batch(probs, precision=64, extra=opts)
    for prob, o in zip(probs, outs):
        ref = runner.run(prob, precision=64, extra=opts)
        e = np.linalg.norm(o["x"] - ref["x"]) / np.linalg.norm(ref["x"])
        print(f"batch member {prob.mesh}: its {o['its']} vs {ref['its']}, x {e:.1e}", flush=True)
        assert o["its"] == ref["its"] and e <= 1e-12
    net = WarmStartNet(probs[0].en_lv1 - 1, seed=5)
    try:
        warm = runner.run(probs[0], precision=32, extra=opts + " -warm_start cnn", net=net)
    finally:
        net.close()
    cold = runner.run(probs[0], precision=32, extra=opts)
    e = np.linalg.norm(warm["x"] - cold["x"]) / np.linalg.norm(cold["x"])
    print(f"warm start: its {warm['its']} vs {cold['its']}, x {e:.1e}", flush=True)
    assert abs(warm["its"] - cold["its"]) <= 1 and e <= 1e-6
