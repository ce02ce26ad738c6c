"""Every benchmark workload at its own size through the GPU library, checked
against the FP64 oracle: directly where the oracle is cheap, through the
committed full-size oracle fixtures (``tests/golden``, written by
``tests/golden/make_fixtures.py``) where it takes minutes to an hour, and
through size-independent properties (true residual by an independent
stencil, symmetry of the preconditioner, subdomain eigenpairs against an
independent eigensolver) at the largest size.

Tolerances (FP64 preconditioner): PCG iterations within 1 of the oracle and
the pressure within 1e-10 relative; where the GPU uses its own subdomain
eigenvectors (converged to the eigensolver tolerance 1e-8, like SLEPc's)
rather than the oracle's, the preconditioner differs from the oracle's by
that much, and on the 1e5-contrast media the pressure within 1e-7.  FP32
preconditioner inside the FP64 CG: iterations within 2 (0 to 4 above on the
pinned 1e5-contrast bricks, the spread measured over LOBPCG seeds), pressure
within 1e-6 (the FP32 rounding of the apply, 1e-7 relative, enters every
search direction).
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

from oracle import premixfem_oracle as O
from oracle import spe10_oracle as S
from premixfem_amd import pc as P
from premixfem_amd import problems, runner

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _fixture(name):
    path = GOLDEN / f"{name}.npz"
    if not path.exists():
        pytest.fail(f"missing fixture {path}: run tests/golden/make_fixtures.py {name}")
    return np.load(path, allow_pickle=False)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_field(prob, f):
    """The generated permeability is the one the fixture was made from."""
    logsum = np.array([np.log(prob.kappa[d]).sum() for d in range(3)])
    np.testing.assert_allclose(logsum, f["kappa_logsum"], rtol=1e-12)
    np.testing.assert_array_equal(prob.kappa[0].ravel()[f["cells"]], f["kappa_sample"])


# ---------------------------------------------------------------------------
# 32^3 GRF (sigma 2, corr 0.1), 4^3 subdomains of 8^3, 4 eigenvectors
# ---------------------------------------------------------------------------
def test_c1_fp64_matches_oracle():
    prob = problems.config("c1")
    f = _fixture("c1")
    _check_field(prob, f)
    print("c1 fp64", flush=True)
    res = runner.run(prob, precision=64, keep_ctx=True)
    try:
        y = P.PC_apply_vec(res["shell"], np.random.default_rng(777).standard_normal(prob.ncell))
    finally:
        P.PC_final(res["ctx"])
    cells = f["cells"]
    ex = _rel(res["x"][cells], f["x_sample"])
    ey = _rel(y[cells], f["By_sample"])
    ev = np.abs(np.asarray(res["stats"]["eigen_max_lv1_range"]) -
                np.array([f["eig_lv1"][:, -1].min(), f["eig_lv1"][:, -1].max()]))
    print(f"c1: its {res['its']} vs {int(f['its'])}, x rel err {ex:.2e}, B v rel err {ey:.2e}, "
          f"eigen_max_lv1 diff {ev}", flush=True)
    assert res["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - int(f["its"])) <= 1
    assert ex <= 1e-10
    assert ey <= 1e-8      # B with the GPU's own eigenvectors (tolerance 1e-8)
    assert abs(np.linalg.norm(res["x"]) - float(f["x_norm"])) <= 1e-10 * float(f["x_norm"])
    assert np.all(ev <= 1e-8 * f["eig_lv1"][:, -1].max())


def test_c1_fp32_preconditioner():
    prob = problems.config("c1")
    f = _fixture("c1")
    res = runner.run(prob, precision=32)
    ex = _rel(res["x"][f["cells"]], f["x_sample"])
    print(f"c1 fp32: its {res['its']} vs {int(f['its'])}, x rel err {ex:.2e}", flush=True)
    assert res["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - int(f["its"])) <= 2
    assert ex <= 1e-6


# ---------------------------------------------------------------------------
# 128^3 binary box medium (1e5 contrast), 8^3 subdomains of 16^3, 6 eigenvectors
# ---------------------------------------------------------------------------
def test_c2_fp64_against_oracle_fixture():
    """The 1e5-contrast medium at full size.  Its pinned coarse matrix has
    lambda_min(A_c) ~ 5.7e-12 (the fixture's first level-3 eigenvalue), so the
    coarse solves amplify rounding by ~1e11 and PCG needs ~1500 iterations,
    a count that moves by a few percent with the rounding of the coarse solve
    alone: the oracle takes 1478 iterations with the full cycle and 1479 with
    the cycle cut to [lv1, lv2, lv1] (both recorded in the fixture; an
    earlier assembly order of the same oracle took 1461).  The GPU uses its own
    subdomain eigenvectors (eigensolver tolerance 1e-8), so its rtol-1e-8
    solution carries a different unconverged error than the oracle's: the
    reference is the GPU's solve to rtol 1e-11, against which the GPU's
    rtol-1e-8 pressure must be as accurate as the oracle's (within 3x) at the
    fixture's cells.  Checked besides: the count within 5 % of the oracle's
    and the true residual at the tolerance."""
    prob = problems.config("c2")
    f = _fixture("c2")
    _check_field(prob, f)
    print("c2 fp64", flush=True)
    res = runner.run(prob, precision=64, keep_ctx=True)
    try:
        _, ev3 = P.PC_get_basis(res["ctx"], 3)
    finally:
        P.PC_final(res["ctx"])
    st = res["stats"]
    # Level 3: the fixture's eigenvalues of A_c (the first is the pinned
    # near-null one, ~5.7e-12), reached within the accepted rounding floor.
    e_lv3 = np.abs(ev3 - f["eig_lv2"]) / f["eig_lv2"]
    print(f"c2 level 3: eigenvalues rel err {e_lv3}, {st['lv3_iters']} its, floor "
          f"{st['lv3_floor']:.2e}", flush=True)
    assert np.all(e_lv3 <= 1e-5)
    assert st["lv3_floor"] <= 100 * 1e-8
    cells = f["cells"]
    ex = _rel(res["x"][cells], f["x_sample"])
    lay = O.Layout(prob.dom, prob.mesh, prob.sd)
    T = O.transmissibilities(lay, prob.kappa)
    pin = (tuple(prob.pin_cell), O.pin_tau(lay, T, prob.pin_cell))
    b = prob.rhs.ravel()
    r = np.linalg.norm(b - O.stencil_apply(lay, T, res["x"], pin)) / np.linalg.norm(b)
    its_o, its_c = int(f["its"]), int(f["its_collapsed"])
    tight = runner.run(prob, precision=64, rtol=1e-11)
    xs = tight["x"][cells]
    e_gpu, e_orc = _rel(res["x"][cells], xs), _rel(f["x_sample"], xs)
    print(f"c2: its {res['its']} vs oracle {its_o} (cut cycle {its_c}), x vs oracle {ex:.2e}; "
          f"true rel residual {r:.2e} (oracle {float(f['true_rel_residual']):.2e}); error vs the "
          f"rtol-1e-11 solve ({tight['its']} its): GPU {e_gpu:.2e}, oracle {e_orc:.2e}",
          flush=True)
    assert res["reason"] == "CONVERGED_RTOL" and tight["reason"] == "CONVERGED_RTOL"
    assert abs(res["its"] - its_o) <= 0.05 * its_o
    assert r <= 2e-8
    assert e_gpu <= 3.0 * e_orc


# ---------------------------------------------------------------------------
# 16^3-cell subdomains (the 128^3 and 256^3 workloads' subdomain size) on a
# 32^3 grid: the 512-thread LOBPCG instantiations at block sizes 8 and 10
# and the quad-cell stream kernels on 16^3 bricks, against the oracle.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def brick16():
    mesh = (32, 32, 32)
    kap = problems.box_inclusions_kappa(mesh, 47, 2, 6, 1.0e5, 2001)
    prob = problems.Problem("brick16", (1.0, 1.0, 1.0), mesh, kap,
                            problems.corner_rhs(mesh, (1.0, 1.0, 1.0)), sd=(2, 2, 2), en_lv1=6,
                            en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell).setup()
    x, its, _, reason = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    assert reason == "CONVERGED_RTOL"
    return prob, orc, x, its


@pytest.mark.parametrize("precision", [64, 32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("mb", [8, 10], ids=["MB8", "MB10"])
def test_16cube_subdomains_lobpcg512_and_quad_kernels(brick16, mb, precision):
    prob, orc, x_ref, its_ref = brick16
    guard = mb - (prob.en_lv1 - 1)
    print("16^3 bricks", mb, precision, flush=True)
    res = runner.run(prob, precision=precision, extra=f"-eig_guard {guard} -eig_autotune 0",
                     keep_ctx=True)
    try:
        _, ev = P.PC_get_basis(res["ctx"], 2)
    finally:
        P.PC_final(res["ctx"])
    st = res["stats"]
    assert st["eig_block"] == mb
    ev = ev.reshape(orc.lay.nsub, prob.en_lv1)
    worst = max(float(np.max(np.abs(ev[s, 1:] - orc.eig_lv1[s][1:]) / orc.eig_lv1[s][1:]))
                for s in range(orc.lay.nsub))
    err = _rel(res["x"], x_ref)
    print(f"MB {mb} fp{precision}: its {res['its']} vs {its_ref}, eig rel err {worst:.2e}, "
          f"x rel err {err:.2e}, lobpcg max its {st['eig_iters_max']}", flush=True)
    assert worst <= 1e-8
    assert res["reason"] == "CONVERGED_RTOL"
    if precision == 64:
        assert abs(res["its"] - its_ref) <= 1
        assert err <= 1e-7
        # The same kernels with the oracle's eigenvectors: one application of
        # the preconditioner against the oracle's cycle cut to [lv1, lv2, lv1]
        # (what the GPU applies on one process; the full cycle equals it in
        # exact arithmetic).  On this pinned 1e5-contrast medium A_c has
        # lambda_min ~1e-9 of its largest eigenvalue, so two backward-stable
        # coarse solves (the GPU's deflated inverse, the oracle's Cholesky)
        # agree to cond(A_c) * eps ~ 1e-7 in one application.
        phi, num, evo = O.basis_columns(orc)
        rb = runner.run(prob, precision=64, basis=(phi, num, evo),
                        extra=f"-eig_guard {guard} -eig_autotune 0", keep_ctx=True)
        try:
            v = np.random.default_rng(3).standard_normal(prob.ncell)
            y = P.PC_apply_vec(rb["shell"], v)
        finally:
            P.PC_final(rb["ctx"])
        y1 = orc.apply_lv1(v)
        y1 = y1 + orc.apply_lv2(v - orc.matvec(y1))
        y_ref = y1 + orc.apply_lv1(v - orc.matvec(y1))
        ea = _rel(y, y_ref)
        eb = _rel(rb["x"], x_ref)
        print(f"MB {mb} fp64 with the oracle basis: B v rel err {ea:.2e}, its {rb['its']}, "
              f"x rel err {eb:.2e}", flush=True)
        assert ea <= 1e-6
        assert abs(rb["its"] - its_ref) <= 1
        assert eb <= 1e-7
    else:
        # FP32 rounding of the preconditioner, amplified by cond(A_c) ~ 1e9 of
        # this pinned medium, spreads the count: over LOBPCG seeds 1..8 the
        # FP64 form takes 96 every time and the FP32 one 97-100 (MB 8 and 10).
        assert 0 <= res["its"] - its_ref <= 4
        assert err <= 1e-6


@pytest.mark.parametrize("mb", [8, 10], ids=["MB8", "MB10"])
def test_16cube_lobpcg_kernels_agree(brick16, mb):
    """The matrix-core LOBPCG (-eig_kernel 1, the default) in both workgroup
    shapes (one 512-thread workgroup per CU; two 256-thread ones with 4-column
    sweep chunks, -eig_shape 1), with W orthogonalised in coefficient space
    (-eig_ortho 1, the default) or by two explicit projection passes, and the
    VALU one (-eig_kernel 0) on the 16^3 subdomains: all within 1e-8 of the
    oracle's eigenvalues, the same PCG iteration count, and pressures within
    the 1e-7 the eigenvector tolerance allows on this medium."""
    prob, orc, x_ref, its_ref = brick16
    guard = mb - (prob.en_lv1 - 1)
    out = {}
    for kern, shape, ortho in ((1, 0, 1), (0, 0, 1), (1, 1, 1), (1, 1, 0)):
        print(f"16^3 bricks MB {mb} eig_kernel {kern} eig_shape {shape} eig_ortho {ortho}",
              flush=True)
        res = runner.run(prob, precision=64, keep_ctx=True,
                         extra=f"-eig_guard {guard} -eig_autotune 0 -eig_kernel {kern} "
                               f"-eig_shape {shape} -eig_ortho {ortho}")
        try:
            _, ev = P.PC_get_basis(res["ctx"], 2)
        finally:
            P.PC_final(res["ctx"])
        ev = ev.reshape(orc.lay.nsub, prob.en_lv1)
        worst = max(float(np.max(np.abs(ev[s, 1:] - orc.eig_lv1[s][1:]) / orc.eig_lv1[s][1:]))
                    for s in range(orc.lay.nsub))
        st = res["stats"]
        print(f"  its {res['its']}, eig rel err {worst:.2e}, lobpcg its max "
              f"{st['eig_iters_max']} mean {st['eig_iters_mean']:.1f}", flush=True)
        assert st["eig_block"] == mb
        assert st["eig_shape"] == shape
        assert worst <= 1e-8
        assert res["reason"] == "CONVERGED_RTOL"
        out[(kern, shape, ortho)] = (ev, res["x"], res["its"])
    ref = out[(1, 0, 1)]
    assert abs(ref[2] - its_ref) <= 1
    for key in ((0, 0, 1), (1, 1, 1), (1, 1, 0)):
        assert abs(out[key][2] - ref[2]) <= 1
        assert _rel(out[key][1], ref[1]) <= 1e-7
        assert np.allclose(out[key][0][:, 1:], ref[0][:, 1:], rtol=1e-8)


# ---------------------------------------------------------------------------
# 60x220x85 reservoir grid, first two IMPES steps of the spe10 fork
# ---------------------------------------------------------------------------
def test_c3_impes_steps_match_oracle_fixture():
    from premixfem_amd import spe10

    f = _fixture("c3")
    mesh, sd = tuple(int(v) for v in f["mesh"]), tuple(int(v) for v in f["sd"])
    perm, poro = problems.reservoir_fields(mesh, 3001)
    np.testing.assert_allclose([np.log(perm[d]).sum() for d in range(3)], f["perm_logsum"],
                               rtol=1e-12)
    print("c3 impes 2 steps", mesh, sd, flush=True)
    res = spe10.Reservoir(mesh, perm, poro, sd, precision=64)
    try:
        rows = res.run(max_steps=2, rtol=1e-8)
        sw, p = res.state()
    finally:
        res.close()
    cells = f["cells"]
    for k in range(2):
        print(f"c3 step {k}: its {rows[k]['its']} vs {int(f['its'][k])}, dt {rows[k]['delta_t']:.8e}"
              f" vs {float(f['delta_t'][k]):.8e}", flush=True)
        assert abs(rows[k]["its"] - int(f["its"][k])) <= 1
        assert rows[k]["delta_t"] == pytest.approx(float(f["delta_t"][k]), rel=1e-7)
    ep = _rel(np.asarray(p).ravel()[cells], f["p_sample"][1])
    es = _rel(np.asarray(sw).ravel()[cells], f["S_w_sample"][1])
    print(f"c3: pressure rel err {ep:.2e}, S_w rel err {es:.2e}", flush=True)
    assert ep <= 1e-8
    assert es <= 1e-7


# ---------------------------------------------------------------------------
# 256^3 fracture medium, 16^3 subdomains of 16^3 cells, 8 eigenvectors
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4():
    return problems.config("c4")


def test_c4_fp64_full_size_properties(c4):
    prob = c4
    print("c4 fp64", flush=True)
    res = runner.run(prob, precision=64, keep_ctx=True)
    ctx, shell = res["ctx"], res["shell"]
    try:
        st = res["stats"]
        assert res["reason"] == "CONVERGED_RTOL"
        # true residual with the oracle's stencil (independent of the GPU SpMV)
        lay = O.Layout(prob.dom, prob.mesh, prob.sd)
        T = O.transmissibilities(lay, prob.kappa)
        pin = (tuple(prob.pin_cell), O.pin_tau(lay, T, prob.pin_cell))
        b = prob.rhs.ravel()
        rres = np.linalg.norm(b - O.stencil_apply(lay, T, res["x"], pin)) / np.linalg.norm(b)
        # symmetry of the preconditioner
        rng = np.random.default_rng(5)
        u, v = rng.standard_normal((2, prob.ncell))
        bu, bv = P.PC_apply_vec(shell, u), P.PC_apply_vec(shell, v)
        asym = abs(bu @ v - u @ bv) / abs(bu @ v)
        # level-2 eigenpairs of three subdomains against the oracle's solver
        phi, ev = P.PC_get_basis(ctx, 2)
    finally:
        P.PC_final(ctx)
    ev = ev.reshape(lay.nsub, prob.en_lv1)
    mdiag = O.mass_diag(lay, prob.kappa, T)
    M, N, Pz = prob.mesh
    gidx = np.arange(M * N * Pz).reshape(Pz, N, M)
    worst_l, worst_r = 0.0, 0.0
    for s in (0, lay.nsub // 2 + 7, lay.nsub - 1):
        stx, nx, sty, ny, stz, nz = lay.corners(s)
        cells = gidx[stz:stz + nz, sty:sty + ny, stx:stx + nx].ravel()
        AN = O.local_neumann(lay, T, s)
        lam, _ = O.ghep(AN, mdiag[cells], prob.en_lv1)
        worst_l = max(worst_l, float(np.max(np.abs(ev[s, 1:] - lam[1:]) / lam[1:])))
        V = phi[:, cells].T
        R = AN @ V - (V * mdiag[cells][:, None]) * ev[s][None, :]
        worst_r = max(worst_r, float(np.max(np.linalg.norm(R[:, 1:], axis=0) /
                                            (ev[s, 1:] * np.linalg.norm(V[:, 1:] *
                                                                        mdiag[cells][:, None],
                                                                        axis=0)))))
    print(f"c4 fp64: its {res['its']}, true residual {rres:.2e}, B asymmetry {asym:.2e}, "
          f"eig rel err {worst_l:.2e}, eig residual {worst_r:.2e}, coarse mode "
          f"{st['coarse_mode']}", flush=True)
    assert rres <= 2e-8
    assert asym <= 1e-10
    assert worst_l <= 1e-7
    assert worst_r <= 1e-4
    path = GOLDEN / "c4.npz"
    if path.exists():
        f = np.load(path, allow_pickle=False)
        _check_field(prob, f)
        ex = _rel(res["x"][f["cells"]], f["x_sample"])
        print(f"c4 vs oracle fixture: its {res['its']} vs {int(f['its'])}, x rel err {ex:.2e}",
              flush=True)
        assert abs(res["its"] - int(f["its"])) <= 1
        assert ex <= 1e-8


def test_c4_fp32_and_learned_solve_alike(c4):
    """The FP32 preconditioner and the learned warm start of the subdomain
    eigensolver (a network trained on one 64^3 field of the same family)
    leave the iteration count and the pressure of the FP32 solve unchanged:
    LOBPCG refines the prediction to the same tolerance."""
    import torch

    torch.cuda.init()     # torch's HIP context before the network's first tensor
    print("c4 learned: torch sees", torch.cuda.device_count(), "device(s)", flush=True)
    from premixfem_amd import cnn

    prob = c4
    data = [cnn.sample_problem(problems.training_problem("c4", 10000, (64, 64, 64)))]
    net = cnn.WarmStartNet(cout=prob.en_lv1 - 1, seed=1)
    net.train(data, epochs=8, batch=64, lr=2e-3)
    try:
        a = runner.run(prob, precision=32)
        b = runner.run(prob, precision=32, extra="-warm_start cnn", net=net)
    finally:
        net.close()
    err = _rel(b["x"], a["x"])
    print(f"c4 fp32 its {a['its']}, learned its {b['its']}, x rel diff {err:.2e}, lobpcg mean its "
          f"{a['stats']['eig_iters_mean']:.1f} / {b['stats']['eig_iters_mean']:.1f}", flush=True)
    assert a["reason"] == "CONVERGED_RTOL" and b["reason"] == "CONVERGED_RTOL"
    assert abs(a["its"] - b["its"]) <= 1
    assert err <= 1e-8


# ---------------------------------------------------------------------------
# 64 independent problems of {32..96}^3 cells in one batched setup
# ---------------------------------------------------------------------------
def test_c5_batch_of_64_matches_members_and_oracle():
    probs = problems.batch_config(64)
    print("c5 batch", sum(p.ncell for p in probs), "cells", flush=True)
    outs, t_setup = runner.run_batch(probs, precision=64, extra="-coarse_composite 0")
    assert all(o["reason"] == "CONVERGED_RTOL" for o in outs)
    sizes = [p.mesh[0] for p in probs]
    # one member of each size against its own setup
    for n in sorted(set(sizes)):
        i = sizes.index(n)
        one = runner.run(probs[i], precision=64, extra="-coarse_composite 0")
        assert one["its"] == outs[i]["its"], (i, n)
        assert _rel(outs[i]["x"], one["x"]) <= 1e-12, (i, n)
    # the smallest members against the oracle
    for n in (32, 48):
        i = sizes.index(n)
        p = probs[i]
        orc = O.Oracle(p.dom, p.mesh, p.kappa, sd=p.sd, en_lv1=p.en_lv1, en_lv2=p.en_lv2,
                       pin_cell=p.pin_cell).setup()
        x, its, _, _ = O.pcg(orc.matvec, p.rhs.ravel(), orc.apply, rtol=p.rtol)
        err = _rel(outs[i]["x"], x)
        print(f"c5 member {i} ({n}^3): its {outs[i]['its']} vs {its}, x rel err {err:.2e}",
              flush=True)
        assert abs(outs[i]["its"] - its) <= 1
        assert err <= 1e-8
    print(f"c5: batched setup {t_setup:.2f}s, its total {sum(o['its'] for o in outs)}", flush=True)
