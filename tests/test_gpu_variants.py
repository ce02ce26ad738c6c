"""GPU parity of the PC_apply_vec variants and options against the FP64
oracle (src/PreMixFEM_3D.c:1840-1936): W-cycle, 2-level, extra smoothing
steps on levels 1 and 2, the MEAN and DIAG eigen mass operators, the
no-shift coarse-coarse matrix and an adaptive eigen count (eb_lv1).  The
oracle's level-2 basis is given to the GPU so that the comparison isolates
the cycle; y = B x within 1e-10, PCG iterations within 1 and the pressure
within 1e-8.  On one process the GPU applies only the first level-2 stage
of the coarse cycle by default (the later stages add exactly zero in exact
arithmetic); the "-full" variants run every stage as the oracle does."""
import numpy as np
import pytest

from oracle import premixfem_oracle as O
from premixfem_amd import pc as P
from premixfem_amd import problems, runner

pytestmark = pytest.mark.gpu

VARIANTS = [
    ("W-cycle", dict(use_W_cycle=True), "-use_W_cycle"),
    ("2-level", dict(use_2level=True), "-use_2level"),
    ("2-level-noshift", dict(use_2level=True, no_shift_A_cc=True), "-use_2level -no_shift_A_cc"),
    ("si_lv1=2", dict(si_lv1=2), "-si_lv1 2"),
    ("si_lv2=2", dict(si_lv2=2), "-si_lv2 2"),
    ("W-si2", dict(use_W_cycle=True, si_lv2=2), "-use_W_cycle -si_lv2 2"),
    # coarse stages one by one (no composed operator), with and without the
    # deflated inverse
    ("stagewise", dict(), "-coarse_composite 0"),
    ("stagewise-plain", dict(), "-coarse_composite 0 -coarse_deflate 0"),
    # exact level-2 factor of the large-coarse-space path: by z-planes of
    # subdomains (its default) and in band storage
    ("planes", dict(), "-coarse_band 1"),
    ("planes-W", dict(use_W_cycle=True), "-coarse_band 1 -use_W_cycle"),
    ("band", dict(), "-coarse_band 1 -coarse_planes 0"),
    ("band-W", dict(use_W_cycle=True), "-coarse_band 1 -coarse_planes 0 -use_W_cycle"),
    # exact level-1 blocks (band Cholesky per brick)
    ("full-Cholesky-lv1", dict(use_full_Cholesky_lv1=True), "-use_full_Cholesky_lv1"),
    ("full-Cholesky-lv1-W", dict(use_full_Cholesky_lv1=True, use_W_cycle=True),
     "-use_full_Cholesky_lv1 -use_W_cycle"),
    # every coarse stage of the cycle run (no cut to the first level-2
    # stage): the composed operator, the stage-wise and the band programs
    ("V-full", dict(), "-coarse_collapse 0"),
    ("W-cycle-full", dict(use_W_cycle=True), "-use_W_cycle -coarse_collapse 0"),
    ("W-si2-full", dict(use_W_cycle=True, si_lv2=2), "-use_W_cycle -si_lv2 2 -coarse_collapse 0"),
    ("stagewise-full", dict(), "-coarse_composite 0 -coarse_collapse 0"),
    ("band-W-full", dict(use_W_cycle=True),
     "-coarse_band 1 -coarse_planes 0 -use_W_cycle -coarse_collapse 0"),
    ("planes-W-full", dict(use_W_cycle=True), "-coarse_band 1 -use_W_cycle -coarse_collapse 0"),
]


def _prob():
    return problems.config("small")


@pytest.mark.parametrize("name,kw,opts", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_variant_matches_oracle(name, kw, opts):
    prob = _prob()
    print("variant", name, opts, flush=True)
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell, **kw)
    orc.setup()
    x_ref, its_ref, _, reason_ref = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    phi, num, ev = O.basis_columns(orc)
    res = runner.run(prob, basis=(phi, num, ev), extra=opts, keep_ctx=True)
    ctx, shell = res["ctx"], res["shell"]
    try:
        x = np.random.default_rng(3).standard_normal(prob.ncell)
        y = P.PC_apply_vec(shell, x)
        y_ref = orc.apply(x)
        e_apply = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    finally:
        P.PC_final(ctx)
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    print(f"{name}: apply rel err {e_apply:.2e}, its {res['its']} vs {its_ref}, x rel err "
          f"{err:.2e}", flush=True)
    assert e_apply <= 1e-10
    assert res["reason"] == "CONVERGED_RTOL" and reason_ref == "CONVERGED_RTOL"
    assert abs(res["its"] - its_ref) <= 1
    assert err <= 1e-8


def test_adaptive_eigen_count():
    """eb_lv1 (:368-384): subdomains keep eigenvectors below the bound only."""
    prob = _prob()
    probe = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                     en_lv2=prob.en_lv2, pin_cell=prob.pin_cell)
    probe.setup()
    allev = np.sort(np.concatenate([e[1:] for e in probe.eig_lv1]))
    eb = float(allev[len(allev) // 2])
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell, eb_lv1=eb)
    orc.setup()
    assert len(set(orc.eigen_num_lv1.tolist())) > 1
    x_ref, its_ref, _, _ = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    phi, num, ev = O.basis_columns(orc)
    res = runner.run(prob, basis=(phi, num, ev), extra=f"-eb_lv1 {eb!r}")
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    assert res["stats"]["size_coarse"] == int(orc.eigen_num_lv1.sum())
    assert abs(res["its"] - its_ref) <= 1
    assert err <= 1e-8


@pytest.mark.parametrize("eg_op", [O.EIG_OP_MEAN, O.EIG_OP_DIAG], ids=["MEAN", "DIAG"])
def test_eigen_mass_operators_full_setup(eg_op):
    """The MEAN and DIAG mass matrices (:265-343) through the GPU's own
    eigensolver: eigenvalues within 1e-6 of the oracle's, PCG within 1."""
    prob = _prob()
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, pin_cell=prob.pin_cell, eg_op=eg_op)
    orc.setup()
    x_ref, its_ref, _, _ = O.pcg(orc.matvec, prob.rhs.ravel(), orc.apply, rtol=prob.rtol)
    res = runner.run(prob, extra=f"-eg_op {eg_op}", keep_ctx=True)
    try:
        _, ev = P.PC_get_basis(res["ctx"], 2)
    finally:
        P.PC_final(res["ctx"])
    ev = ev.reshape(orc.lay.nsub, prob.en_lv1)
    for s in range(orc.lay.nsub):
        assert np.allclose(ev[s, 1:], orc.eig_lv1[s][1:], rtol=1e-6), (s, ev[s], orc.eig_lv1[s])
    err = np.linalg.norm(res["x"] - x_ref) / np.linalg.norm(x_ref)
    assert abs(res["its"] - its_ref) <= 1
    assert err <= 1e-7


def test_lobpcg_autotuner_keeps_the_basis():
    """-eig_autotune 1 picks a LOBPCG block size by timing a sample; the
    eigenvalues and the solve are those of the untuned run (to tolerance)."""
    prob = _prob()
    base = runner.run(prob)
    tuned = runner.run(prob, extra="-eig_autotune 1")
    st = tuned["stats"]
    print("autotuned block", st["eig_block"], "default", base["stats"]["eig_block"], flush=True)
    assert st["eig_block"] in (4, 6, 8, 10)
    assert abs(tuned["its"] - base["its"]) <= 1
    assert st["eigen_max_lv1_range"][1] == pytest.approx(base["stats"]["eigen_max_lv1_range"][1],
                                                         rel=1e-6)
    err = np.linalg.norm(tuned["x"] - base["x"]) / np.linalg.norm(base["x"])
    assert err <= 1e-8


def _uneven_prob():
    """Bricks with x-extents 12 and uneven z-extents (6, 5, 5): the quad-cell
    kernels' halo rows come from neighbours of other shapes."""
    mesh = (24, 20, 16)
    kap = problems.grf_kappa(mesh, 1.5, 0.2, 77)
    return problems.Problem("uneven", (1.0, 1.0, 1.0), mesh, kap,
                            problems.corner_rhs(mesh, (1.0, 1.0, 1.0)), sd=(2, 2, 3), en_lv1=4,
                            en_lv2=3, pin_cell=(mesh[0] - 1, 0, 0))


@pytest.mark.parametrize("name", ["small", "c2_small", "uneven"])
def test_quad_stream_kernels_match_per_cell(name):
    """The quad-cell stream kernels (x-extents multiple of 4) against the
    per-cell ones: the per-cell arithmetic is the same, the block partial sums
    run in another order, so the iterations agree within 1 and the solutions
    within the solver tolerance."""
    prob = _uneven_prob() if name == "uneven" else problems.config(name)
    print("quad vs per-cell", name, prob.mesh, prob.sd, flush=True)
    r1 = runner.run(prob, extra="-quad 1")
    r0 = runner.run(prob, extra="-quad 0")
    err = np.linalg.norm(r1["x"] - r0["x"]) / np.linalg.norm(r0["x"])
    print(f"{name}: its {r1['its']} vs {r0['its']}, x rel diff {err:.2e}", flush=True)
    assert r1["reason"] == "CONVERGED_RTOL" and r0["reason"] == "CONVERGED_RTOL"
    assert abs(r1["its"] - r0["its"]) <= 1
    assert err <= 1e-8


def test_bicc_prefetch_solve_is_bit_identical():
    """The block ICC(0) level-2 solve that loads its next round's rows ahead
    (en_lv1 = 8) gives bit for bit the plain kernel's PCG solution."""
    import dataclasses

    prob = dataclasses.replace(_prob(), en_lv1=8, en_lv2=4)
    opts = "-lv2_icc -eg_op 2 -no_shift_A_cc -bicc_prefetch "
    a = runner.run(prob, precision=64, extra=opts + "1")
    b = runner.run(prob, precision=64, extra=opts + "0")
    print(f"bicc prefetch: its {a['its']} vs {b['its']}", flush=True)
    assert a["its"] == b["its"]
    assert np.array_equal(a["x"], b["x"])


def test_bicc_prefetch_solve_is_bit_identical_with_multi_round_planes():
    """The same on the reservoir grid's subdomain split (10 x 20 x 17
    subdomains, here of 4^3 cells): up to 170 subdomains share a coarse
    hyperplane, so a plane takes several rounds of the prefetching sweep
    (its same-plane prefetch, the double-buffered right-hand side and the
    backward reads of D^-1 rows)."""
    mesh = (40, 80, 68)
    dom = (1.0, 1.0, 1.0)
    kap = problems.layered_kappa(mesh, 2.0, (10, 10, 2), 1e-3, 2e4, 31)
    prob = problems.Problem("planes", dom, mesh, kap, problems.corner_rhs(mesh, dom),
                            sd=(10, 20, 17), en_lv1=8, en_lv2=4, eg_op=2,
                            pin_cell=(mesh[0] - 1, 0, 0))
    sx, sy, sz = np.meshgrid(*(np.arange(n) for n in prob.sd), indexing="ij")
    widest = np.bincount((sx + sy + sz).ravel()).max()
    assert widest > 128, widest
    opts = "-lv2_icc -no_shift_A_cc -bicc_prefetch "
    print("bicc prefetch on", mesh, prob.sd, "widest plane", widest, flush=True)
    a = runner.run(prob, precision=64, extra=opts + "1")
    b = runner.run(prob, precision=64, extra=opts + "0")
    print(f"bicc prefetch (multi-round planes): its {a['its']} vs {b['its']}", flush=True)
    assert a["reason"] == "CONVERGED_RTOL"
    assert a["its"] == b["its"]
    assert np.array_equal(a["x"], b["x"])


@pytest.mark.parametrize("mesh,sd,k", [((32, 32, 32), (4, 4, 4), 4), ((30, 26, 33), (3, 2, 5), 5)],
                         ids=["32^3-sd4", "uneven-sd3x2x5"])
def test_plane_factor_matches_dense_inverse(mesh, sd, k):
    """The coarse solve by z-planes (block LDL^T over the planes of
    subdomains, explicit inverse Schur complements) against the dense
    deflated inverse on the same setup: level-3 eigenvalues, iterations and
    the pressure.  Uneven planes: 5 planes of 6 subdomains, odd eigen count."""
    kap = problems.grf_kappa(mesh, 2.0, 0.1, 91)
    prob = problems.Problem("planes", (1.0, 1.0, 1.0), mesh, kap,
                            problems.corner_rhs(mesh, (1.0, 1.0, 1.0)), sd=sd, en_lv1=k,
                            en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    print("planes vs dense", mesh, sd, k, flush=True)
    a = runner.run(prob, extra="-coarse_band 1")
    b = runner.run(prob)
    assert a["stats"]["coarse_mode"] == 4 and b["stats"]["coarse_mode"] in (0, 1)
    err = np.linalg.norm(a["x"] - b["x"]) / np.linalg.norm(b["x"])
    print(f"planes: its {a['its']} vs {b['its']}, x rel diff {err:.2e}, lv3 "
          f"{a['stats']['eigen_min_lv2']:.6e} vs {b['stats']['eigen_min_lv2']:.6e}", flush=True)
    assert a["reason"] == "CONVERGED_RTOL"
    assert abs(a["its"] - b["its"]) <= 1
    assert err <= 1e-9
    assert a["stats"]["eigen_max_lv2"] == pytest.approx(b["stats"]["eigen_max_lv2"], rel=1e-6)


def test_explicit_projections_refused_on_small_subdomains():
    """-eig_ortho 0 (the matrix-core LOBPCG's explicit projections) is off for
    subdomains of at most 512 cells: refused with PETSC_ERR_SUP at setup,
    before any launch."""
    prob = problems.config("small")
    with pytest.raises(P.PMXError) as e:
        runner.run(prob, extra="-eig_ortho 0 -eig_autotune 0")
    assert e.value.code == 56
