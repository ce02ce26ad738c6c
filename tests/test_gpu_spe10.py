"""GPU parity of the spe10 fork's pieces against the oracle: velocity
recovery (face fluxes), the pressure system with the producer-well term, the
fork's preconditioner (well term in level 1 and the eigenproblems, DIAG
mass, ICC(0) level-2 factor, no NIL shift), the upwind transport and the
IMPES loop (src/spe10.c:197-274)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import premixfem_oracle as O
from oracle import spe10_oracle as S
from premixfem_amd import pc as P
from premixfem_amd import problems as PB
from premixfem_amd import spe10
from premixfem_amd._lib import check, lib

pytestmark = pytest.mark.gpu

MESH = (16, 18, 15)      # the reference's DEBUG_LOCAL grid (src/spe10.c:351-354)
SD = (2, 3, 3)


def _fields(mesh=MESH, seed=3001):
    perm, poro = PB.reservoir_fields(mesh, seed)
    return perm, poro


def _oracle(lay, kappa, wd, sd=SD, en_lv1=8, en_lv2=4, **kw):
    return O.Oracle(lay.dom, lay.mesh, kappa, sd=sd, en_lv1=en_lv1, en_lv2=en_lv2,
                    eg_op=O.EIG_OP_DIAG, wd=wd, lv2_icc=True, no_shift_A_cc=True, **kw).setup()


def test_velocity_matches_oracle_and_balances_mass():
    prob = PB.config("small")
    print("velocity small", flush=True)
    ctx = P.PCCtx()
    P.PC_init(ctx, prob.dom, prob.mesh, "-sd 2")
    try:
        ctx.kappa = prob.kappa
        P._push_kappa(ctx)
        rng = np.random.default_rng(4)
        p = rng.standard_normal(prob.ncell)
        f = [np.zeros(prob.ncell) for _ in range(3)]
        check(lib().pmx_velocity(ctx.handle, p, *f))
        lay = O.Layout(prob.dom, prob.mesh, prob.sd)
        T = O.transmissibilities(lay, prob.kappa)
        Fx, Fy, Fz = O.face_fluxes(lay, T, p)
        M, N, Pz = prob.mesh
        fx, fy, fz = (a.reshape(Pz, N, M) for a in f)
        # fx[c] is the flux from c-x into c = F of the face below c
        np.testing.assert_allclose(fx[:, :, 1:], Fx, rtol=1e-14, atol=1e-14 * np.abs(Fx).max())
        np.testing.assert_allclose(fy[:, 1:, :], Fy, rtol=1e-14, atol=1e-14 * np.abs(Fy).max())
        np.testing.assert_allclose(fz[1:, :, :], Fz, rtol=1e-14, atol=1e-14 * np.abs(Fz).max())
        assert not fx[:, :, 0].any() and not fy[:, 0, :].any() and not fz[0].any()
        # mass balance: outflow of every cell = A p (no pin, no wells)
        out = np.zeros((Pz, N, M))
        out[:, :, :-1] += fx[:, :, 1:]
        out[:, :, 1:] -= fx[:, :, 1:]
        out[:, :-1, :] += fy[:, 1:, :]
        out[:, 1:, :] -= fy[:, 1:, :]
        out[:-1] += fz[1:]
        out[1:] -= fz[1:]
        np.testing.assert_allclose(out.ravel(), O.stencil_apply(lay, T, p), rtol=1e-12,
                                   atol=1e-12 * np.abs(out).max())
    finally:
        P.PC_final(ctx)


def test_pressure_system_matches_oracle():
    perm, poro = _fields()
    print("spe10 system", MESH, flush=True)
    res = spe10.Reservoir(MESH, perm, poro, SD)
    try:
        rng = np.random.default_rng(6)
        M, N, Pz = MESH
        Sw = 0.2 + 0.5 * rng.random((Pz, N, M))
        res.set_state(S_w=Sw)
        kap, wd, rhs = res.system()
        prm = S.Spe10Params()
        lay = S.layout(MESH, SD, prm.H)
        kref, T, wref, rref = S.linear_system(prm, lay, list(perm), Sw)
        for a in range(3):
            np.testing.assert_allclose(kap[a], kref[a].ravel(), rtol=1e-14)
        np.testing.assert_allclose(wd, wref, rtol=1e-14)
        np.testing.assert_allclose(rhs, rref, rtol=1e-14)
        # the operator the context assembles includes the well term
        res.assemble()
        n = M * N * Pz
        diag, off = np.zeros(n), np.zeros(6 * n)
        check(lib().pmx_create_A(res.ctx.handle, diag, off))
        A = O.assemble_A(lay, T) + sp.diags(wref)
        np.testing.assert_allclose(diag, A.diagonal(), rtol=1e-13)
    finally:
        res.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_spe10_preconditioner_parity(precision):
    """Well term + DIAG mass + ICC(0) level 2 + no shift: eigenvalues, PCG
    iteration count and pressure against the oracle."""
    perm, poro = _fields()
    print("spe10 pc parity precision", precision, flush=True)
    prm = S.Spe10Params()
    lay = S.layout(MESH, SD, prm.H)
    M, N, Pz = MESH
    Sw = np.full((Pz, N, M), prm.S_wi)
    kref, T, wref, rref = S.linear_system(prm, lay, list(perm), Sw)
    orc = _oracle(lay, kref, wref)
    x0 = np.zeros_like(rref)
    x_ref, its_ref, _, reason_ref = O.pcg(orc.matvec, rref, orc.apply, rtol=1e-8, x0=x0)
    res = spe10.Reservoir(MESH, perm, poro, SD, precision=precision)
    try:
        res.assemble()
        its, rn = res.solve(rtol=1e-8)
        st = P.PC_get_stats(res.ctx)
        _, x = res.state()
    finally:
        res.close()
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"its {its} vs {its_ref}, x rel err {err:.2e}, lv3 its {st['lv3_iters']}, "
          f"lv2 eig {st['eigen_min_lv2']:.6e} vs {orc.eig_lv2[0]:.6e}", flush=True)
    assert reason_ref == "CONVERGED_RTOL"
    assert abs(its - its_ref) <= 1
    assert err <= 1e-8
    assert st["eigen_min_lv2"] == pytest.approx(orc.eig_lv2[0], rel=1e-6)
    assert st["eigen_min_lv1_range"][0] == pytest.approx(min(e[0] for e in orc.eig_lv1),
                                                         rel=1e-6, abs=1e-12)


def test_icc_two_level_apply_exact_with_oracle_basis():
    """With the oracle's level-2 basis and the 2-level cycle, y = B x of the
    ICC(0) path matches the oracle to rounding: pins the block ICC(0)
    factorisation and sweeps."""
    perm, poro = _fields()
    print("icc 2-level apply", flush=True)
    prm = S.Spe10Params()
    lay = S.layout(MESH, SD, prm.H)
    M, N, Pz = MESH
    Sw = 0.2 + 0.3 * np.random.default_rng(8).random((Pz, N, M))
    kref, T, wref, rref = S.linear_system(prm, lay, list(perm), Sw)
    orc = _oracle(lay, kref, wref, use_2level=True)
    phi, num, ev = O.basis_columns(orc)
    ctx = P.PCCtx()
    P.PC_init(ctx, lay.dom, MESH, spe10.pc_options(SD, extra="-use_2level"))
    try:
        ctx.kappa = np.stack([k.reshape(Pz, N, M) for k in kref])
        P._push_kappa(ctx)
        check(lib().pmx_set_well_term(ctx.handle, wref.ctypes.data_as(C.c_void_p)))
        shell = P.PCShell(ctx)
        P.PC_setup_from_basis(shell, phi, num, ev)
        x = np.random.default_rng(9).standard_normal(lay.mesh[0] * lay.mesh[1] * lay.mesh[2])
        y = P.PC_apply_vec(shell, x)
    finally:
        P.PC_final(ctx)
    y_ref = orc.apply(x)
    e = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    print(f"apply rel err {e:.2e}", flush=True)
    assert e <= 1e-10


def test_transport_matches_oracle():
    perm, poro = _fields()
    print("spe10 transport", flush=True)
    prm = S.Spe10Params()
    lay = S.layout(MESH, SD, prm.H)
    M, N, Pz = MESH
    rng = np.random.default_rng(11)
    Sw = 0.2 + 0.2 * rng.random((Pz, N, M))
    p = 4000.0 + 500.0 * rng.random(M * N * Pz)
    res = spe10.Reservoir(MESH, perm, poro, SD)
    try:
        res.set_state(S_w=Sw, P_=p)
        res.assemble()   # kappa of this saturation, as at the start of an IMPES step
        dt = res.transport(steps=7, DS_max=0.001)
        sw, _ = res.state()
    finally:
        res.close()
    kref = S.kappa_from(prm, list(perm), Sw)
    s_ref, dt_ref = S.transport(prm, lay, list(perm), kref, S.modified_poro(poro), p, Sw, 7, 0.001)
    print(f"dt {dt:.6e} vs {dt_ref:.6e}", flush=True)
    assert dt == pytest.approx(dt_ref, rel=1e-12)
    np.testing.assert_allclose(sw, s_ref.ravel(), rtol=1e-12, atol=1e-13)


def test_impes_loop_matches_oracle():
    """Three IMPES steps (assemble, setup, solve from the previous pressure,
    50 transport substeps) on the reference's debug grid."""
    perm, poro = _fields()
    print("impes 3 steps", flush=True)
    res = spe10.Reservoir(MESH, perm, poro, SD)
    try:
        rows = res.run(max_steps=3, rtol=1e-8)
        sw, p = res.state()
    finally:
        res.close()
    prm = S.Spe10Params()
    lay = S.layout(MESH, SD, prm.H)
    M, N, Pz = MESH
    Sw = np.full((Pz, N, M), prm.S_wi)
    x = np.zeros(M * N * Pz)
    t = 0.0
    for k in range(3):
        kref, T, wref, rref = S.linear_system(prm, lay, list(perm), Sw)
        orc = _oracle(lay, kref, wref)
        x, its, _, reason = O.pcg(orc.matvec, rref, orc.apply, rtol=1e-8, x0=x)
        print(f"step {k}: its {rows[k]['its']} vs {its}, S_w range "
              f"[{rows[k]['S_w_min']:.5f}, {rows[k]['S_w_max']:.5f}]", flush=True)
        assert abs(rows[k]["its"] - its) <= 1
        assert rows[k]["simu_time"] == pytest.approx(t, rel=1e-9, abs=1e-12)
        Sw, dt = S.transport(prm, lay, list(perm), kref, S.modified_poro(poro), x, Sw, 50, 0.001)
        assert rows[k]["delta_t"] == pytest.approx(dt, rel=1e-7)
        t += dt
    assert np.linalg.norm(p - x) / np.linalg.norm(x) <= 1e-8
    np.testing.assert_allclose(sw, Sw.ravel(), rtol=1e-7, atol=1e-9)


def test_impes_outputs_written_by_the_loop(tmp_path):
    """The driver's files (src/spe10.c:203-274) from the device loop: the
    CSV log has one row per step, the snapshots are taken after the solve
    and before the transport of their step."""
    from premixfem_amd import io_vtk

    perm, poro = _fields()
    print("impes outputs", flush=True)
    res = spe10.Reservoir(MESH, perm, poro, SD, precision=32)
    try:
        rows = res.run(max_steps=3, rtol=1e-8, output_dir=str(tmp_path), save_files_freq=2,
                       output_P=True)
        _, p_final = res.state()
    finally:
        res.close()
    assert len(rows) == 3
    lines = open(tmp_path / "P_solver.csv").read().splitlines()
    assert lines[0] == io_vtk.CSV_HEADER.strip() and len(lines) == 4
    assert [int(l.split(",")[6]) for l in lines[1:]] == [r["its"] for r in rows]
    s0, _, name = io_vtk.read_vtr(tmp_path / "S_w0.vtr")
    assert name == "S_w" and s0.shape == (MESH[2], MESH[1], MESH[0])
    np.testing.assert_array_equal(s0, 0.2)               # initial saturation
    p2, _, _ = io_vtk.read_vtr(tmp_path / "P2.vtr")
    np.testing.assert_array_equal(p2.ravel(), p_final)  # the last step's pressure
    series = io_vtk.read_series(tmp_path / "S_w.vtr.series")
    assert [f["name"] for f in series["files"]] == ["S_w0.vtr", "S_w2.vtr"]
    # written with %.9f, as the reference does
    assert series["files"][1]["time"] == pytest.approx(rows[2]["simu_time"], abs=1e-9)
