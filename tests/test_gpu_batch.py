"""Batched setup of independent problems (pmx_setup_batch): the subdomain
eigenproblems of every problem in one LOBPCG launch.  Each problem must end
up exactly as its own PC_setup (src/PreMixFEM_3D.c:1826-1838) leaves it:
the same level-2 basis and eigenvalues, the same PCG iteration count and
solution; the smallest member is also checked against the oracle."""
import numpy as np
import pytest

from oracle import premixfem_oracle as O
from premixfem_amd import pc as P
from premixfem_amd import problems, runner

pytestmark = pytest.mark.gpu

SIZES = (16, 24, 32)   # 8^3-cell subdomains: 8, 27 and 64 per problem


def _members(n=4):
    return [problems.batch_member(i, sizes=SIZES) for i in range(n)]


def _single(prob, precision, extra, net=None):
    res = runner.run(prob, precision=precision, extra=extra, keep_ctx=True, net=net)
    try:
        phi, ev = P.PC_get_basis(res["ctx"], 2)
    finally:
        P.PC_final(res["ctx"])
    return res, phi, ev


def _compare(probs, precision, extra="", net=None):
    outs, _ = runner.run_batch(probs, precision=precision, extra=extra, keep_ctx=True, net=net)
    try:
        got = []
        for o in outs:
            phi, ev = P.PC_get_basis(o["ctx"], 2)
            got.append((o, phi, ev))
    finally:
        for o in outs:
            P.PC_final(o["ctx"])
    for prob, (o, phi, ev) in zip(probs, got):
        ref, phi_r, ev_r = _single(prob, precision, extra, net)
        e_phi = np.abs(phi - phi_r).max()
        e_x = np.linalg.norm(o["x"] - ref["x"]) / np.linalg.norm(ref["x"])
        print(f"{prob.name} {prob.mesh} p{precision}: its {o['its']} vs {ref['its']}, "
              f"basis max diff {e_phi:.1e}, x rel diff {e_x:.1e}, MB "
              f"{o['stats']['eig_block']}", flush=True)
        assert o["reason"] == "CONVERGED_RTOL"
        assert o["its"] == ref["its"]
        assert o["stats"]["eig_iters_max"] == ref["stats"]["eig_iters_max"]
        assert e_phi <= 1e-12
        np.testing.assert_allclose(ev, ev_r, rtol=1e-12, atol=0.0)
        assert e_x <= 1e-12
    return got


@pytest.mark.parametrize("precision", [64, 32])
def test_batch_equals_individual_setups(precision):
    probs = _members()
    print("batch of", [p.mesh[0] for p in probs], flush=True)
    _compare(probs, precision)


def test_batch_member_matches_oracle():
    probs = _members()
    small = min(probs, key=lambda p: p.ncell)
    outs, _ = runner.run_batch(probs, precision=64)
    o = outs[probs.index(small)]
    orc = O.Oracle(small.dom, small.mesh, small.kappa, sd=small.sd, en_lv1=small.en_lv1,
                   en_lv2=small.en_lv2, pin_cell=small.pin_cell).setup()
    x_ref, its_ref, _, reason = O.pcg(orc.matvec, small.rhs.ravel(), orc.apply, rtol=small.rtol)
    err = np.linalg.norm(o["x"] - x_ref) / np.linalg.norm(x_ref)
    print(f"{small.name} {small.mesh}: its {o['its']} vs oracle {its_ref}, x rel err {err:.1e}",
          flush=True)
    assert reason == "CONVERGED_RTOL"
    assert abs(o["its"] - its_ref) <= 1
    assert err <= 1e-8


def test_batch_with_learned_warm_start():
    """All members seeded by one (untrained) network: the warm-start rows of
    every problem land in its slice of the batch."""
    from premixfem_amd.cnn import WarmStartNet

    probs = _members(3)
    net = WarmStartNet(probs[0].en_lv1 - 1, seed=5)
    try:
        _compare(probs, 32, extra="-warm_start cnn", net=net)
    finally:
        net.close()


def test_batch_rejects_mixed_options():
    probs = _members(2)
    ctxs = []
    try:
        for prob, extra in zip(probs, ("", "-eig_guard 1")):
            ctx = P.PCCtx()
            ctxs.append(ctx)
            P.PC_init(ctx, prob.dom, prob.mesh, runner.options_for(prob, 64, extra))
            ctx.kappa = prob.kappa
        with pytest.raises(P.PMXError, match="eigensolver options"):
            P.PC_setup_batch([P.PCShell(c) for c in ctxs])
    finally:
        for ctx in ctxs:
            P.PC_final(ctx)
