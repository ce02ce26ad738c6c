"""The learned warm start on the GPU: the HIP convolution network's forward
pass, loss and gradient against the CPU oracle (BF16-emulating), training
lowering the loss, and the warm start inside pmx_setup leaving the basis and
the PCG result unchanged while reporting its effect on the eigensolver."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cnn_oracle as N

torch = pytest.importorskip("torch")


def _bits(x):
    """float array -> BF16 bit patterns (int16), via the oracle's rounding."""
    f = N.bf16(x).astype(np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def _random_case(seed, nb, shape, cout):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.zeros((nb, nz, ny, nx, N.CH))
    x[..., :4] = rng.standard_normal((nb, nz, ny, nx, 4))
    tgt = np.zeros((nb, nz, ny, nx, N.CH))
    tgt[..., :cout] = rng.standard_normal((nb, nz, ny, nx, cout))
    p = (rng.standard_normal(N.NPARAMS) * 0.25 * N.live_mask(cout)).astype(np.float32)
    return p, x, tgt


@pytest.mark.gpu
@pytest.mark.parametrize("nb,shape,cout", [(3, (16, 16, 16), 5), (2, (6, 11, 5), 7),
                                           (2, (20, 9, 4), 3)])
def test_network_matches_oracle(nb, shape, cout):
    from premixfem_amd.cnn import WarmStartNet

    print("cnn case", nb, shape, cout, flush=True)
    p, x, tgt = _random_case(7, nb, shape, cout)
    dev = torch.device("cuda:0")
    xin = torch.from_numpy(_bits(x).ravel()).to(dev)
    tt = torch.from_numpy(tgt.astype(np.float32).ravel()).to(dev)
    net = WarmStartNet(cout, params=p)
    out = net.predict(xin, nb, shape).cpu().numpy().reshape(x.shape)
    loss = net.loss_grad(xin, tt, nb, shape, loss="mse")
    g = net.grad().astype(np.float64)
    _, _, out_ref = N.forward(p.astype(np.float64), x, emulate_bf16=True)
    loss_ref, g_ref = N.loss_grad(p.astype(np.float64), x, tgt, cout, emulate_bf16=True)
    # subspace loss against orthogonal unit-RMS targets
    tq = np.zeros_like(tgt)
    for b in range(nb):
        Q, _ = np.linalg.qr(tgt[b].reshape(-1, N.CH)[:, :cout])
        tq[b].reshape(-1, N.CH)[:, :cout] = Q * np.sqrt(Q.shape[0])
    ttq = torch.from_numpy(tq.astype(np.float32).ravel()).to(dev)
    lsub = net.loss_grad(xin, ttq, nb, shape, loss="subspace")
    gsub = net.grad().astype(np.float64)
    net.close()
    lsub_ref, gsub_ref = N.loss_grad_subspace(p.astype(np.float64), x, tq, cout,
                                              emulate_bf16=True)
    live = N.live_mask(cout)
    e_out = np.linalg.norm(out - out_ref) / np.linalg.norm(out_ref)
    e_loss = abs(loss - loss_ref) / loss_ref
    e_g = np.linalg.norm(g[live] - g_ref[live]) / np.linalg.norm(g_ref[live])
    e_ls = abs(lsub - lsub_ref) / lsub_ref
    e_gs = np.linalg.norm(gsub[live] - gsub_ref[live]) / np.linalg.norm(gsub_ref[live])
    print(f"rel err out {e_out:.2e} loss {e_loss:.2e} grad {e_g:.2e} | subspace loss "
          f"{e_ls:.2e} grad {e_gs:.2e}", flush=True)
    assert e_out <= 2e-3
    assert e_loss <= 2e-3
    assert e_g <= 2e-2
    assert e_ls <= 2e-3
    assert e_gs <= 2e-2
    assert np.all(g[~live] == 0.0) and np.all(gsub[~live] == 0.0)


@pytest.mark.gpu
def test_training_lowers_loss():
    from premixfem_amd.cnn import Dataset, WarmStartNet

    nb, shape, cout = 16, (8, 8, 8), 3
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    x = np.zeros((nb, nz, ny, nx, N.CH))
    x[..., :4] = rng.standard_normal((nb, nz, ny, nx, 4))
    # a learnable target: smooth functions of the inputs
    tgt = np.zeros_like(x)
    tgt[..., 0] = np.tanh(x[..., 0] + 0.5 * x[..., 1])
    tgt[..., 1] = x[..., 2] * x[..., 3]
    tgt[..., 2] = np.maximum(x[..., 1], 0.0)
    dev = torch.device("cuda:0")
    d = Dataset(torch.from_numpy(_bits(x).ravel()).to(dev),
                torch.from_numpy(tgt.astype(np.float32).ravel()).to(dev), nb, shape, cout)
    net = WarmStartNet(cout, seed=3)
    hist = net.train([d], epochs=40, batch=8, lr=3e-3, loss="mse")
    net.close()
    print("loss history", [round(h, 4) for h in hist[::5]], flush=True)
    assert hist[-1] < 0.5 * hist[0]


@pytest.mark.gpu
def test_warm_start_keeps_solution_and_reports_gain():
    """Train on the eigenvectors of seeded binary-medium fields, then set up
    held-out fields with and without the warm start.  The warm start may
    change only what the eigensolver's start changes: the eigenvalues agree
    to the eigen-tolerance, the PCG converges to the same solution (its
    iteration count on these 1e5-contrast media moves by a few per cent with
    any O(tol) change of the basis, e.g. another random seed).  The
    eigensolver's iteration counts with and without the network are
    reported."""
    from premixfem_amd import cnn, problems, runner
    from premixfem_amd import pc as P

    train = [problems.training_problem("c2", 10000 + s) for s in range(6)]
    data = [cnn.sample_problem(p) for p in train]
    net = cnn.WarmStartNet(cout=5, seed=1)
    hist = net.train(data, epochs=30, batch=32, lr=2e-3)
    print("loss", [round(h, 3) for h in hist[::5]], flush=True)
    assert hist[-1] < 0.8 * hist[0]
    tot = {"random": [0, 0], "cnn": [0, 0]}
    for seed in (20000, 20001):
        test = problems.training_problem("c2", seed)
        base = runner.run(test, keep_ctx=True)
        warm = runner.run(test, extra="-warm_start cnn", net=net, keep_ctx=True)
        _, eb = P.PC_get_basis(base["ctx"], 2)
        _, ew = P.PC_get_basis(warm["ctx"], 2)
        P.PC_final(base["ctx"])
        P.PC_final(warm["ctx"])
        sb, sw = base["stats"], warm["stats"]
        rel = np.abs(ew - eb)[eb > 0] / eb[eb > 0]
        err = np.linalg.norm(base["x"] - warm["x"]) / np.linalg.norm(base["x"])
        print(f"field {seed}: eigenvalues max rel diff {rel.max():.2e}; PCG its "
              f"{base['its']} / {warm['its']}; x rel diff {err:.2e}; LOBPCG its random max "
              f"{sb['eig_iters_max']} mean {sb['eig_iters_mean']:.1f}, cnn max "
              f"{sw['eig_iters_max']} mean {sw['eig_iters_mean']:.1f}, fallbacks "
              f"{sw['eig_fallbacks']}", flush=True)
        assert base["reason"] == warm["reason"] == "CONVERGED_RTOL"
        assert rel.max() <= 1e-7
        assert abs(base["its"] - warm["its"]) <= 0.05 * base["its"]
        assert err <= 1e-5
        tot["random"][0] += sb["eig_iters_mean"]
        tot["cnn"][0] += sw["eig_iters_mean"]
    net.close()
    print("mean LOBPCG iterations over the held-out fields: random "
          f"{tot['random'][0] / 2:.1f}, cnn {tot['cnn'][0] / 2:.1f}", flush=True)
