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


def _warm_start_pair(kind, seeds, train_fields, fixed):
    from premixfem_amd import cnn, problems, runner
    from premixfem_amd import pc as P

    train = [problems.training_problem(kind, 10000 + s) for s in range(train_fields)]
    data = [cnn.sample_problem(p, extra=fixed) for p in train]
    cout = train[0].en_lv1 - 1
    net = cnn.WarmStartNet(cout=cout, seed=1)
    hist = net.train(data, epochs=30, batch=32, lr=2e-3)
    print(kind, "loss", [round(h, 3) for h in hist[::5]], flush=True)
    out = []
    try:
        for seed in seeds:
            test = problems.training_problem(kind, seed)
            base = runner.run(test, extra=fixed, keep_ctx=True)
            warm = runner.run(test, extra=fixed + " -warm_start cnn", net=net, keep_ctx=True)
            _, eb = P.PC_get_basis(base["ctx"], 2)
            _, ew = P.PC_get_basis(warm["ctx"], 2)
            P.PC_final(base["ctx"])
            P.PC_final(warm["ctx"])
            sb, sw = base["stats"], warm["stats"]
            rel = np.abs(ew - eb)[eb > 0] / eb[eb > 0]
            err = np.linalg.norm(base["x"] - warm["x"]) / np.linalg.norm(base["x"])
            print(f"{kind} field {seed}: eigenvalues max rel diff {rel.max():.2e}; PCG its "
                  f"{base['its']} / {warm['its']}; x rel diff {err:.2e}; LOBPCG its random max "
                  f"{sb['eig_iters_max']} mean {sb['eig_iters_mean']:.1f}, cnn max "
                  f"{sw['eig_iters_max']} mean {sw['eig_iters_mean']:.1f}, fallbacks "
                  f"{sw['eig_fallbacks']}", flush=True)
            out.append((base, warm, rel.max(), err))
    finally:
        net.close()
    return hist, out


@pytest.mark.gpu
def test_warm_start_gain_on_held_out_fields():
    """The learned warm start on the fracture family (16^3 subdomains, 8
    eigenvectors, the 256^3 workload's), trained on 16 seeded 64^3 fields
    (1024 eigenvector sets) and set up on two held-out fields with the block
    size fixed (8): the network must lower both the mean and the largest
    LOBPCG iteration count of the subdomain eigenproblems, while the
    eigenvalues, the PCG iteration count (+-1) and the pressure (1e-8) stay
    those of the random start."""
    hist, res = _warm_start_pair("c4", (20000, 20001), 16, "-eig_guard 1 -eig_autotune 0")
    assert hist[-1] < 0.7 * hist[0]
    mean_r = sum(b["stats"]["eig_iters_mean"] for b, *_ in res) / len(res)
    mean_w = sum(w["stats"]["eig_iters_mean"] for _, w, *_ in res) / len(res)
    max_r = max(b["stats"]["eig_iters_max"] for b, *_ in res)
    max_w = max(w["stats"]["eig_iters_max"] for _, w, *_ in res)
    print(f"held-out LOBPCG iterations: mean {mean_r:.1f} -> {mean_w:.1f}, max {max_r} -> "
          f"{max_w}", flush=True)
    for base, warm, rel, err in res:
        assert base["reason"] == warm["reason"] == "CONVERGED_RTOL"
        assert warm["stats"]["eig_block"] == base["stats"]["eig_block"] == 8
        assert rel <= 1e-7
        assert abs(base["its"] - warm["its"]) <= 1
        assert err <= 1e-8
    assert mean_w < mean_r
    assert max_w < max_r


@pytest.mark.gpu
@pytest.mark.xfail(reason="on the pinned 1e5-contrast binary medium (lambda_min(A_c) ~1e-9 of its "
                          "largest eigenvalue) an O(eig-tol) change of the subdomain basis, which "
                          "any change of LOBPCG's start makes, moves the PCG count by a few per "
                          "cent", strict=False)
def test_warm_start_keeps_solve_on_binary_medium():
    """The same pair on the binary box medium (the 128^3 workload's family):
    iterations within 1 and the pressure within 1e-8."""
    hist, res = _warm_start_pair("c2", (20000, 20001), 6, "")
    for base, warm, rel, err in res:
        assert base["reason"] == warm["reason"] == "CONVERGED_RTOL"
        assert rel <= 1e-7
        assert abs(base["its"] - warm["its"]) <= 1
        assert err <= 1e-8
