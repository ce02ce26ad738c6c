"""The CNN oracle pinned without the GPU: its gradient against central finite
differences (FP64, no rounding), the data-gradient convolution as the adjoint
of the forward convolution, and the BF16 rounding against exact values."""
from __future__ import annotations

import numpy as np

from oracle import cnn_oracle as N


def _problem(seed=0, nb=2, shape=(3, 4, 5), cout=3):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    x = np.zeros((nb, nz, ny, nx, N.CH))
    x[..., :4] = rng.standard_normal((nb, nz, ny, nx, 4))
    tgt = np.zeros((nb, nz, ny, nx, N.CH))
    tgt[..., :cout] = rng.standard_normal((nb, nz, ny, nx, cout))
    p = rng.standard_normal(N.NPARAMS) * 0.3 * N.live_mask(cout)
    return p, x, tgt, cout


def test_gradient_matches_finite_differences():
    p, x, tgt, cout = _problem()
    loss, g = N.loss_grad(p, x, tgt, cout)
    rng = np.random.default_rng(1)
    live = np.flatnonzero(N.live_mask(cout))
    idx = rng.choice(live, 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (N.loss_grad(pp, x, tgt, cout)[0] - N.loss_grad(pm, x, tgt, cout)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_padding_gradient_is_zero():
    p, x, tgt, cout = _problem(seed=2)
    _, g = N.loss_grad(p, x, tgt, cout)
    assert np.all(g[~N.live_mask(cout)] == 0.0)


def test_data_gradient_is_adjoint_of_conv():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 4, 5, N.CH))
    y = rng.standard_normal((2, 3, 4, 5, N.CH))
    W = rng.standard_normal((N.CH, N.TAPS, N.CH))
    lhs = np.sum(N.conv(x, W) * y)
    rhs = np.sum(x * N.conv_data_grad(y, W))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_conv_identity_tap():
    # a kernel with only the centre tap (13) = identity on the channels
    W = np.zeros((N.CH, N.TAPS, N.CH))
    W[:, 13, :] = np.eye(N.CH)
    x = np.random.default_rng(4).standard_normal((1, 3, 3, 3, N.CH))
    assert np.array_equal(N.conv(x, W), x)
    # the +x tap (14) shifts by one cell, zero at the face
    W[:] = 0
    W[:, 14, :] = np.eye(N.CH)
    y = N.conv(x, W)
    assert np.array_equal(y[..., :-1, :], x[..., 1:, :])
    assert np.all(y[..., -1, :] == 0)


def test_bf16_rounding():
    assert N.bf16(1.0) == 1.0
    # 1 + 2^-8 is exactly half-way between 1 and 1 + 2^-7: ties to even (1.0)
    assert N.bf16(1.0 + 2.0 ** -8) == 1.0
    assert N.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    v = np.random.default_rng(5).standard_normal(1000)
    assert np.max(np.abs(N.bf16(v) - v) / np.abs(v)) <= 2.0 ** -8


def _orthonormal_targets(tgt, cout):
    nb = tgt.shape[0]
    out = np.zeros_like(tgt)
    for b in range(nb):
        T = tgt[b].reshape(-1, N.CH)[:, :cout]
        Q, _ = np.linalg.qr(T)
        out[b].reshape(-1, N.CH)[:, :cout] = Q * np.sqrt(T.shape[0])
    return out


def test_subspace_gradient_matches_finite_differences():
    p, x, tgt, cout = _problem(seed=6)
    tgt = _orthonormal_targets(tgt, cout)
    loss, g = N.loss_grad_subspace(p, x, tgt, cout)
    rng = np.random.default_rng(7)
    idx = rng.choice(np.flatnonzero(N.live_mask(cout)), 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (N.loss_grad_subspace(pp, x, tgt, cout)[0] -
              N.loss_grad_subspace(pm, x, tgt, cout)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_subspace_loss_invariances():
    p, x, tgt, cout = _problem(seed=8)
    tgt = _orthonormal_targets(tgt, cout)
    l0, _ = N.loss_grad_subspace(p, x, tgt, cout)
    # rotating and flipping the target columns leaves the loss unchanged
    R, _ = np.linalg.qr(np.random.default_rng(9).standard_normal((cout, cout)))
    t2 = tgt.copy()
    t2[..., :cout] = tgt[..., :cout] @ R
    l1, _ = N.loss_grad_subspace(p, x, t2, cout)
    assert abs(l0 - l1) <= 1e-10
    assert 0.0 <= l0 <= 1.0
