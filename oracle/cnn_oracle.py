"""CPU oracle of the learned warm-start network (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module; the
product path runs the HIP kernels of ``premixfem_amd/csrc/pmx_cnn.hip``.

The network has no counterpart in the reference (it is an addition of this
implementation: a learned initial block for the per-subdomain eigenproblems
of ``src/PreMixFEM_3D.c:224-421``).  This module restates it so its HIP
forward and backward passes can be checked:

* three 3x3x3 convolutions with zero padding at the patch faces,
  4 -> 16 (+ReLU), 16 -> 16 (+ReLU), 16 -> cout;
* parameters laid out as the library's flat FP32 array: per layer
  ``W[co][tap][ci]`` (16 x 27 x 16, tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) then
  ``bias[16]``;
* activations ``[brick][z][y][x][16]``;
* MSE loss over the first ``cout`` channels,
  ``L = sum (out - target)^2 / (ncell * cout)``.

``emulate_bf16=True`` rounds weights, activations and back-propagated
gradients to BF16 (round to nearest even) where the HIP kernels do, so the
two agree to FP32-accumulation accuracy; with ``False`` the oracle is plain
FP64 and its gradient is pinned by finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

CH = 16
TAPS = 27
LAYER_W = CH * TAPS * CH
LAYER = LAYER_W + CH
NPARAMS = 3 * LAYER


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out = []
    for l in range(3):
        W = p[l * LAYER:l * LAYER + LAYER_W].reshape(CH, TAPS, CH)
        b = p[l * LAYER + LAYER_W:(l + 1) * LAYER]
        out.append((W, b))
    return out


def _offsets():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def conv(x, W):
    """out[b,z,y,x,co] = sum_t sum_ci W[co,t,ci] x[b, z+dz, y+dy, x+dx, ci]."""
    nb, nz, ny, nx, _ = x.shape
    xp = np.zeros((nb, nz + 2, ny + 2, nx + 2, CH))
    xp[:, 1:-1, 1:-1, 1:-1, :] = x
    out = np.zeros((nb, nz, ny, nx, CH))
    for t, (dz, dy, dx) in enumerate(_offsets()):
        sl = xp[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :]
        out += sl @ W[:, t, :].T
    return out


def conv_data_grad(g, W):
    """gin[b,v,ci] = sum_t sum_co W[co,t,ci] g[b, v - off_t, co]."""
    nb, nz, ny, nx, _ = g.shape
    gp = np.zeros((nb, nz + 2, ny + 2, nx + 2, CH))
    gp[:, 1:-1, 1:-1, 1:-1, :] = g
    out = np.zeros((nb, nz, ny, nx, CH))
    for t, (dz, dy, dx) in enumerate(_offsets()):
        sl = gp[:, 1 - dz:1 - dz + nz, 1 - dy:1 - dy + ny, 1 - dx:1 - dx + nx, :]
        out += sl @ W[:, t, :]
    return out


def conv_weight_grad(x, g):
    """dW[co,t,ci] = sum_v g[v,co] x[v + off_t, ci]."""
    nb, nz, ny, nx, _ = x.shape
    xp = np.zeros((nb, nz + 2, ny + 2, nx + 2, CH))
    xp[:, 1:-1, 1:-1, 1:-1, :] = x
    dW = np.zeros((CH, TAPS, CH))
    g2 = g.reshape(-1, CH)
    for t, (dz, dy, dx) in enumerate(_offsets()):
        sl = xp[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx, :].reshape(-1, CH)
        dW[:, t, :] = g2.T @ sl
    return dW


def forward(params, x, emulate_bf16=False):
    """Returns (a1, a2, out): hidden activations and the FP64 output."""
    (W1, b1), (W2, b2), (W3, b3) = unpack(params)
    r = bf16 if emulate_bf16 else (lambda v: v)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    x = r(x)
    a1 = r(np.maximum(conv(x, W1) + b1, 0.0))
    a2 = r(np.maximum(conv(a1, W2) + b2, 0.0))
    out = conv(a2, W3) + b3
    return a1, a2, out


def loss_grad(params, x, target, cout, emulate_bf16=False):
    """MSE loss over the first cout channels and its gradient (flat array)."""
    r = bf16 if emulate_bf16 else (lambda v: v)
    (W1, b1), (W2, b2), (W3, b3) = unpack(params)
    W2r, W3r = r(W2), r(W3)
    x = r(np.asarray(x, dtype=np.float64))
    a1, a2, out = forward(params, x, emulate_bf16)
    ncell = out.shape[0] * out.shape[1] * out.shape[2] * out.shape[3]
    scale = 1.0 / (ncell * cout)
    d = np.zeros_like(out)
    d[..., :cout] = out[..., :cout] - target[..., :cout]
    loss = scale * float(np.sum(d * d))
    g3 = r(2.0 * scale * d)
    grad = np.zeros(NPARAMS)

    def put(l, dW, db):
        grad[l * LAYER:l * LAYER + LAYER_W] = dW.ravel()
        grad[l * LAYER + LAYER_W:(l + 1) * LAYER] = db

    put(2, conv_weight_grad(a2, g3), g3.reshape(-1, CH).sum(axis=0))
    g2 = r(conv_data_grad(g3, W3r) * (a2 > 0))
    put(1, conv_weight_grad(a1, g2), g2.reshape(-1, CH).sum(axis=0))
    g1 = r(conv_data_grad(g2, W2r) * (a1 > 0))
    put(0, conv_weight_grad(x, g1), g1.reshape(-1, CH).sum(axis=0))
    return loss, grad


def loss_grad_subspace(params, x, target, cout, emulate_bf16=False):
    """Subspace loss (sign and rotation invariant) of the first cout output
    channels against targets with orthogonal unit-RMS columns, per brick:
    L_b = cout - tr(C^T (G + d I)^{-1} C), G = Y^T Y, C = Y^T U/sqrt(n),
    d = 1e-6 n; total sum_b L_b / (nb * cout).  Returns (loss, grad)."""
    r = bf16 if emulate_bf16 else (lambda v: v)
    (W1, b1), (W2, b2), (W3, b3) = unpack(params)
    W2r, W3r = r(W2), r(W3)
    x = r(np.asarray(x, dtype=np.float64))
    a1, a2, out = forward(params, x, emulate_bf16)
    nb = out.shape[0]
    n = out.shape[1] * out.shape[2] * out.shape[3]
    lscale = 1.0 / (nb * cout)
    loss = 0.0
    g3 = np.zeros_like(out)
    for b in range(nb):
        Y = out[b].reshape(n, CH)[:, :cout]
        U = target[b].reshape(n, CH)[:, :cout] / np.sqrt(n)
        G = Y.T @ Y + 1e-6 * n * np.eye(cout)
        Cm = Y.T @ U
        Gi = np.linalg.inv(G)
        loss += cout - np.trace(Cm.T @ Gi @ Cm)
        K = Gi @ Cm @ Cm.T @ Gi
        dY = -2.0 * (U @ Cm.T @ Gi - Y @ K)
        g3[b].reshape(n, CH)[:, :cout] = dY * lscale
    loss *= lscale
    g3 = r(g3)
    grad = np.zeros(NPARAMS)

    def put(l, dW, db):
        grad[l * LAYER:l * LAYER + LAYER_W] = dW.ravel()
        grad[l * LAYER + LAYER_W:(l + 1) * LAYER] = db

    put(2, conv_weight_grad(a2, g3), g3.reshape(-1, CH).sum(axis=0))
    g2 = r(conv_data_grad(g3, W3r) * (a2 > 0))
    put(1, conv_weight_grad(a1, g2), g2.reshape(-1, CH).sum(axis=0))
    g1 = r(conv_data_grad(g2, W2r) * (a1 > 0))
    put(0, conv_weight_grad(x, g1), g1.reshape(-1, CH).sum(axis=0))
    return loss, grad


def live_mask(cout):
    """Parameters the network actually uses (the rest is channel padding)."""
    m = np.zeros(NPARAMS, dtype=bool)
    for l in range(3):
        W = np.zeros((CH, TAPS, CH), dtype=bool)
        cin = 4 if l == 0 else CH
        co = cout if l == 2 else CH
        W[:co, :, :cin] = True
        m[l * LAYER:l * LAYER + LAYER_W] = W.ravel()
        m[l * LAYER + LAYER_W:l * LAYER + LAYER_W + co] = True
    return m
