"""Learned warm start for the level-2 eigensolver: data generation and
training driver around the HIP network of ``csrc/pmx_cnn.hip``.

The network is trained on eigenvectors the method itself produces: for each
training field the GPU preconditioner is set up (batched LOBPCG), and every
subdomain contributes one sample (input: its log-permeability patch and cell
coordinates; target: its eigenvectors 1..cout scaled to unit RMS).  Forward,
backward and the Adam update all run in the library's HIP kernels; PyTorch
only holds the device buffers.

Typical use::

    net = WarmStartNet(cout=prob.en_lv1 - 1, seed=0)
    data = [sample_problem(p) for p in training_problems]
    net.train(data, epochs=30)
    ... PC_init(ctx, ..., "-warm_start cnn ..."); net.attach(ctx)
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from . import pc as P
from . import runner
from ._lib import check, lib

CH = 16


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Dataset:
    """Device-resident samples of one field: uniform bricks of nx*ny*nz."""

    def __init__(self, inputs, targets, nb, shape, cout):
        self.inputs = inputs      # torch.int16 [nb*n*16] (BF16 bits)
        self.targets = targets    # torch.float32 [nb*n*16]
        self.nb = nb
        self.shape = shape        # (nx, ny, nz)
        self.cout = cout

    @property
    def ncell_brick(self):
        nx, ny, nz = self.shape
        return nx * ny * nz


def uniform_brick_shape(prob):
    shapes = set()
    for a in range(3):
        n, sd = prob.mesh[a], prob.sd[a]
        shapes.add((a, n // sd, n % sd))
    ext = []
    for a in range(3):
        n, sd = prob.mesh[a], prob.sd[a]
        if n % sd:
            raise ValueError("training needs uniform subdomains (mesh divisible by sd)")
        ext.append(n // sd)
    return tuple(ext)


def sample_problem(prob, precision=64, extra=""):
    """Set up the preconditioner on ``prob`` (GPU LOBPCG) and return its
    subdomains as a :class:`Dataset` (device tensors)."""
    import torch

    shape = uniform_brick_shape(prob)
    cout = prob.en_lv1 - 1
    ctx = P.PCCtx()
    P.PC_init(ctx, prob.dom, prob.mesh, runner.options_for(prob, precision, extra))
    try:
        ctx.kappa = prob.kappa
        P._push_kappa(ctx)
        check(lib().pmx_setup(ctx.handle))
        n = prob.ncell
        dev = torch.device("cuda:0")
        inp = torch.empty(n * CH, dtype=torch.int16, device=dev)
        tgt = torch.empty(n * CH, dtype=torch.float32, device=dev)
        check(lib().pmx_cnn_inputs(ctx.handle, _ptr(inp)))
        check(lib().pmx_cnn_targets(ctx.handle, cout, _ptr(tgt)))
        nb = int(np.prod(prob.sd))
        return Dataset(inp, tgt, nb, shape, cout)
    finally:
        P.PC_final(ctx)


class WarmStartNet:
    """The HIP network (FP32 master weights, BF16 compute)."""

    def __init__(self, cout, seed=0, params=None):
        self.cout = int(cout)
        self.handle = C.c_void_p()
        buf = None
        if params is not None:
            buf = np.ascontiguousarray(params, dtype=np.float32)
            if buf.size != self.num_params():
                raise ValueError("parameter count mismatch")
        check(lib().pmx_cnn_create(C.byref(self.handle), self.cout,
                                   buf.ctypes.data_as(C.c_void_p) if buf is not None else None,
                                   int(seed)))

    @staticmethod
    def num_params():
        return int(lib().pmx_cnn_num_params())

    def params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().pmx_cnn_get_params(self.handle, p.ctypes.data_as(C.c_void_p), None))
        return p

    def grad(self):
        g = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().pmx_cnn_get_params(self.handle, None, g.ctypes.data_as(C.c_void_p)))
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().pmx_cnn_set_params(self.handle, p.ctypes.data_as(C.c_void_p)))

    def loss_grad(self, inputs, targets, nb, shape, loss="subspace"):
        """Forward + backward; loss "subspace" (sign/rotation invariant) or
        "mse"."""
        nx, ny, nz = shape
        out = C.c_double(0.0)
        kind = {"mse": 0, "subspace": 1}[loss]
        check(lib().pmx_cnn_loss_grad(self.handle, nb, nx, ny, nz, _ptr(inputs), _ptr(targets),
                                      kind, C.byref(out)))
        return out.value

    def predict(self, inputs, nb, shape):
        import torch

        nx, ny, nz = shape
        out = torch.empty(nb * nx * ny * nz * CH, dtype=torch.float32, device=inputs.device)
        check(lib().pmx_cnn_predict(self.handle, nb, nx, ny, nz, _ptr(inputs), _ptr(out)))
        return out

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().pmx_cnn_adam_step(self.handle, lr, beta1, beta2, eps))

    def train(self, datasets, epochs=20, batch=64, lr=2e-3, seed=0, log=None, loss="subspace"):
        """Mini-batch Adam over all bricks of ``datasets``; returns the mean
        loss of each epoch."""
        rng = np.random.default_rng(seed)
        batches = []
        for d in datasets:
            n = d.ncell_brick * CH
            for b0 in range(0, d.nb, batch):
                nb = min(batch, d.nb - b0)
                batches.append((d, b0, nb, n))
        hist = []
        for ep in range(epochs):
            t0 = time.perf_counter()
            tot, cnt = 0.0, 0
            for i in rng.permutation(len(batches)):
                d, b0, nb, n = batches[i]
                val = self.loss_grad(d.inputs[b0 * n:(b0 + nb) * n],
                                     d.targets[b0 * n:(b0 + nb) * n], nb, d.shape, loss)
                self.adam_step(lr)
                tot += val * nb
                cnt += nb
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4f} ({time.perf_counter() - t0:.2f}s)")
        return hist

    def attach(self, ctx):
        """Use this network for -warm_start cnn on ``ctx`` (a PCCtx)."""
        check(lib().pmx_set_warm_start(ctx._require(), self.handle))

    def close(self):
        if self.handle:
            lib().pmx_cnn_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
