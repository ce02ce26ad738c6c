"""Seeded synthetic Darcy problems (permeability fields, sources, pin).

The reference fills ``s_ctx.kappa[0..2]`` and the right-hand side in its
drivers (``src/Expe_scal.c:11-64``, ``src/test.c:7-38``); the data files its
other drivers read (fracture masks, SPE10 rock) are not part of the reference
snapshot, so every field here is generated from a seed.

Arrays are in the reference's natural layout ``arr[ez][ey][ex]`` (x fastest),
i.e. numpy shape ``(P, N, M)`` for an ``M x N x P`` cell grid.
"""
from __future__ import annotations

import dataclasses
import numpy as np


@dataclasses.dataclass
class Problem:
    """One pressure solve: grid, permeability, right-hand side and PC options.

    ``kappa`` has shape (3, P, N, M) (x, y, z components).  ``pin_cell`` is a
    natural-order (x, y, z) cell that gets a diagonal term ``tau`` equal to its
    largest face transmissibility (the producer-well mechanism of
    ``src/spe10.c:780-785``); ``None`` keeps the pure-Neumann reference system.
    """
    name: str
    dom: tuple
    mesh: tuple
    kappa: np.ndarray
    rhs: np.ndarray
    sd: tuple = (2, 2, 2)
    en_lv1: int = 4
    en_lv2: int = 4
    eg_op: int = 0
    pin_cell: tuple | None = None
    rtol: float = 1e-8

    @property
    def ncell(self) -> int:
        return int(np.prod(self.mesh))


def _exp_cov_grf(shape, corr, sigma, rng):
    """Gaussian random field with covariance sigma^2 exp(-r/corr).

    FFT circulant embedding on a 2x padded periodic grid; negative embedding
    eigenvalues are clipped to zero.  ``corr`` is in cell units per axis
    (tuple) and ``shape`` is (P, N, M).
    """
    pad = tuple(2 * s for s in shape)
    axes = [np.minimum(np.arange(p), p - np.arange(p)).astype(np.float64) for p in pad]
    dz, dy, dx = np.meshgrid(*axes, indexing="ij")
    cz, cy, cx = corr
    r = np.sqrt((dz / cz) ** 2 + (dy / cy) ** 2 + (dx / cx) ** 2)
    cov = sigma ** 2 * np.exp(-r)
    lam = np.fft.fftn(cov).real
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal(pad) + 1j * rng.standard_normal(pad)
    field = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(np.prod(pad))
    field = field.real[: shape[0], : shape[1], : shape[2]]
    return np.ascontiguousarray(field)


def corner_rhs(mesh, dom):
    """Source +|E| at cell (0,0,0) and -|E| at (M-1,N-1,P-1).

    Scaled by the element measure as the reference scales its well sources
    (``src/Expe_scal.c:56``)."""
    M, N, P = mesh
    E = (dom[0] / M) * (dom[1] / N) * (dom[2] / P)
    b = np.zeros((P, N, M))
    b[0, 0, 0] = E
    b[P - 1, N - 1, M - 1] = -E
    return b


def xxy_rhs(mesh, dom):
    """Reference RHS: +1e3|E| in the four corner columns, -1e3|E| in the four
    centre columns, then normalised (``src/Expe_scal.c:44-64,92``)."""
    M, N, P = mesh
    E = (dom[0] / M) * (dom[1] / N) * (dom[2] / P)
    b = np.zeros((P, N, M))
    for ex in (0, M - 1):
        for ey in (0, N - 1):
            b[:, ey, ex] = 1.0e3 * E
    for ex in (M // 2 - 1, M // 2):
        for ey in (N // 2 - 1, N // 2):
            b[:, ey, ex] = -1.0e3 * E
    return b / np.linalg.norm(b)


def cross_kappa(mesh, cr):
    """Periodic cross channels of contrast 10^cr (``src/Expe_scal.c:11-42``)."""
    M, N, P = mesh
    ex = np.arange(M) % 8
    ey = np.arange(N) % 8
    ez = np.arange(P) % 8
    inx = (ex >= 3) & (ex < 5)
    iny = (ey >= 3) & (ey < 5)
    inz = (ez >= 3) & (ez < 5)
    Z, Y, X = np.meshgrid(inz, iny, inx, indexing="ij")
    chan = (X & Y) | (X & Z) | (Y & Z)
    k = np.where(chan, 10.0 ** cr, 1.0)
    return np.stack([k, k, k])


def laplace_kappa(mesh):
    """Constant anisotropic kappa = (2, 1, 0.5) (``src/test.c:7-16``)."""
    M, N, P = mesh
    one = np.ones((P, N, M))
    return np.stack([2.0 * one, 1.0 * one, 0.5 * one])


def grf_kappa(mesh, sigma, corr_len, seed):
    """Scalar k = exp(g), g a GRF with exponential covariance on the unit cube
    (correlation length in domain units)."""
    M, N, P = mesh
    rng = np.random.default_rng(seed)
    corr = (corr_len * P, corr_len * N, corr_len * M)
    g = _exp_cov_grf((P, N, M), corr, sigma, rng)
    k = np.exp(g)
    return np.stack([k, k, k])


def box_inclusions_kappa(mesh, nbox, side_lo, side_hi, k_in, seed):
    """Background 1 with ``nbox`` axis-aligned boxes of integer sides
    U{side_lo..side_hi} at uniform positions, k = k_in inside."""
    M, N, P = mesh
    rng = np.random.default_rng(seed)
    k = np.ones((P, N, M))
    sides = rng.integers(side_lo, side_hi + 1, size=(nbox, 3))
    for b in range(nbox):
        sx, sy, sz = sides[b]
        x0 = rng.integers(0, M - sx + 1)
        y0 = rng.integers(0, N - sy + 1)
        z0 = rng.integers(0, P - sz + 1)
        k[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = k_in
    return np.stack([k, k.copy(), k.copy()])


def layered_kappa(mesh, sigma, corr_cells, kmin, kmax, seed, kz_ratio=0.1):
    """Layered lognormal field clipped to [kmin, kmax]; kz = kz_ratio * kx."""
    M, N, P = mesh
    rng = np.random.default_rng(seed)
    cx, cy, cz = corr_cells
    g = _exp_cov_grf((P, N, M), (cz, cy, cx), sigma, rng)
    kx = np.clip(np.exp(g), kmin, kmax)
    return np.stack([kx, kx.copy(), kz_ratio * kx])


MD_TO_REF = 9.869233e-16 * 10.7639   # 1 mD in m^2, times the 10.7639 scaling of src/spe10.c:379


def reservoir_fields(mesh, seed, sigma=2.5, corr_cells=(20, 20, 2), kmin=1e-3, kmax=2e4,
                     kz_ratio=0.1):
    """Synthetic rock for the IMPES workload (the reference reads SPE10
    model 2 from ``data/spe10/SPE10_rock.h5``, which is not available):
    layered lognormal kx = ky in mD clipped to [kmin, kmax], kz = kz_ratio kx,
    converted to the reference's units (m^2 times 10.7639); porosity from the
    same field, 0.05 + 0.25 * (rank of log k), so high-permeability layers are
    also porous.  Returns (perm (3, P, N, M), poro (P, N, M))."""
    k = layered_kappa(mesh, sigma, corr_cells, kmin, kmax, seed, kz_ratio)
    lk = np.log(k[0])
    rank = np.argsort(np.argsort(lk.ravel())).reshape(lk.shape) / max(lk.size - 1, 1)
    poro = 0.05 + 0.25 * rank
    return k * MD_TO_REF, poro


def fracture_kappa(mesh, sigma, corr_cells, nfrac, k_frac, seed):
    """Lognormal background plus ``nfrac`` one-cell-thick axis-aligned planes
    spanning 30-100 % of the domain in each in-plane direction."""
    M, N, P = mesh
    rng = np.random.default_rng(seed)
    g = _exp_cov_grf((P, N, M), (corr_cells,) * 3, sigma, rng)
    k = np.exp(g)
    dims = (M, N, P)
    for _ in range(nfrac):
        normal = rng.integers(0, 3)
        pos = rng.integers(0, dims[normal])
        lo, hi = [], []
        for a in range(3):
            if a == normal:
                lo.append(pos)
                hi.append(pos + 1)
            else:
                ext = int(round(rng.uniform(0.3, 1.0) * dims[a]))
                s = rng.integers(0, dims[a] - ext + 1)
                lo.append(s)
                hi.append(s + ext)
        k[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = k_frac
    return np.stack([k, k.copy(), k.copy()])


def config(name: str) -> Problem:
    """The named workloads.

    * ``c1``: 32^3 GRF (sigma 2, corr 0.1), sd 4, k 4;
    * ``c2``: 128^3 binary box medium (3000 boxes, k=1e5), sd 8, k 6;
    * ``c3``: 60x220x85 layered lognormal reservoir grid, sd (10,20,17), k 8;
    * ``c4``: 256^3 lognormal + 64 fractures, sd 16, k 8;
    * ``tiny``/``small``: quick cases for tests.
    All use the corner source/sink and a pin at cell (M-1, 0, 0).
    """
    dom = (1.0, 1.0, 1.0)
    if name == "c1":
        mesh = (32, 32, 32)
        kap = grf_kappa(mesh, 2.0, 0.1, 1001)
        return Problem("c1", dom, mesh, kap, corner_rhs(mesh, dom), sd=(4, 4, 4),
                       en_lv1=4, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "c2":
        mesh = (128, 128, 128)
        kap = box_inclusions_kappa(mesh, 3000, 2, 6, 1.0e5, 2001)
        return Problem("c2", dom, mesh, kap, corner_rhs(mesh, dom), sd=(8, 8, 8),
                       en_lv1=6, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "c3":
        mesh = (60, 220, 85)
        dom = (60 * 20.0, 220 * 10.0, 85 * 2.0)
        kap = layered_kappa(mesh, 2.5, (20, 20, 2), 1e-3, 2e4, 3001)
        return Problem("c3", dom, mesh, kap, corner_rhs(mesh, dom), sd=(10, 20, 17),
                       en_lv1=8, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "c4":
        mesh = (256, 256, 256)
        kap = fracture_kappa(mesh, 1.0, 8.0, 64, 1.0e4, 4001)
        return Problem("c4", dom, mesh, kap, corner_rhs(mesh, dom), sd=(16, 16, 16),
                       en_lv1=8, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "c2_small":
        mesh = (32, 32, 32)
        kap = box_inclusions_kappa(mesh, 47, 2, 6, 1.0e5, 2001)
        return Problem("c2_small", dom, mesh, kap, corner_rhs(mesh, dom), sd=(2, 2, 2),
                       en_lv1=6, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "tiny":
        mesh = (8, 8, 8)
        kap = grf_kappa(mesh, 1.0, 0.2, 7)
        return Problem("tiny", dom, mesh, kap, corner_rhs(mesh, dom), sd=(2, 2, 2),
                       en_lv1=3, en_lv2=2, pin_cell=(mesh[0] - 1, 0, 0))
    if name == "small":
        mesh = (16, 16, 16)
        kap = grf_kappa(mesh, 2.0, 0.1, 11)
        return Problem("small", dom, mesh, kap, corner_rhs(mesh, dom), sd=(2, 2, 2),
                       en_lv1=4, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
    raise KeyError(name)


BATCH_SIZES = (32, 48, 64, 80, 96)


def batch_member(i: int, sizes=BATCH_SIZES, brick: int = 8, en_lv1: int = 6) -> Problem:
    """Member i of the batched workload of independent Darcy problems: a cube
    of n^3 cells with n drawn uniformly from ``sizes``, a lognormal
    permeability exp(g) with sigma drawn uniformly from [1, 3] (correlation
    length 0.1 of the domain), seed 5000 + i, brick^3-cell subdomains,
    en_lv1 eigenvectors, corner source/sink and the pin at (n-1, 0, 0)."""
    rng = np.random.default_rng(5000 + i)
    n = int(rng.choice(np.asarray(sizes)))
    sigma = float(rng.uniform(1.0, 3.0))
    mesh = (n, n, n)
    dom = (1.0, 1.0, 1.0)
    kap = grf_kappa(mesh, sigma, 0.1, 5000 + i)
    sd = (n // brick,) * 3
    return Problem(f"batch-{i}", dom, mesh, kap, corner_rhs(mesh, dom), sd=sd,
                   en_lv1=en_lv1, en_lv2=4, pin_cell=(n - 1, 0, 0))


def batch_config(count: int = 64, sizes=BATCH_SIZES) -> list:
    """The batched workload: ``count`` members (64: sum of cells ~ 23 M)."""
    return [batch_member(i, sizes) for i in range(count)]


def training_problem(kind: str, seed: int, mesh=(64, 64, 64)) -> Problem:
    """A field of the same family as workload ``kind`` ("c2": binary box
    medium, "c4": lognormal + fractures, both with 16^3 subdomains; "batch":
    the batched members' lognormal family with 8^3 subdomains) on ``mesh``,
    for generating warm-start training samples (seeds disjoint from the
    workloads')."""
    dom = (1.0, 1.0, 1.0)
    vol = mesh[0] * mesh[1] * mesh[2]
    sd = tuple(m // 16 for m in mesh)
    if kind == "batch":
        # lognormal, sigma drawn from [1, 3] as in batch_member, 8^3 bricks
        sigma = float(np.random.default_rng(seed).uniform(1.0, 3.0))
        kap = grf_kappa(mesh, sigma, 0.1, seed)
        return Problem(f"batch-train-{seed}", dom, mesh, kap, corner_rhs(mesh, dom),
                       sd=tuple(m // 8 for m in mesh), en_lv1=6, en_lv2=4,
                       pin_cell=(mesh[0] - 1, 0, 0))
    if kind == "c2":
        kap = box_inclusions_kappa(mesh, max(1, round(3000 * vol / 128 ** 3)), 2, 6, 1.0e5, seed)
        en = 6
    elif kind == "c4":
        nfrac = max(1, round(64 * (vol / 256 ** 3) ** (1.0 / 3.0)))
        kap = fracture_kappa(mesh, 1.0, 8.0, nfrac, 1.0e4, seed)
        en = 8
    else:
        raise KeyError(kind)
    return Problem(f"{kind}-train-{seed}", dom, mesh, kap, corner_rhs(mesh, dom), sd=sd,
                   en_lv1=en, en_lv2=4, pin_cell=(mesh[0] - 1, 0, 0))
