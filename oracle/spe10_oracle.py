"""CPU oracle of the spe10 fork's reservoir loop (IMPES): total-mobility
permeability, wells, pressure system and explicit upwind saturation update.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` (and the ``cpu_baseline`` leg of
``bench.py``) may import this module; the product path runs the HIP kernels
of ``premixfem_amd/csrc/pmx_spe10.hip``.

Restated from ``src/spe10.c``:

* ``Spe10Params``           -- ``SPE10_init`` (:345-405): fluid constants,
                               cell sizes, well index, injection rate
* ``mobility``              -- ``get_mobi`` (:802-814), Corey exponent 2
* ``kappa_from``            -- ``INTER_upd_kappa`` (:620-648): perm * lambda_t(S_w)
* ``linear_system``         -- ``get_linear_system`` (:723-800): TPFA operator
                               + Peaceman producers (diagonal) and the
                               injector / producer right-hand side
* ``transport``             -- ``get_next_S_w_multisteps`` (:1942-2063)
* ``modified_poro``         -- ``get_modified_poro`` (:2084-2101)

Arrays are in the natural layout arr[ez][ey][ex] (numpy shape (P, N, M)).
The rock data of the reference (``data/spe10/SPE10_rock.h5``) is not
available; callers pass synthetic perm / poro fields of the same shape.
"""
from __future__ import annotations

import math

import numpy as np

from .premixfem_oracle import Layout, transmissibilities

PORO_MIN = 0.05
S_W_BOUND_MIN, S_W_BOUND_MAX = 0.2, 0.8
MAX_TIMESTEP = 200.0
NIL = 1.0e-12
CP_TO_PSI_DAY = 1.6786805555555556e-12


class Spe10Params:
    """``SPE10_init`` constants (:345-405)."""

    def __init__(self, H=(20.0, 10.0, 2.0)):
        self.H = tuple(float(h) for h in H)
        self.mu_w = 0.3 * CP_TO_PSI_DAY
        self.mu_o = 3.0 * CP_TO_PSI_DAY
        self.B_w = 1.01
        self.injection_rate = 5.0e3 * 5.61458 * self.B_w
        self.bottomhole_pressure = 4.0e3
        self.S_wc = 0.2
        self.S_or = 0.2
        self.S_wi = 0.2
        self.wellbore_radius = 1.0
        self.well_index = well_index(self.H[0], self.H[1], self.wellbore_radius)


def well_index(Hx, Hy, rw):
    """Peaceman-type well index of :399-405."""
    r1 = math.sqrt(9.0 * Hx * Hx + Hy * Hy) / 2.0
    r2 = math.sqrt(Hx * Hx + 9.0 * Hy * Hy) / 2.0
    t = Hy / Hx * math.log(r1 / rw) + Hx / Hy * math.log(r2 / rw)
    t = t * (2.0 / math.pi) - 1.0
    return (1.0 / Hx / Hx + 1.0 / Hy / Hy) / t


def mobility(prm: Spe10Params, S):
    """(lambda_w, lambda_o, lambda_t) of get_mobi (:802-814)."""
    s = (S - prm.S_wc) / (1.0 - prm.S_wc - prm.S_or)
    krw = s * s
    kro = (1.0 - s) * (1.0 - s)
    return krw / prm.mu_w, kro / prm.mu_o, krw / prm.mu_w + kro / prm.mu_o


def producer_mask(mesh):
    """The four corner columns (ex, ey) in {0, M-1} x {0, N-1} (:780)."""
    M, N, P = mesh
    m = np.zeros((P, N, M), dtype=bool)
    for x in (0, M - 1):
        for y in (0, N - 1):
            m[:, y, x] = True
    return m


def injector_mask(mesh):
    """The four centre columns (M/2-1 .. M/2) x (N/2-1 .. N/2) (:788)."""
    M, N, P = mesh
    m = np.zeros((P, N, M), dtype=bool)
    for x in (M // 2 - 1, M // 2):
        for y in (N // 2 - 1, N // 2):
            m[:, y, x] = True
    return m


def layout(mesh, sd, H):
    M, N, P = mesh
    return Layout((M * H[0], N * H[1], P * H[2]), mesh, sd)


def kappa_from(prm: Spe10Params, perm, S):
    """kappa_d = perm_d * lambda_t(S_w) (:620-648)."""
    lam = mobility(prm, S)[2]
    return [np.asarray(p) * lam for p in perm]


def linear_system(prm: Spe10Params, lay: Layout, perm, S):
    """get_linear_system (:723-800): face coefficients T of kappa, the well
    diagonal wd (|E| perm_x lambda_t WI at the producers) and the rhs
    (wd * BHP at the producers, q_inj / (4 P) at the injector cells)."""
    kappa = kappa_from(prm, perm, S)
    T = transmissibilities(lay, kappa)
    lam = mobility(prm, S)[2]
    prod = producer_mask(lay.mesh)
    inj = injector_mask(lay.mesh)
    P = lay.mesh[2]
    wd = np.where(prod, lay.meas_elem * perm[0] * lam * prm.well_index, 0.0)
    rhs = np.where(prod, wd * prm.bottomhole_pressure, 0.0)
    rhs = rhs + np.where(inj, prm.injection_rate / (4 * P), 0.0)
    return kappa, T, wd.ravel(), rhs.ravel()


def modified_poro(poro):
    return np.maximum(np.asarray(poro, dtype=np.float64), PORO_MIN)


def saturation_rate(prm: Spe10Params, lay: Layout, perm, kappa, poro, p, S):
    """One evaluation of the saturation right-hand side of
    get_next_S_w_multisteps (:1960-2040): injection, producer outflow and the
    upwind fractional-flow face fluxes, divided by the porosity."""
    M, N, P = lay.mesh
    p = np.asarray(p).reshape(P, N, M)
    lw, _, lt = mobility(prm, S)
    fw = lw / lt
    H = (lay.Hx, lay.Hy, lay.Hz)
    r = np.zeros((P, N, M))
    r += np.where(injector_mask(lay.mesh), prm.injection_rate / (4 * P) / lay.meas_elem, 0.0)
    r += np.where(producer_mask(lay.mesh),
                  perm[0] * lw * prm.well_index * (prm.bottomhole_pressure - p), 0.0)
    for axis, k in ((2, kappa[0]), (1, kappa[1]), (0, kappa[2])):
        h = H[2 - axis]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(None, -1)
        hi[axis] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        kb = 2.0 / (1.0 / k[lo] + 1.0 / k[hi])
        # Face between a = lo cell and b = hi cell; the flux into a is
        # f_up * kb * (p_b - p_a) / h^2 with the upwind cell of the pair.
        dp = p[hi] - p[lo]
        f_a = np.where(p[lo] > p[hi], fw[lo], fw[hi])   # :1991-1994 seen from a
        f_b = np.where(p[hi] > p[lo], fw[hi], fw[lo])   # seen from b
        r[lo] += f_a * kb * dp / (h * h)
        r[hi] += f_b * kb * (-dp) / (h * h)
    poro = np.asarray(poro)
    return np.where(poro > 0.0, r / np.where(poro > 0.0, poro, 1.0), 0.0)


def transport(prm: Spe10Params, lay: Layout, perm, kappa, poro, p, S, steps, DS_max):
    """get_next_S_w_multisteps: ``steps`` explicit substeps with
    dt = DS_max / max(rate); returns (S_new, total dt)."""
    S = np.asarray(S, dtype=np.float64).copy()
    dt_tot = 0.0
    for _ in range(steps):
        r = saturation_rate(prm, lay, perm, kappa, poro, p, S)
        rmax = float(r.max())
        if rmax < 0.0:
            raise RuntimeError("Water saturation is decreasing everywhere")
        S = S + (DS_max / rmax) * r
        dt_tot += DS_max / rmax
    return S, dt_tot
