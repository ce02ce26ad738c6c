#!/usr/bin/env python3
"""Writes the full-size oracle fixtures of the benchmark workloads.

The FP64 oracle (``oracle/premixfem_oracle.py``, ``oracle/spe10_oracle.py``)
is run once here, on the CPU, at each workload's own size; the GPU tests
then compare the HIP path with these small files instead of re-running an
oracle that takes minutes to an hour per workload.  Usage::

    python tests/golden/make_fixtures.py c1 c2 c3 c4 [--workers 6]

Each ``<name>.npz`` holds data only (no code): the oracle's PCG iteration
count and reason, the pressure at 4096 seeded cells and its norm, the true
relative residual, y = B v at the same cells for a seeded v (the
preconditioner itself at full size), the level-2 eigenvalues of every
subdomain and the level-3 eigenvalues, plus checksums of the generated
permeability so a drift of the field generator is noticed.  ``c2`` also
records the iteration count of the cycle cut to [lv1, lv2, lv1] (equal to
the full cycle in exact arithmetic on one process), which with the full
cycle's count brackets what rounding alone does on that ill-conditioned
system.  ``c3`` holds the first two IMPES steps of the spe10 fork on the
60x220x85 reservoir grid (iterations, time steps, pressure and saturation
samples).
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import premixfem_oracle as O  # noqa: E402
from oracle import spe10_oracle as S  # noqa: E402
from premixfem_amd import problems  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NSAMPLE = 4096

_G = {}


def sample_cells(n, count=NSAMPLE):
    """The fixture's cells: ``count`` distinct cells drawn with seed 12345,
    sorted (the tests recompute them the same way)."""
    rng = np.random.default_rng(12345)
    return np.sort(rng.choice(n, size=min(count, n), replace=False))


def probe_vector(n):
    """The v of the stored y = B v (seed 777)."""
    return np.random.default_rng(777).standard_normal(n)


def _ghep_job(s):
    from threadpoolctl import threadpool_limits

    lay, T, mdiag, wd, k = _G["lay"], _G["T"], _G["mdiag"], _G["wd"], _G["k"]
    M, N, P = lay.mesh
    stx, nx, sty, ny, stz, nz = lay.corners(s)
    gidx = np.arange(M * N * P).reshape(P, N, M)
    cells = gidx[stz:stz + nz, sty:sty + ny, stx:stx + nx].ravel()
    with threadpool_limits(1):   # one BLAS thread per worker
        return O.ghep(O.local_neumann(lay, T, s, wd), mdiag[cells], k)


def parallel_ghep(orc, workers):
    """Every subdomain GHEP of ``orc`` in a process pool (forked, so the
    fields are shared); returns a ghep_fn for ``Oracle.setup`` that hands
    them out in subdomain order."""
    _G.update(lay=orc.lay, T=orc.T, mdiag=O.mass_diag(orc.lay, orc.kappa, orc.T, orc.eg_op, orc.wd),
              wd=orc.wd, k=orc.en_lv1)
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_ghep_job, range(orc.lay.nsub), chunksize=8)
    it = iter(res)
    return lambda A_N, m, k: next(it)


def kappa_checks(kappa, idx):
    k0 = kappa[0].ravel()
    return dict(kappa_logsum=np.array([float(np.log(kappa[d]).sum()) for d in range(3)]),
                kappa_sample=k0[idx])


def collapsed_apply(orc):
    """[lv1, lv2, lv1] with fine-level residuals: the cycle after its coarse
    part is cut to the first level-2 stage."""
    def apply(x):
        x = np.asarray(x, dtype=np.float64).ravel()
        y = orc.apply_lv1(x)
        y = y + orc.apply_lv2(x - orc.matvec(y))
        return y + orc.apply_lv1(x - orc.matvec(y))
    return apply


def fixture_setup_solve(name, workers, collapsed):
    prob = problems.config(name)
    n = prob.ncell
    idx = sample_cells(n)
    t0 = time.perf_counter()
    orc = O.Oracle(prob.dom, prob.mesh, prob.kappa, sd=prob.sd, en_lv1=prob.en_lv1,
                   en_lv2=prob.en_lv2, eg_op=prob.eg_op, pin_cell=prob.pin_cell)
    fn = parallel_ghep(orc, workers) if workers > 1 else None
    orc.setup(fn)
    t_setup = time.perf_counter() - t0
    print(f"{name}: setup {t_setup:.1f}s nc {int(orc.dof_idx[-1])}", flush=True)
    b = prob.rhs.ravel()
    t0 = time.perf_counter()
    x, its, hist, reason = O.pcg(orc.matvec, b, orc.apply, rtol=prob.rtol)
    t_pcg = time.perf_counter() - t0
    res = np.linalg.norm(b - orc.matvec(x)) / np.linalg.norm(b)
    print(f"{name}: PCG {its} its ({reason}) in {t_pcg:.1f}s, true rel residual {res:.3e}",
          flush=True)
    v = probe_vector(n)
    y = orc.apply(v)
    out = dict(name=name, mesh=np.array(prob.mesh), sd=np.array(prob.sd), en_lv1=prob.en_lv1,
               en_lv2=prob.en_lv2, rtol=prob.rtol, its=its, reason=reason,
               rnorm_hist_last=hist[-1], rnorm0=hist[0], true_rel_residual=res,
               cells=idx, x_sample=x[idx], x_norm=np.linalg.norm(x),
               By_sample=y[idx], By_norm=np.linalg.norm(y),
               eig_lv1=np.array([e[:prob.en_lv1] for e in orc.eig_lv1]),
               eig_lv2=np.asarray(orc.eig_lv2), size_coarse=int(orc.dof_idx[-1]),
               t_setup=t_setup, t_pcg=t_pcg, **kappa_checks(prob.kappa, idx))
    if collapsed:
        t0 = time.perf_counter()
        xc, its_c, _, reason_c = O.pcg(orc.matvec, b, collapsed_apply(orc), rtol=prob.rtol)
        print(f"{name}: PCG with the cycle cut to [lv1, lv2, lv1]: {its_c} its ({reason_c}) in "
              f"{time.perf_counter() - t0:.1f}s", flush=True)
        out.update(its_collapsed=its_c, x_collapsed_sample=xc[idx],
                   x_collapsed_norm=np.linalg.norm(xc))
    return out


IMPES_MESH = (60, 220, 85)
IMPES_SD = (10, 20, 17)


def fixture_impes(workers, steps=2):
    """Steps 0..steps-1 of the spe10 fork's IMPES loop (src/spe10.c:197-274)
    on the 60x220x85 reservoir grid: assemble, set up the fork's
    preconditioner (DIAG mass, wells, ICC(0) level 2, LU level 3 without
    shift), PCG from the previous pressure, 50 transport substeps of
    DS_max 0.001 (the first stage of the reference's tables)."""
    perm, poro = problems.reservoir_fields(IMPES_MESH, 3001)
    prm = S.Spe10Params()
    lay = S.layout(IMPES_MESH, IMPES_SD, prm.H)
    M, N, P = IMPES_MESH
    n = M * N * P
    idx = sample_cells(n)
    Sw = np.full((P, N, M), prm.S_wi)
    x = np.zeros(n)
    out = dict(name="c3", mesh=np.array(IMPES_MESH), sd=np.array(IMPES_SD), cells=idx,
               perm_logsum=np.array([float(np.log(perm[d]).sum()) for d in range(3)]),
               poro_sample=poro.ravel()[idx])
    its_l, dt_l, xs, ss, xn, rs = [], [], [], [], [], []
    for k in range(steps):
        t0 = time.perf_counter()
        kref, T, wref, rref = S.linear_system(prm, lay, list(perm), Sw)
        orc = O.Oracle(lay.dom, lay.mesh, kref, sd=IMPES_SD, en_lv1=8, en_lv2=4,
                       eg_op=O.EIG_OP_DIAG, wd=wref, lv2_icc=True, no_shift_A_cc=True)
        fn = parallel_ghep(orc, workers) if workers > 1 else None
        orc.setup(fn)
        t1 = time.perf_counter()
        x, its, _, reason = O.pcg(orc.matvec, rref, orc.apply, rtol=1e-8, x0=x)
        t2 = time.perf_counter()
        res = np.linalg.norm(rref - orc.matvec(x)) / np.linalg.norm(rref)
        Sw, dt = S.transport(prm, lay, list(perm), kref, S.modified_poro(poro), x, Sw, 50, 0.001)
        print(f"c3 step {k}: setup {t1 - t0:.1f}s PCG {its} its ({reason}) {t2 - t1:.1f}s, "
              f"dt {dt:.6e}, S_w [{Sw.min():.5f}, {Sw.max():.5f}]", flush=True)
        its_l.append(its)
        dt_l.append(dt)
        xs.append(x[idx])
        xn.append(np.linalg.norm(x))
        ss.append(Sw.ravel()[idx])
        rs.append(res)
    out.update(its=np.array(its_l), delta_t=np.array(dt_l), p_sample=np.array(xs),
               p_norm=np.array(xn), S_w_sample=np.array(ss), true_rel_residual=np.array(rs))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("names", nargs="+", choices=("c1", "c2", "c3", "c4"))
    ap.add_argument("--workers", type=int, default=6)
    args = ap.parse_args()
    for name in args.names:
        t0 = time.perf_counter()
        if name == "c3":
            out = fixture_impes(args.workers)
        else:
            out = fixture_setup_solve(name, args.workers, collapsed=name in ("c1", "c2"))
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: wrote {path} ({os.path.getsize(path)} bytes) after "
              f"{time.perf_counter() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
