"""Drive one pressure solve through the reference call sequence
(PC_init -> kappa -> PC_create_A -> PCSHELL setup -> KSPSolve, as in
``src/Expe_scal.c:66-160``) on the GPU, with wall-clock stage times."""
from __future__ import annotations

import time

import numpy as np

from . import pc as P
from .problems import Problem


def options_for(prob: Problem, precision: int = 64, extra: str = "") -> str:
    sdx, sdy, sdz = prob.sd
    s = (f"-sdx {sdx} -sdy {sdy} -sdz {sdz} -en_lv1 {prob.en_lv1} -en_lv2 {prob.en_lv2} "
         f"-eg_op {prob.eg_op} -pc_precision {precision}")
    if prob.pin_cell is not None:
        s += " -pin {},{},{}".format(*prob.pin_cell)
    return (s + " " + extra).strip()


def run(prob: Problem, precision: int = 64, extra: str = "", basis=None, rtol=None,
        keep_ctx=False, sync=None, max_it=10000, net=None):
    """Set up and solve ``prob``; returns a dict with the solution (natural
    order), iteration count, residual, setup / solve seconds and stats."""
    ctx = P.PCCtx()
    P.PC_init(ctx, prob.dom, prob.mesh, options_for(prob, precision, extra))
    ctx.kappa = prob.kappa
    if net is not None:
        net.attach(ctx)
    shell = P.PCShell(ctx)
    try:
        t0 = time.perf_counter()
        if basis is None:
            P.PC_setup(shell)
        else:
            P.PC_setup_from_basis(shell, *basis)
        t1 = time.perf_counter()
        x, its, rnorm, reason = P.KSPSolve(shell, prob.rhs.ravel(),
                                           rtol=prob.rtol if rtol is None else rtol,
                                           max_it=max_it)
        t2 = time.perf_counter()
        st = P.PC_get_stats(ctx)
        out = dict(x=x, its=its, rnorm=rnorm, reason=reason, t_setup=t1 - t0, t_solve=t2 - t1,
                   stats=st)
        if keep_ctx:
            out["ctx"] = ctx
            out["shell"] = shell
        return out
    finally:
        if not keep_ctx:
            P.PC_final(ctx)


def run_batch(probs, precision: int = 64, extra: str = "", net=None, rtol=None,
              keep_ctx=False, max_it=10000):
    """Set up several independent problems with one batched setup call
    (PC_setup_batch: one LOBPCG launch for every subdomain of every problem),
    then solve each.  Returns (list of per-problem dicts as ``run``, setup
    seconds of the whole batch)."""
    ctxs, shells = [], []
    try:
        for prob in probs:
            ctx = P.PCCtx()
            ctxs.append(ctx)
            P.PC_init(ctx, prob.dom, prob.mesh, options_for(prob, precision, extra))
            ctx.kappa = prob.kappa
            if net is not None:
                net.attach(ctx)
            shells.append(P.PCShell(ctx))
        t0 = time.perf_counter()
        P.PC_setup_batch(shells)
        t_setup = time.perf_counter() - t0
        outs = []
        for prob, ctx, shell in zip(probs, ctxs, shells):
            t1 = time.perf_counter()
            x, its, rnorm, reason = P.KSPSolve(shell, prob.rhs.ravel(),
                                               rtol=prob.rtol if rtol is None else rtol,
                                               max_it=max_it)
            out = dict(x=x, its=its, rnorm=rnorm, reason=reason,
                       t_solve=time.perf_counter() - t1, stats=P.PC_get_stats(ctx))
            if keep_ctx:
                out["ctx"] = ctx
                out["shell"] = shell
            outs.append(out)
        return outs, t_setup
    finally:
        if not keep_ctx:
            for ctx in ctxs:
                P.PC_final(ctx)
