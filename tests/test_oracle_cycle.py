"""Pins of the oracle's cycle composition (CPU only).

``Oracle.apply`` runs PC_apply_vec's stages (src/PreMixFEM_3D.c:1840-1936),
each y += B_k (x - A y).  After the stages the error propagates as
E = prod_k (I - B_k A) (last stage leftmost), so the preconditioner is
B = (I - E) A^{-1}, i.e. B A v = v - E v for every v.  The stage operators
are built here densely and independently of ``Oracle.apply``'s loop:

* B1 = the level-1 block ICC(0) solve (columns of the identity through the
  DIC factor, which ``test_oracle.py`` pins against IC(0) on the pattern);
* B2 = Phi A_c^{-1} Phi^T with A_c = Phi^T A Phi formed as a dense product,
  with the couplings between process boxes dropped (_PC_setup_lv2_J,
  :806-888);
* B3 = Phi_cc (Phi_cc^T A Phi_cc + NIL I)^{-1} Phi_cc^T (:1177-1390,
  :1597-1648), or Phi (A_c + NIL I)^{-1} Phi^T for the 2-level variant
  (:423-735, :1481-1537).

On process boxes level 2 is box-local, so a dropped level-3 stage, a
missing W-cycle repetition or a wrong smoothing count changes E (checked
below by planting such mistakes in the oracle).  Stage products that are
projections (I - B A with an exact A-orthogonal B) are idempotent, so a
repeated exact stage is the same operator in exact arithmetic.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import premixfem_oracle as O
from premixfem_amd import problems

MESH = (10, 8, 6)


def _orc(procs=(2, 1, 1), **kw):
    k = problems.grf_kappa(MESH, 1.5, 0.2, 29)
    args = dict(sd=2, en_lv1=3, en_lv2=2, pin_cell=(MESH[0] - 1, 0, 0), procs=procs)
    args.update(kw)
    return O.Oracle((1.0, 1.0, 1.0), MESH, k, **args).setup()


def _stage_mats(orc):
    n = orc.A.shape[0]
    A = orc.A.toarray()
    B1 = np.column_stack([orc.dic.solve(e) for e in np.eye(n)])
    Phi = orc.Phi.toarray()
    Ac = Phi.T @ A @ Phi
    box = orc.dof_box
    Ac[box[:, None] != box[None, :]] = 0.0
    if orc.use_2level:
        shift = 0.0 if orc.no_shift_A_cc else O.NIL
        B2 = Phi @ np.linalg.solve(Ac + shift * np.eye(Ac.shape[0]), Phi.T)
        return A, B1, B2, None
    B2 = Phi @ np.linalg.solve(Ac, Phi.T)
    Pcc = np.asarray(orc.Phi_cc)
    Acc = Pcc.T @ A @ Pcc
    if not orc.no_shift_A_cc:
        Acc = Acc + O.NIL * np.eye(Acc.shape[0])
    B3 = Pcc @ np.linalg.solve(Acc, Pcc.T)
    return A, B1, B2, B3


def _sequence(orc, B1, B2, B3):
    """The stage list of PC_apply_vec written out from the reference."""
    s1, s2 = orc.si_lv1, orc.si_lv2
    if orc.use_2level:
        return [B1] * s1 + [B2] + [B1] * s1
    if orc.use_W_cycle:
        return [B1] * s1 + [B2] * s2 + [B3] + [B2] * s2 + [B3] + [B2] * s2 + [B1] * s1
    return [B1] * s1 + [B2] * s2 + [B3] + [B2] * s2 + [B1] * s1


def _propagate(A, seq, v):
    e = v.copy()
    for Bk in seq:
        e = e - Bk @ (A @ e)
    return e


def _check(orc, seq=None, tol=1e-9):
    A, B1, B2, B3 = _stage_mats(orc)
    seq = seq if seq is not None else _sequence(orc, B1, B2, B3)
    rng = np.random.default_rng(5)
    V = rng.standard_normal((3, A.shape[0]))
    worst = 0.0
    for v in V:
        lhs = orc.apply(A @ v)
        rhs = v - _propagate(A, seq, v)
        worst = max(worst, np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs))
    return worst


CASES = [
    ("V", {}),
    ("W", dict(use_W_cycle=True)),
    ("V-si_lv2=2", dict(si_lv2=2)),
    ("W-si_lv2=2", dict(use_W_cycle=True, si_lv2=2)),
    ("V-si_lv1=2", dict(si_lv1=2)),
    ("V-no_shift", dict(no_shift_A_cc=True)),
]


@pytest.mark.parametrize("procs", [(2, 1, 1), (2, 2, 1)], ids=["2x1x1", "2x2x1"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_apply_is_the_product_of_stage_propagators(procs, name, kw):
    orc = _orc(procs=procs, **kw)
    assert _check(orc) <= 1e-9


@pytest.mark.parametrize("kw", [{}, dict(use_W_cycle=True), dict(use_2level=True)],
                         ids=["V", "W", "2-level"])
def test_one_process_apply_is_the_product_of_stage_propagators(kw):
    orc = _orc(procs=(1, 1, 1), **kw)
    assert _check(orc) <= 1e-9


def test_one_process_level3_and_second_level2_vanish():
    """On one process the first level-2 stage leaves Phi^T (x - A y) = 0,
    so the cycle equals [lv1, lv2, lv1] (the reason the GPU cuts the coarse
    program there)."""
    orc = _orc(procs=(1, 1, 1))
    A, B1, B2, B3 = _stage_mats(orc)
    assert _check(orc, seq=[B1, B2, B1]) <= 1e-9


def test_planted_mistakes_are_caught():
    """The check fails for a cycle with the level-3 stage dropped, with one
    W-cycle repetition missing or with one smoothing step too few (the
    mistakes are planted in the stage list, i.e. in what the oracle is
    compared with)."""
    orc = _orc(procs=(2, 2, 1), use_W_cycle=True)
    A, B1, B2, B3 = _stage_mats(orc)
    assert _check(orc) <= 1e-9
    assert _check(orc, seq=[B1, B2, B2, B2, B1]) > 1e-6           # level 3 dropped
    assert _check(orc, seq=[B1, B2, B3, B2, B1]) > 1e-6           # V instead of W
    assert _check(orc, seq=[B1, B2, B3, B2, B3, B2]) > 1e-6       # post-smoothing missing
    orc2 = _orc(procs=(2, 2, 1), si_lv2=2)
    A, B1, B2, B3 = _stage_mats(orc2)
    assert _check(orc2, seq=[B1, B2, B3, B2, B1]) > 1e-6          # si_lv2 ignored


def test_planted_mistake_in_the_oracle_is_caught(monkeypatch):
    """An oracle whose level-3 stage is dropped fails the product check."""
    orc = _orc(procs=(2, 2, 1))
    monkeypatch.setattr(orc, "apply_lv3", lambda r: np.zeros_like(r))
    assert _check(orc) > 1e-6
