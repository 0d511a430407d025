"""CPU checks that pin the oracle's implicit (reference Newton) step: the frozen Jacobian against
the reference's own utils::NewtonMatrices (golden products and diagonals), the Newton update
against a direct solve on a problem whose residual is affine, the reference's convergence
test and iterate rotation. No GPU needed."""
import copy

import numpy as np
import pytest

from helpers import nn, orc, rel_err
from test_oracle import load_golden

O = orc.Oracle(nn.Params)


@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_jacobian_matches_reference_newton_matrices(name):
    """utils.h:844-1849: every block K_f applied to two random vectors, and its diagonal. The
    golden state has distinct iterate and previous state, so a Jacobian that read the iterate
    where the reference reads U^n would fail."""
    d, case = load_golden(name)
    for m in range(2):
        out, dg = O.jacobian_apply(case, d["row_ptr"], d["col"], d["phase"], d["Un"],
                                   d["jac_v"][m], float(d["voln1rel"]), want_diag=True)
        for f in range(11):
            assert rel_err(out[f], d["jac_kv"][f, m]) < 1e-13, (f, m)
            assert rel_err(dg[f], d["jac_diag"][f]) < 1e-13, f


def _affine_case(ng=6, S=3, seed=2):
    """Generic chemistry without the dissolution sink: every residual row is affine in the
    iterate, so one exact Newton update lands on the implicit-Euler solution."""
    case = nn.synth.regular(ng, S=S, seed=seed, r_diss=0.0)
    case.params.front_threshold = -1.0  # no flips: the phase stays fixed
    p = case.params
    st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    return case, rp, col


def _dense_jacobian(case, rp, col, phase, Un, voln1rel):
    F = case.params.S + 1
    n = case.n
    K = np.zeros((F, n, n))
    for j in range(n):
        e = np.zeros((F, n))
        e[:, j] = 1.0
        out, _ = O.jacobian_apply(case, rp, col, phase, Un, e, voln1rel)
        K[:, :, j] = out
    return K


def test_implicit_step_affine_problem_matches_direct_solve():
    case, rp, col = _affine_case()
    S = case.params.S
    U0, ph0, sc0, v0 = O.init_state(case)
    st = O.run_implicit(case, rp, col, 1, newton_tol=1e-12, lin_rtol=1e-14, lin_maxit=2000)
    # the iterate starts at U^n with the Dirichlet values, which init_state already wrote
    F0 = O.bc_override(case, U0, O.residuals(case, rp, col, ph0, U0, U0, 0.0))
    K = _dense_jacobian(case, rp, col, ph0, U0, 0.0)
    exact = np.stack([U0[f] - np.linalg.solve(K[f], F0[f]) for f in range(S + 1)])
    assert rel_err(st["U"], exact) < 1e-11
    # the residual of the implicit-Euler equations vanishes at the new state
    F1 = O.bc_override(case, st["U"], O.residuals(case, rp, col, ph0, st["U"], U0, 0.0))
    for f in range(S + 1):
        assert np.abs(F1[f]).max() <= 1e-9 * max(np.abs(F0[f]).max(), 1e-30), f
    # main.cpp:863-911: the test compares ||dU|| with the previous iteration's, so an exact
    # update still takes three iterations (large, ~0, ~0)
    assert st["log"][0]["newton"] == 3 and st["log"][0]["converged"] == 1


def test_implicit_rotation_follows_the_reference():
    """main.cpp:1002-1024 rotates (U^{n-1}, U^n, U^{n+1}): the iterate of the second step starts
    from the zero-initialised U^{n-1} buffer, the third from U^0."""
    case, rp, col = _affine_case(ng=5)
    U0, ph0, sc0, v0 = O.init_state(case)
    s1 = O.run_implicit(case, rp, col, 1, newton_tol=1e-12, lin_rtol=1e-14, lin_maxit=2000)
    assert np.array_equal(s1["Uprev"], U0) and not s1["Uit"].any() and s1["first"] == 0
    s2 = O.run_implicit(case, rp, col, 1, state=s1, newton_tol=1e-12, lin_rtol=1e-14,
                        lin_maxit=2000)
    assert np.array_equal(s2["Uprev"], s1["U"]) and np.array_equal(s2["Uit"], U0)
    assert s2["step"] == 3
    # on an affine problem the start does not change the converged state
    s2b = O.run_implicit(case, rp, col, 1, state=dict(s1, Uit=s1["U"].copy()), newton_tol=1e-12,
                         lin_rtol=1e-14, lin_maxit=2000)
    assert rel_err(s2["U"], s2b["U"]) < 1e-10


def test_implicit_newton_stops_on_the_reference_criterion():
    """A loose tolerance stops early: max_f | ||dU_f|| - ||dU_f||_prev | < tol."""
    case, rp, col = _affine_case(ng=5)
    st = O.run_implicit(case, rp, col, 1, newton_tol=1e300, lin_rtol=1e-14, lin_maxit=2000)
    assert st["log"][0]["newton"] == 1 and st["log"][0]["converged"] == 1
    st = O.run_implicit(case, rp, col, 1, newton_tol=0.0, newton_maxit=4, lin_rtol=1e-14,
                        lin_maxit=2000)
    assert st["log"][0]["newton"] == 4 and st["log"][0]["converged"] == 0


def test_implicit_reference_chemistry_drives_the_residual_down():
    """Reference chemistry on the golden lattice, all liquid: there the reference's diagonal
    terms R'_k are the reactions' derivatives up to the tiny k1b..k6b terms, so modified Newton
    at a tight tolerance must end with a vanishing residual. (At interface solid nodes the
    reference's H+ diagonal carries -(k7b U3n + k8b U5n + k9b U7n) while the residual there has
    no iterate dependence, utils.h:241-253 against 1089-1100; Newton then contracts slowly and
    the reference's ||dU|| test ends the step early.)"""
    d, case = load_golden("reg7")
    c = copy.deepcopy(case)
    c.params.front_threshold = -1.0
    c.params.dt = 1.0
    phase = np.zeros(case.n, dtype=np.int32)
    U = d["Un"].copy()
    U[:10] = np.maximum(U[:10], 1e-6)  # away from the clamp floor
    state = {"U": U, "phase": phase.copy(), "step": 1, "voln1rel": 0.0, "vol0rel": 1.0,
             "solid": 0}
    st = O.run_implicit(c, d["row_ptr"], d["col"], 1, state=state, newton_tol=1e-14,
                        newton_maxit=60, lin_rtol=1e-14, lin_maxit=4000)
    assert st["log"][0]["converged"] == 1
    Ui = st["Uprev"]  # U^n; the iterate started at U^n with the Dirichlet values
    F0 = O.bc_override(c, Ui, O.residuals(c, d["row_ptr"], d["col"], phase, Ui, Ui, 0.0))
    F1 = O.bc_override(c, st["U"], O.residuals(c, d["row_ptr"], d["col"], phase, st["U"], Ui,
                                                0.0))
    for f in range(11):
        assert np.abs(F1[f]).max() <= 1e-10 * np.abs(F0[f]).max(), f
