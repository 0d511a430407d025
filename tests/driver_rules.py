"""States at the thresholds of the step driver's rules, with the outcomes the reference's driver
gives them (main.cpp:805-837 clamp, 915-922 front update, 944-949 voln1rel refresh), for the
explicit step of the oracle (CPU) and of the library (GPU).

The case has every diffusivity and the dissolution rate set to 0, so a step leaves each species
value where it was before the Dirichlet rows and the clamp: every outcome below follows from the
rules alone, and each is chosen so that a plausible misreading fails it:
  * the clamp `U < floor -> floor` holds species 1..9 everywhere and species 10 only at liquid
    nodes (main.cpp:805-837: `U10n1[i] < 1e-8/CSolid && nodePhase[i] == 0`), with the phase
    of the Newton loop, i.e. before the front update of the same step; a value equal to the
    floor is kept; Dirichlet values are clamped too (the first Newton iterate holds them);
  * the front flips solid nodes with `U1n1 <= CSat/CSolid` (main.cpp:915-922: `<=`, the
    clamped value, solid nodes only, Dirichlet nodes included); a liquid node never turns
    solid; one ulp above the threshold stays solid;
  * voln1rel is recomputed only at steps s with s % sub_step_no == 0 (main.cpp:944-949), over
    the nodes that are solid after that step's front update.
"""
import numpy as np

from helpers import nn

FLOOR = 1e-8
THR = 0.25
SUB = 3
VOL0REL = 2.5
VR0 = 0.123


def rules_case(ng=6):
    case = nn.synth.regular(ng, S=11, seed=0)
    p = case.params
    S = p.S
    for k in range(S):
        p.D_liquid[k] = 0.0
        p.D_solid[k] = 0.0
        p.clamp[k] = nn._abi.CLAMP_NONE
    for k in range(9):
        p.clamp[k] = nn._abi.CLAMP_ALWAYS  # species 1..9
    p.clamp[9] = nn._abi.CLAMP_LIQUID  # species 10 (Cl-): liquid nodes only
    p.r_diss = 0.0
    p.clamp_floor = FLOOR
    p.front_threshold = THR
    p.sub_steps = SUB
    return case


def pick_nodes(case, phase):
    """Interior (non-Dirichlet) solid nodes A, C, D, F and liquid node B; Dirichlet node G."""
    inner = case.bctypes == 1
    solid = np.flatnonzero(inner & (phase == 1))
    liquid = np.flatnonzero(inner & (phase == 0))
    face = np.flatnonzero(case.bctypes == 0)
    assert solid.size >= 5 and liquid.size >= 1 and face.size >= 1
    return dict(A=solid[0], C=solid[1], D=solid[2], F=solid[3], E=solid[4], B=liquid[0],
                G=face[0])


def initial_state(case):
    """Start state: icvalues, phase = (Mg > 0), the special values at the picked nodes."""
    p = case.params
    S = p.S
    U = case.icvalues.copy()
    ph = (U[0] > 0).astype(np.int32)
    U[:, case.bctypes == 0] = np.where(case.bcvalues[:, case.bctypes == 0] != -1,
                                       case.bcvalues[:, case.bctypes == 0], 0.0)
    U[:S, case.bctypes == 0] /= p.CSolid
    nd = pick_nodes(case, ph)
    U[0, (ph == 0) & (case.bctypes == 1)] = 1e-3  # liquid Mg above the floor (and never flips)
    low = 0.5 * FLOOR
    U[:S, nd["A"]] = low                  # solid, every species below the floor
    U[:S, nd["B"]] = low                  # liquid, every species below the floor
    U[0, nd["C"]] = THR                   # solid, Mg exactly at the threshold: flips
    U[0, nd["D"]] = np.nextafter(THR, 1.0)  # one ulp above: stays solid
    U[0, nd["E"]] = FLOOR                 # solid, Mg equal to the floor: kept, flips
    ph[nd["G"]] = 1                       # a solid Dirichlet node whose Mg value is below
    case.bcvalues[0, nd["G"]] = low * p.CSolid  # the floor: Dirichlet, clamp, then flip
    return U, ph, nd


def check_step1(case, U0, ph0, nd, U1, ph1, vr1):
    """Outcome of step s = 1 from (U0, ph0)."""
    p = case.params
    S = p.S
    low = 0.5 * FLOOR
    A, B, C, D, E, G = (nd[k] for k in "ABCDEG")
    # clamp: species 1..9 at A (solid) and B (liquid); species 10 at B only; species 11 never
    assert np.all(U1[:9, A] == FLOOR) and np.all(U1[:9, B] == FLOOR)
    assert U1[9, A] == low, "species 10 is not clamped at a node solid during the step"
    assert U1[9, B] == FLOOR
    assert U1[10, A] == low and U1[10, B] == low
    assert U1[0, E] == FLOOR
    # Dirichlet value below the floor: clamped
    assert U1[0, G] == FLOOR
    # every other species value unchanged (no flux, no source)
    keep = np.ones(U0[:S].shape, dtype=bool)
    keep[:, [A, B]] = False
    keep[0, G] = False
    assert np.array_equal(U1[:S][keep], U0[:S][keep])
    # front: A (clamped 1e-8), C (== threshold), E, G flip; D (one ulp above) and B stay
    flipped = np.flatnonzero((ph0 == 1) & (ph1 == 0))
    assert sorted(flipped.tolist()) == sorted([A, C, E, G]), flipped
    assert ph1[D] == 1 and ph1[B] == 0
    assert np.all(ph1[ph0 == 0] == 0), "a liquid node turned solid"
    # s = 1: no voln1rel refresh
    assert vr1 == VR0


def expected_voln1rel(case, U, ph):
    """Hand sum over the solid nodes: sum (U_Mg - CSat/CSolid) / vol0rel."""
    p = case.params
    idx = np.flatnonzero(ph == 1)
    return float(sum(U[0, i] - p.CSat / p.CSolid for i in idx)) / VOL0REL


def close(a, b):
    return abs(a - b) <= 1e-13 * max(1.0, abs(b))


def run_scenario(stepper):
    """stepper(U, ph, step, voln1rel) -> (U, ph, step, voln1rel) after one explicit step.

    Runs steps s = 1..6 with edits between steps and checks every outcome; returns the
    trajectory of voln1rel."""
    case = rules_case()
    U, ph, nd = initial_state(case)
    U0, ph0 = U.copy(), ph.copy()
    U, ph, s, vr = stepper(case, U, ph, 1, VR0)
    assert s == 2
    check_step1(case, U0, ph0, nd, U, ph, vr)
    vrs = [vr]
    U, ph, s, vr = stepper(case, U, ph, s, vr)  # s = 2: no refresh
    assert vr == VR0
    vrs.append(vr)
    # before s = 3: D drops to the threshold, so it flips in the refresh step itself
    U[0, nd["D"]] = THR
    solid_before = ph.copy()
    U, ph, s, vr = stepper(case, U, ph, s, vr)  # s = 3: refresh with the post-flip phases
    assert ph[nd["D"]] == 0 and solid_before[nd["D"]] == 1
    assert close(vr, expected_voln1rel(case, U, ph))
    assert not close(vr, expected_voln1rel(case, U, solid_before)), "pre-flip phases used"
    vrs.append(vr)
    vr3 = vr
    # before s = 4: raise F's Mg; s = 4 and 5 do not see it, s = 6 does
    U[0, nd["F"]] = 0.6
    for _ in range(2):
        U, ph, s, vr = stepper(case, U, ph, s, vr)
        assert vr == vr3
        vrs.append(vr)
    U, ph, s, vr = stepper(case, U, ph, s, vr)  # s = 6
    assert s == 7
    assert close(vr, expected_voln1rel(case, U, ph)) and not close(vr, vr3)
    vrs.append(vr)
    return vrs


def run_clamp_before_front(stepper):
    """The front tests the clamped value: with the threshold below the floor, a solid node
    whose Mg is below both is clamped to the floor and stays solid."""
    case = rules_case()
    case.params.front_threshold = 0.9 * FLOOR
    U, ph, nd = initial_state(case)
    case.bcvalues[0, nd["G"]] = 1.0 * case.params.CSolid
    U[0, nd["G"]] = 1.0
    U, ph2, s, vr = stepper(case, U, ph, 1, VR0)
    assert U[0, nd["A"]] == FLOOR and ph2[nd["A"]] == 1
    assert ph2[nd["E"]] == 1  # equal to the floor, above the threshold
