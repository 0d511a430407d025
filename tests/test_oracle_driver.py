"""The oracle's explicit step (or_step) against the driver rules of main.cpp at their
thresholds: clamp (805-837), front update (915-922), voln1rel cadence (944-949). The states and
expected outcomes are in driver_rules.py; the GPU runs the same scenario in
test_gpu_driver_rules.py."""
import numpy as np

import driver_rules as R
from helpers import nn, orc

O = orc.Oracle(nn.Params)


def oracle_stepper():
    cache = {}

    def step(case, U, ph, s, vr):
        if "nb" not in cache:
            p = case.params
            st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5),
                                               p.dx * 10., p.dy * 10., p.dz * 10.)
            assert st == 0
            cache["nb"] = (rp, col)
        rp, col = cache["nb"]
        state = {"U": U.copy(), "phase": ph.copy(), "step": s, "voln1rel": vr,
                 "vol0rel": R.VOL0REL, "solid": int((ph == 1).sum())}
        out = O.run(case, rp, col, 1, rtol=1e-12, state=state)
        return out["U"], out["phase"], out["step"], out["voln1rel"]

    return step


def test_driver_rules_at_thresholds():
    vrs = R.run_scenario(oracle_stepper())
    assert len(vrs) == 6


def test_front_uses_clamped_value():
    R.run_clamp_before_front(oracle_stepper())


def test_rules_catch_a_planted_mistake():
    """The scenario fails an oracle whose front test is `<` instead of `<=`: emulated by a
    stepper that moves the threshold one ulp down (the same comparison outcomes as `<`)."""
    inner = oracle_stepper()

    def wrong(case, U, ph, s, vr):
        case.params.front_threshold = np.nextafter(R.THR, 0.0)
        try:
            return inner(case, U, ph, s, vr)
        finally:
            case.params.front_threshold = R.THR

    try:
        R.run_scenario(wrong)
    except AssertionError:
        return
    raise AssertionError("the scenario did not detect a strict front comparison")


def test_rules_catch_a_shifted_cadence():
    """The scenario fails a driver that refreshes voln1rel one step late (s % 3 == 1)."""
    inner = oracle_stepper()

    def late(case, U, ph, s, vr):
        U2, ph2, s2, vr2 = inner(case, U, ph, s - 1, vr)
        return U2, ph2, s2 + 1, vr2

    try:
        R.run_scenario(late)
    except AssertionError:
        return
    raise AssertionError("the scenario did not detect a shifted voln1rel cadence")
