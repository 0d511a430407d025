"""The library's explicit step (nnpp_step, FP64 and mixed) against the driver rules of main.cpp
at their thresholds: clamp (805-837), front update (915-922), voln1rel cadence (944-949).
Same states and expected outcomes as the oracle's test (driver_rules.py)."""
import sys

import numpy as np
import pytest

import driver_rules as R
from helpers import nn, orc

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def gpu_stepper(precision):
    cache = {}

    def step(case, U, ph, s, vr):
        if "pb" not in cache:
            p = case.params
            st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5),
                                               p.dx * 10., p.dy * 10., p.dz * 10.)
            assert st == 0
            pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
            pb.set_neighbors(rp, col)
            pb.set_solver(rtol=1e-12)
            cache["pb"] = pb
        pb = cache["pb"]
        pb.set_params(case.params)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.set_state(U)
        pb.set_phase(ph)
        pb.set_time(s, vr, R.VOL0REL)
        st = pb.step(1, precision)
        assert st["status"] == 0
        s2, vr2, _ = pb.get_time()
        return pb.get_state(), pb.get_phase(), s2, vr2

    return step, cache


@pytest.mark.parametrize("precision", [nn.FP64, nn.MIXED])
def test_gpu_driver_rules_at_thresholds(precision):
    print("case driver rules", precision, flush=True)
    step, cache = gpu_stepper(precision)
    try:
        R.run_scenario(step)
        R.run_clamp_before_front(step)
    finally:
        if "pb" in cache:
            cache["pb"].close()
