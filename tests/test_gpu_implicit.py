"""GPU parity of the implicit (reference Newton) step, through the C ABI.

The frozen Jacobian (utils.h:844-1849) is checked against the reference's own NewtonMatrices
(golden products) and against the oracle, to 1e-12 relative. Implicit trajectories are checked
against the oracle's or_step_implicit with tight linear and Newton tolerances (both then solve
the same equations; the species solvers are the same BiCGSTAB, the potential solvers differ:
CG on the eliminated system here, BiCGSTAB on the full K11 in the oracle), to 1e-9 relative.
"""
import sys

import numpy as np
import pytest

from helpers import nn, orc, rel_err, reference_params
from test_gpu_parity import load_golden, oracle_nb, problem

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)

TIGHT = dict(newton_tol=1e-11, newton_maxit=100, lin_rtol=1e-13, lin_maxit=5000)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_jacobian_vs_reference_golden(name):
    say("case jacobian golden", name)
    d, case = load_golden(name)
    with problem(case) as pb:
        pb.set_neighbors(d["row_ptr"], d["col"])
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.set_phase(d["phase"])
        pb.set_state(d["Un"], nn.SLOT_N)
        pb.set_state(d["U1"], nn.SLOT_N1)
        for m in range(2):
            out, dg = pb.jacobian_apply(d["jac_v"][m], float(d["voln1rel"]), want_diag=True)
            for f in range(11):
                assert rel_err(out[f], d["jac_kv"][f, m]) < 1e-12, (f, m)
                assert rel_err(dg[f], d["jac_diag"][f]) < 1e-12, f


@pytest.mark.parametrize("kind,ng,S", [("regular", 10, 11), ("screw", 12, 4)])
def test_jacobian_generic_vs_oracle(kind, ng, S):
    say("case jacobian oracle", kind, ng, S)
    case = getattr(nn.synth, kind)(ng, S=S, seed=3)
    rp, col, _ = oracle_nb(case)
    U, ph, _, _ = O.init_state(case)
    rng = np.random.default_rng(4)
    U[S] = rng.uniform(-0.05, 0.05, case.n)  # nonzero migration terms
    v = rng.standard_normal(U.shape)
    ko, do = O.jacobian_apply(case, rp, col, ph, U, v, 0.3, want_diag=True)
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.set_phase(ph)
        pb.set_state(U, nn.SLOT_N)
        k, dg = pb.jacobian_apply(v, 0.3, want_diag=True)
    for f in range(S + 1):
        assert rel_err(k[f], ko[f]) < 1e-12, f
        assert rel_err(dg[f], do[f]) < 1e-12, f


def _flip_case(ng=12, S=3, seed=1):
    case = nn.synth.regular(ng, S=S, seed=seed)
    case.params.front_threshold = 1.0  # solid nodes dissolve within a few steps
    case.params.sub_steps = 2          # several voln1rel (BV) updates
    return case


def _run_gpu(case, rp, col, nsteps, precision=nn.FP64, **opts):
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000)
        st = pb.step_implicit(nsteps, precision, **opts)
        return (st, pb.get_state(nn.SLOT_N), pb.get_state(nn.SLOT_N1), pb.get_phase(),
                pb.get_time())


def test_implicit_steps_match_oracle():
    say("case implicit generic regular12 with flips")
    case = _flip_case()
    rp, col, _ = oracle_nb(case)
    ref = O.run_implicit(case, rp, col, 4, **TIGHT)
    st, U, U1, ph, (step, vr, v0) = _run_gpu(case, rp, col, 4, **TIGHT)
    assert st["status"] == 0 and st["steps"] == 4 and st["unconverged_steps"] == 0
    assert sum(r["flips"] for r in ref["log"]) > 0 and st["flipped"] == sum(
        r["flips"] for r in ref["log"])
    assert np.array_equal(ph, ref["phase"])
    assert step == ref["step"] and abs(vr - ref["voln1rel"]) <= 1e-10 * abs(ref["voln1rel"])
    for f in range(case.params.S + 1):
        assert rel_err(U[f], ref["U"][f]) < 1e-9, f
        # slot N1 holds the next start of the rotation: U^{n-1}
        assert rel_err(U1[f], ref["Uit"][f]) < 1e-9, f
    newton_ref = sum(r["newton"] for r in ref["log"])
    assert abs(st["newton_iterations"] - newton_ref) <= 4


def test_implicit_reference_chemistry_matches_oracle():
    say("case implicit reference chemistry")
    case = nn.synth.regular(8, S=10, seed=4)
    p = reference_params()
    case.params = p
    s = p.dx * 8
    case.x, case.y, case.z = case.x * s, case.y * s, case.z * s
    p.n_nodes = case.n
    rng = np.random.default_rng(5)
    case.icvalues = 10.0 ** rng.uniform(-8, -1, size=(11, case.n))
    case.icvalues[0][case.x > s / 2] = 0.0  # half the box solid
    case.icvalues[10] = 0.0
    case.bcvalues[:10][case.bcvalues[:10] != -1] = 1e-3 * p.CSolid
    p.dt = 1.0
    rp, col, _ = oracle_nb(case)
    # the interface H+ rows contract slowly (utils.h:1089-1100 against 241-253): both sides run
    # the same ten unconverged iterations per step
    opts = dict(TIGHT, newton_maxit=10)
    ref = O.run_implicit(case, rp, col, 2, **opts)
    st, U, U1, ph, _ = _run_gpu(case, rp, col, 2, **opts)
    assert st["status"] == 0 and st["steps"] == 2
    assert st["newton_iterations"] == sum(r["newton"] for r in ref["log"])
    assert np.array_equal(ph, ref["phase"])
    for f in range(11):
        assert rel_err(U[f], ref["U"][f]) < 1e-9, f


def test_implicit_start_state_and_mixed_potential():
    say("case implicit start=state, mixed")
    case = _flip_case(ng=10)
    rp, col, _ = oracle_nb(case)
    st64, U64, _, ph64, _ = _run_gpu(case, rp, col, 3, start=nn._abi.NEWTON_START_STATE,
                                     **TIGHT)
    stm, Um, _, phm, _ = _run_gpu(case, rp, col, 3, precision=nn.MIXED,
                                  start=nn._abi.NEWTON_START_STATE, **TIGHT)
    st_r, U_r, _, ph_r, _ = _run_gpu(case, rp, col, 3, **TIGHT)
    assert st64["status"] == 0 and stm["status"] == 0
    assert np.array_equal(ph64, phm) and np.array_equal(ph64, ph_r)
    for f in range(case.params.S + 1):
        # converged Newton: the start only changes the iteration path
        assert rel_err(U64[f], U_r[f]) < 1e-9, f
        assert rel_err(Um[f], U64[f]) < 1e-8, f


def test_implicit_reference_defaults_and_guards():
    say("case implicit defaults")
    case = _flip_case(ng=10)
    rp, col, _ = oracle_nb(case)
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        st = pb.step_implicit(2)  # EPS = 1e-5 everywhere, as main.cpp
        assert st["status"] in (0, nn._abi.ERR_NOCONV) and st["steps"] == 2
        assert st["newton_iterations"] >= 2 and st["lin_iterations"] > 0
        assert np.isfinite(pb.get_state()).all()
        with pytest.raises(nn.NnppError) as e:
            pb.step_implicit(1, newton_maxit=-1)
        assert e.value.code == nn._abi.ERR_ARG
        pb.tape_begin(1)
        with pytest.raises(nn.NnppError) as e:
            pb.step_implicit(1)
        assert e.value.code == nn._abi.ERR_UNSUPPORTED
        pb.tape_end()


def test_cpp_driver_implicit(tmp_path):
    """host/nnpp_main --implicit on reference-format job files: the raw initial values as U^n,
    the Dirichlet values on the first iterate (main.cpp:556-640), against or_step_implicit."""
    import os
    import subprocess
    say("case nnpp_main --implicit")
    case = _flip_case(ng=10)
    case.params.n_steps = 8
    job = str(tmp_path / "NNPP")
    nn.formats.write_job(job, case)
    exe = os.path.join(os.path.dirname(nn._abi.LIB_PATH), "bin", "nnpp_main")
    r = subprocess.run([exe, "-01", job, "0", "--species", "3", "--chem", "generic", "--rtol",
                        "1e-13", "--implicit", "--newton-tol", "1e-11", "--lin-rtol", "1e-13"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rp, col, _ = oracle_nb(case)
    _, ph0, sc0, v0 = O.init_state(case)
    state = {"U": case.icvalues.copy(), "phase": ph0, "step": 1, "voln1rel": 0.0,
             "vol0rel": v0, "solid": sc0}
    ref = O.run_implicit(case, rp, col, 8, state=state, **TIGHT)
    U = np.zeros((4, case.n))
    ph = np.zeros(case.n, dtype=np.int32)
    assert nn.lib.nnpp_read_start_csv((job + "-01-8.output.csv").encode(), case.n, 3,
                                      nn._abi.dptr(U), nn._abi.iptr(ph)) == 0
    assert np.array_equal(ph, ref["phase"])
    for f in range(4):
        assert rel_err(U[f], ref["U"][f]) < 1e-9, f
    vol = open(job + "-01.volume").read().split()
    assert len(vol) == 4  # sub_steps = 2: outputs after steps 2, 4, 6, 8
    assert abs(float(vol[-1].rstrip(",")) - ref["voln1rel"]) <= 1e-10 * abs(ref["voln1rel"])


def test_jacobian_and_implicit_steps_stencil_form():
    """Regular 16^3 (spacing 1/16, exact bond vectors): the Jacobian apply and the implicit
    steps take the lattice stencil form of the bond sums; they agree bit for bit with the
    column form (the same additions) and with the oracle."""
    say("case jacobian / implicit stencil form")
    case = nn.synth.regular(16, S=3, seed=6)
    case.params.front_threshold = 1.0
    case.params.sub_steps = 2
    rp, col, _ = oracle_nb(case)
    U, ph, _, _ = O.init_state(case)
    rng = np.random.default_rng(11)
    U[3] = rng.uniform(-0.05, 0.05, case.n)  # nonzero migration terms
    v = rng.standard_normal(U.shape)
    ko, do = O.jacobian_apply(case, rp, col, ph, U, v, 0.3, want_diag=True)
    res = {}
    for fmt in (nn.FORMAT_AUTO, nn.FORMAT_NO_ZMARCH, nn.FORMAT_NO_BALL, nn.FORMAT_NO_STENCIL):
        with problem(case) as pb:
            pb.set_format(fmt)
            pb.set_neighbors(rp, col)
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.set_phase(ph)
            pb.set_state(U, nn.SLOT_N)
            k, dg = pb.jacobian_apply(v, 0.3, want_diag=True)
            pb.init_state(case.icvalues)
            pb.set_solver(rtol=1e-13, maxit=100000)
            st = pb.step_implicit(3, **TIGHT)
            assert st["status"] == 0 and st["flipped"] > 0
            res[fmt] = (k, dg, pb.get_state(nn.SLOT_N), pb.get_phase())
    with problem(case) as pb:  # the geometry qualifies (the same check as the Jacobian's)
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        assert pb.step(1)["flux_stencil"] == 3  # (the explicit step: the z-marching form)
    ka, da, Ua, pa = res[nn.FORMAT_NO_ZMARCH]
    # the potential SpMV's table-driven stencil loop; the column forms
    for other in (nn.FORMAT_NO_BALL, nn.FORMAT_NO_STENCIL):
        kc, dc, Uc, pc = res[other]
        assert np.array_equal(ka, kc) and np.array_equal(da, dc), other
        assert np.array_equal(Ua, Uc) and np.array_equal(pa, pc), other
    # the z-marching potential SpMV (dot products summed in another order)
    kz, dz, Uz, pz = res[nn.FORMAT_AUTO]
    assert np.array_equal(ka, kz) and np.array_equal(da, dz) and np.array_equal(pa, pz)
    for f in range(case.params.S + 1):
        assert rel_err(Uz[f], Ua[f]) < 1e-12, f
    for f in range(case.params.S + 1):
        assert rel_err(ka[f], ko[f]) < 1e-12, f
        assert rel_err(da[f], do[f]) < 1e-12, f
