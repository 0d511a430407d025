"""GPU parity of libnnpp against the CPU oracle and the reference golden vectors.

Tolerances: neighbour lists bit-exact; FP64 operators (residuals, potential operator) within
1e-12 relative (max-norm per field); FP64 multi-step trajectories within 1e-12 relative with
both solves at rtol 1e-13; mixed precision within 1e-6 relative of FP64."""
import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, nn, orc, rel_err, reference_params

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    p = reference_params()
    p.n_nodes = d["x"].size
    case = nn.synth.Case(p, d["x"], d["y"], d["z"], None, d["bctypes"], d["bcvalues"],
                         d["Un"].copy())
    return d, case


def problem(case):
    return nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)


def oracle_nb(case):
    p = case.params
    st, rp, col, ml = O.build_neighbors(case.x, case.y, case.z, p.delta * (1.00 + 1e-5),
                                        p.dx * 10., p.dy * 10., p.dz * 10.)
    assert st == 0
    return rp, col, ml


@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_neighbors_bit_exact_vs_reference(name):
    say("case", name)
    d, case = load_golden(name)
    with problem(case) as pb:
        nnz, ml = pb.build_clouds_hash()
        rp, col = pb.neighbors()
    assert np.array_equal(rp, d["row_ptr"]) and np.array_equal(col, d["col"])
    assert ml == int(d["max_len"])


@pytest.mark.parametrize("kind,ng", [("regular", 24), ("regular", 17), ("screw", 20)])
def test_neighbors_bit_exact_vs_oracle(kind, ng):
    say("case", kind, ng)
    case = getattr(nn.synth, kind)(ng, S=2)
    rp_o, col_o, ml_o = oracle_nb(case)
    with problem(case) as pb:
        nnz, ml = pb.build_clouds_hash()
        rp, col = pb.neighbors()
    assert nnz == int(rp_o[-1]) and ml == ml_o
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)


def test_neighbors_overflow_status():
    say("case overflow")
    case = nn.synth.regular(8, S=1)
    with problem(case) as pb:
        with pytest.raises(nn.NnppError) as e:
            pb.build_clouds_hash(max_node_no=100)
        assert e.value.code == nn._abi.ERR_OVERFLOW


@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_residuals_vs_reference_golden(name):
    say("case residuals", name)
    d, case = load_golden(name)
    with problem(case) as pb:
        pb.set_neighbors(d["row_ptr"], d["col"])
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.set_phase(d["phase"])
        pb.set_state(d["Un"], nn.SLOT_N)
        pb.set_state(d["U1"], nn.SLOT_N1)
        F = pb.residuals(float(d["voln1rel"]))
        Fb = pb.residuals(float(d["voln1rel"]), apply_bc=True)
    for f in range(11):
        assert rel_err(F[f], d["F"][f]) < 1e-12, f
    Fo = O.bc_override(case, d["U1"], d["F"])
    for f in range(11):
        assert rel_err(Fb[f], Fo[f]) < 1e-12, f


def test_residuals_generic_vs_oracle_and_mixed():
    say("case residuals generic")
    case = nn.synth.regular(16, S=11, seed=5)
    rp, col, _ = oracle_nb(case)
    rng = np.random.default_rng(9)
    U, ph, _, _ = O.init_state(case)
    U1 = U * rng.uniform(0.98, 1.02, size=U.shape)
    U1[11] = rng.uniform(-0.05, 0.05, case.n)
    U[11] = rng.uniform(-0.05, 0.05, case.n)
    Fo = O.residuals(case, rp, col, ph, U1, U, 0.4)
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_phase(ph)
        pb.set_state(U, nn.SLOT_N)
        pb.set_state(U1, nn.SLOT_N1)
        F = pb.residuals(0.4)
        Fm = pb.residuals(0.4, precision=nn.MIXED)
    for f in range(12):
        assert rel_err(F[f], Fo[f]) < 1e-12, f
        assert rel_err(Fm[f], Fo[f]) < 1e-5, f


def test_poisson_apply_vs_oracle():
    say("case poisson apply")
    d, case = load_golden("jit7")
    rng = np.random.default_rng(1)
    v = rng.standard_normal(case.n)
    yo = O.poisson_apply(case, d["row_ptr"], d["col"], d["phase"], v)
    with problem(case) as pb:
        pb.set_neighbors(d["row_ptr"], d["col"])
        pb.set_phase(d["phase"])
        y = pb.poisson_apply(v)
    assert rel_err(y, yo) < 1e-12


@pytest.mark.parametrize("precision", [nn.FP64, nn.MIXED])
def test_poisson_solve_vs_oracle(precision):
    say("case poisson solve", precision)
    case = nn.synth.regular(20, S=11, seed=2)
    rp, col, _ = oracle_nb(case)
    U, ph, _, _ = O.init_state(case)
    phi_o, it_o, rr_o, st = O.poisson_solve(case, rp, col, ph, U, U[11], rtol=1e-13)
    assert st == 0
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000, inner_rtol=1e-4)
        it, rr, st = pb.poisson_solve(precision)
        phi = pb.get_state()[11]
    assert st == 0 and rr <= 1e-13
    assert rel_err(phi, phi_o) < 1e-10
    if precision == nn.FP64:
        assert abs(it - it_o) <= 2


def test_mixed_inner_rtol_sets_refinement_passes():
    """inner_rtol is the residual-norm reduction of each FP32 pass of the mixed solve: a larger
    one ends passes sooner (more FP64 residual replacements); both stop at the same FP64
    relative residual, so the states agree to the solver tolerance. Values outside (0, 1) are
    refused."""
    say("case mixed inner_rtol")
    case = nn.synth.regular(20, S=11, seed=2)
    rp, col, _ = oracle_nb(case)
    out = {}
    for ir in (1e-1, 1e-4):
        with problem(case) as pb:
            pb.set_neighbors(rp, col)
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.init_state(case.icvalues)
            pb.set_solver(rtol=1e-12, maxit=100000, inner_rtol=ir)
            st = pb.step(3, nn.MIXED)
            assert st["status"] == 0 and st["last_relres"] <= 1e-12
            out[ir] = (st["cg_spmv_fp64"], pb.get_state())
    assert out[1e-1][0] > out[1e-4][0], (out[1e-1][0], out[1e-4][0])
    for f in range(case.params.S + 1):
        assert rel_err(out[1e-1][1][f], out[1e-4][1][f]) < 1e-9, f
    with problem(case) as pb:
        for bad in (0.0, 1.0, -1e-3):
            with pytest.raises(nn.NnppError):
                pb.set_solver(rtol=1e-10, maxit=100, inner_rtol=bad)


def test_step_fp64_matches_oracle():
    say("case step fp64 regular16")
    case = nn.synth.regular(16, S=11, seed=0)
    rp, col, _ = oracle_nb(case)
    ref = O.run(case, rp, col, 10, rtol=1e-13)
    with problem(case) as pb:
        pb.build_clouds_hash()
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000)
        st = pb.step(10, nn.FP64)
        U = pb.get_state()
        ph = pb.get_phase()
    assert st["status"] == 0 and st["steps"] == 10
    assert np.array_equal(ph, ref["phase"])
    for f in range(12):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, f


def test_step_reference_chemistry_matches_oracle():
    say("case step reference chemistry")
    case = nn.synth.regular(12, S=10, seed=4)
    p = reference_params()
    # reference chemistry on the lattice of the synthetic case (production dx, dt)
    case.params = p
    s = p.dx * 12
    case.x, case.y, case.z = case.x * s, case.y * s, case.z * s
    case.params.n_nodes = case.n
    rng = np.random.default_rng(5)
    case.icvalues = 10.0 ** rng.uniform(-8, -1, size=(11, case.n))
    case.icvalues[0][case.x > s / 2] = 0.0  # half the box solid
    case.icvalues[10] = 0.0
    case.bcvalues[:10][case.bcvalues[:10] != -1] = 1e-3 * p.CSolid
    p.dt = 1.0
    rp, col, _ = oracle_nb(case)
    ref = O.run(case, rp, col, 5, rtol=1e-13)
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000)
        st = pb.step(5, nn.FP64)
        U = pb.get_state()
    assert st["status"] == 0
    # This case is ill-conditioned on purpose (states over 7 decades, dt = 1 s against fast
    # reactions): the oracle's own 5-step result moves by ~3e-13 (field max) when its initial
    # state is perturbed by 1e-15 relative. The GPU (whose per-step operators agree to ~1e-15)
    # must stay within 1e-12 after one step, and after five within 1e-12 or 20x that
    # measured sensitivity, whichever is larger.
    U0, ph0, sc0, v0 = O.init_state(case)
    st0 = {"U": U0, "phase": ph0, "step": 1, "voln1rel": 0.0, "vol0rel": v0, "solid": sc0}
    Up = U0.copy()
    Up[:10] *= 1 + 1e-15 * np.random.default_rng(6).standard_normal(Up[:10].shape)
    pert = O.run(case, rp, col, 5, rtol=1e-13, state=dict(st0, U=Up))["U"]
    sens = max(rel_err(pert[f], ref["U"][f]) for f in range(11))
    one = O.run(case, rp, col, 1, rtol=1e-13)["U"]
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000)
        pb.step(1, nn.FP64)
        U1 = pb.get_state()
    for f in range(11):
        assert rel_err(U1[f], one[f]) < 1e-12, f
    tol = max(1e-12, 20 * sens)
    say("reference chemistry: sensitivity", sens, "tolerance", tol)
    for f in range(11):
        assert rel_err(U[f], ref["U"][f]) < tol, (f, rel_err(U[f], ref["U"][f]), sens)


def test_front_update_and_volume_cadence():
    say("case front")
    case = nn.synth.regular(14, S=3, seed=1)
    case.params.front_threshold = 1.0  # about half the solid nodes flip in the first step
    rp, col, _ = oracle_nb(case)
    ref = O.run(case, rp, col, 9, rtol=1e-13)
    with problem(case) as pb:
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-13, maxit=100000)
        st = pb.step(9, nn.FP64)
        U = pb.get_state()
        ph = pb.get_phase()
    assert st["flipped"] > 0
    assert np.array_equal(ph, ref["phase"])
    assert abs(st["voln1rel"] - ref["voln1rel"]) <= 1e-12 * abs(ref["voln1rel"])
    for f in range(4):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, f


def test_mixed_step_close_to_fp64():
    say("case mixed vs fp64")
    case = nn.synth.regular(20, S=11, seed=3)
    out = {}
    for prec in (nn.FP64, nn.MIXED):
        with problem(case) as pb:
            pb.build_clouds_hash()
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.init_state(case.icvalues)
            pb.set_solver(rtol=1e-10)
            st = pb.step(20, prec)
            assert st["status"] == 0
            out[prec] = pb.get_state()
    for f in range(12):
        assert rel_err(out[nn.MIXED][f], out[nn.FP64][f]) < 1e-6, f


@pytest.mark.parametrize("kind", ["regular", "screw"])
def test_storage_formats_bit_identical(kind):
    """Stencil forms (lattices: the potential operator and the explicit flux; the SpMV's
    lattice-ball and table-driven loops), 16-bit column steps vs 32-bit columns, dictionary
    weight classes vs stored weights: the same additions in the same order, so the trajectories
    agree bit for bit."""
    say("case storage formats", kind)
    # 16^3 regular: spacing 1/16 makes every bond vector exact (dictionary and stencil apply);
    # the screw case has per-node volumes (column form only)
    case = nn.synth.regular(16, S=11, seed=9) if kind == "regular" else nn.synth.screw(16, S=11, seed=9)
    case.params.front_threshold = 1.0  # flips: the operator is rebuilt on the way
    rp, col, _ = oracle_nb(case)
    W, D, T = nn.FORMAT_WIDE_COLUMNS, nn.FORMAT_NO_DICTIONARY, nn.FORMAT_NO_STENCIL
    B = nn.FORMAT_NO_BALL  # the table-driven stencil loop instead of the lattice-ball one
    Z = nn.FORMAT_NO_ZMARCH  # the tile SpMV of the lattice ball instead of the z-marching one
    formats = [nn.FORMAT_AUTO, Z, W | Z, D | Z, T, T | W, T | D, T | W | D, B, B | T]
    res = {}
    for prec in (nn.FP64, nn.MIXED):
        out = res[prec] = {}
        for fmt in formats:
            with problem(case) as pb:
                pb.set_format(fmt)
                pb.set_neighbors(rp, col)
                pb.set_bc(case.bctypes, case.bcvalues)
                pb.init_state(case.icvalues)
                pb.set_solver(rtol=1e-12, maxit=100000)
                st = pb.step(4, prec)
                assert st["status"] == 0
                out[fmt] = (pb.get_state(), st)
        U0, st0 = out[Z]  # tile form: the column forms' additions, dot products included
        Uc = out[T][0]  # column form
        assert st0["flipped"] > 0
        for fmt, (U, st) in out.items():
            stencil = kind == "regular" and not (fmt & T)
            zm = stencil and not (fmt & (B | Z))
            # 16^3 lattice with delta = 3 spacings: the offsets are the radius-3 ball on the
            # full box, so the z-marching kernels run unless a flag turns them off
            assert st["pot_stencil"] == (0 if not stencil else (1 if fmt & B else (3 if fmt & Z
                                                                                    else 4))), fmt
            # the explicit flux takes the stencil form wherever the operator does (FP32 bond
            # sums in mixed precision)
            # (FP64 on the full box lattice: the z-marching flux)
            assert st["flux_stencil"] == (0 if not stencil else (2 if prec == nn.MIXED
                                                                   else (3 if zm else 1))), fmt
            if stencil:
                assert st["pot_col_bytes"] == 0 and st["pot_weight_bytes"] == 0
            else:
                assert st["pot_col_bytes"] == (4 if fmt & W else 2)
                assert st["pot_weight_bytes"] == (4 if (fmt & D or kind == "screw") else 1)
            if zm:
                # the z-marching SpMV: FP64 rows as the other forms, dot products summed in
                # another order; mixed: FP32 rows with their own chunks and the FP32 flux
                for f in range(case.params.S + 1):
                    assert rel_err(U[f], res[nn.FP64][Z][0][f]) < (1e-12 if prec == nn.FP64
                                                                 else 1e-6), (prec, fmt, f)
            elif prec == nn.FP64 or not stencil:
                # FP64: every form performs the same additions; FP32 column forms too
                assert np.array_equal(U, U0 if prec == nn.FP64 else Uc), (prec, fmt)
                if prec == nn.FP64:
                    assert st["cg_iterations"] == st0["cg_iterations"], fmt
            else:
                # mixed stencil forms: FP32 stencil SpMVs with the same additions, FP64 residual
                # replacements that bring the solves to the FP64 rtol; the FP32 stencil flux
                # walks the lattice ball plane by plane, or (B) the offset table in chunks of 16
                if not fmt & B:
                    assert np.array_equal(U, U0), fmt
                for f in range(case.params.S + 1):
                    assert rel_err(U[f], res[nn.FP64][nn.FORMAT_AUTO][0][f]) < 1e-6, (fmt, f)
    with problem(case) as pb:
        with pytest.raises(nn.NnppError):
            pb.set_format(32)


@pytest.mark.parametrize("kind", ["regular", "screw"])
def test_potential_operator_kernel(kind):
    """The solver's SpMV kernel alone: FP64 against the oracle's K11 action on the free rows
    (1e-12), FP32 + Kahan against FP64 (1e-6), and every storage form bit-identical."""
    say("case potential operator kernel", kind)
    case = nn.synth.regular(16, S=11, seed=4) if kind == "regular" else nn.synth.screw(16, S=11, seed=4)
    rp, col, _ = oracle_nb(case)
    U, ph, _, _ = O.init_state(case)
    S = case.params.S
    liquid = ph == 0
    rows = np.repeat(np.arange(case.n), np.diff(rp))
    has_ll = np.zeros(case.n, bool)
    has_ll[rows[liquid[rows] & liquid[col]]] = True
    dirphi = (case.bctypes == 0) & (case.bcvalues[S] != -1)
    free = liquid & has_ll & ~dirphi
    rng = np.random.default_rng(3)
    x = np.where(free, rng.standard_normal(case.n), 0.0)
    yo = O.poisson_apply(case, rp, col, ph, x)
    sc = case.vol / (case.params.dx * case.params.dy * case.params.dz) if case.vol is not None else 1.0
    yo = np.where(free, yo * sc, 0.0)
    T = nn.FORMAT_NO_STENCIL
    out = {}
    B = nn.FORMAT_NO_BALL
    Z = nn.FORMAT_NO_ZMARCH
    for fmt in (nn.FORMAT_AUTO, Z, B, T, T | nn.FORMAT_NO_DICTIONARY, T | nn.FORMAT_WIDE_COLUMNS):
        with problem(case) as pb:
            pb.set_format(fmt)
            pb.set_neighbors(rp, col)
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.init_state(case.icvalues)
            out[fmt] = (pb.potential_operator_apply(x, nn.FP64),
                        pb.potential_operator_apply(x, nn.MIXED))
    y64, y32 = out[Z]
    assert rel_err(y64, yo) < 1e-12
    assert np.abs(y32 - y64).max() <= 1e-6 * np.abs(y64).max()
    yc32 = out[T][1]
    yz32 = out[nn.FORMAT_AUTO][1]
    for fmt, (a, b) in out.items():
        assert np.array_equal(a, y64), fmt  # FP64: the same additions in every form
        # FP32: the column forms agree among themselves, the two stencil tile loops between
        # them; the z-marching rows (plane chunks) are their own
        exp = {nn.FORMAT_AUTO: yz32 if kind == "regular" else yc32, Z: y32, B: y32}
        assert np.array_equal(b, exp.get(fmt, yc32)), fmt
    # FP32 stencil rows (chunk sums, compensated across chunks) against FP32 column rows
    assert np.abs(y32 - yc32).max() <= 4e-7 * np.abs(y64).max()
    assert np.abs(yz32 - yc32).max() <= 4e-7 * np.abs(y64).max()


def test_deterministic_and_error_paths():
    say("case determinism / errors")
    case = nn.synth.regular(12, S=5, seed=6)
    runs = []
    for _ in range(2):
        with problem(case) as pb:
            pb.build_clouds_hash()
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.init_state(case.icvalues)
            pb.step(3, nn.FP64)
            runs.append(pb.get_state())
    assert np.array_equal(runs[0], runs[1])
    with problem(case) as pb:
        pb.build_clouds_hash()
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=1e-14, maxit=1)
        st = pb.step(1, nn.FP64)
        assert st["status"] == nn._abi.ERR_NOCONV
        U = pb.get_state()
        U[2, 5] = np.nan
        pb.set_state(U)
        pb.set_solver(rtol=1e-10)
        with pytest.raises(nn.NnppError) as e:
            pb.step(1, nn.FP64)
        assert e.value.code == nn._abi.ERR_NONFINITE
    with problem(case) as pb:
        with pytest.raises(nn.NnppError) as e:
            pb.step(1)
        assert e.value.code == nn._abi.ERR_STATE


def test_cpp_driver_end_to_end(tmp_path):
    """host/nnpp_main on reference-format job files vs the oracle's trajectory."""
    import subprocess
    say("case nnpp_main driver")
    case = nn.synth.regular(10, S=11, seed=8)
    case.params.n_steps = 16
    job = str(tmp_path / "NNPP")
    nn.formats.write_job(job, case)
    exe = os.path.join(os.path.dirname(nn._abi.LIB_PATH), "bin", "nnpp_main")
    r = subprocess.run([exe, "-01", job, "0", "--species", "11", "--rtol", "1e-13"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rp, col, _ = oracle_nb(case)
    ref = O.run(case, rp, col, 16, rtol=1e-13)
    U = np.zeros((12, case.n))
    ph = np.zeros(case.n, dtype=np.int32)
    assert nn.lib.nnpp_read_start_csv((job + "-01-16.output.csv").encode(), case.n, 11,
                                      nn._abi.dptr(U), nn._abi.iptr(ph)) == 0
    assert np.array_equal(ph, ref["phase"])
    for f in range(12):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, f
    vol = open(job + "-01.volume").read().split()
    assert len(vol) == 2 and abs(float(vol[-1].rstrip(",")) - ref["voln1rel"]) <= 1e-12 * abs(ref["voln1rel"])
    # .energy (main.cpp:962-998): sum of U_Mg^2 dx dy dz over the nodes with |x| or |y| below
    # the largest coordinate of the node file (energyBound = nodes.maxCoeff(), main.cpp:242),
    # written with the stream's default 6 significant digits
    bound = max(case.x.max(), case.y.max(), case.z.max())
    sel = (np.abs(case.x) < bound) | (np.abs(case.y) < bound)
    assert not sel.all()  # the nodes at x = y = max are left out
    p = case.params
    e_hand = float(np.sum(U[0][sel] ** 2) * p.dx * p.dy * p.dz)
    ene = open(job + "-01.energy").read().split()
    assert len(ene) == 2 and abs(float(ene[-1]) - e_hand) <= 5e-6 * e_hand, (ene, e_hand)
