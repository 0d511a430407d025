"""The five benchmark workloads at their own sizes, on the GPU, through the C ABI.

  * 24^3 lattice, 11 species, 100 FP64 steps: every field within 1e-12 (max norm, relative)
    of the oracle, phases identical;
  * 128^3 lattice: a 32^3 crop (touching a Dirichlet face and the Mg sphere) within 1e-12 of
    the oracle after 10 FP64 steps; at full size, one FP64 step and one potential-operator
    application with the stencil forms against the column forms (NNPP_FORMAT_NO_STENCIL);
  * 128^3 mixed against FP64 after 1000 steps: per-field max and L2 relative error <= 1e-6;
  * the jittered screw on a 128^3 lattice (about 1.5M nodes): neighbour lists bit-exact
    against the oracle's restatement of grid.h at full size, Voronoi-approximate volumes,
    FP64 and mixed for 200 steps (front flips included), per-field error <= 1e-6;
  * 64^3 adjoint, 200 FP64 steps, checkpoints every 20: bit-identical to the full tape,
    central differences (h = 1e-6 relative) within 1e-5 on five parameters; the GPU adjoint
    against the oracle's adjoint on a 16^3 case.
Solves run to rtol 1e-13 where the oracle is compared, 1e-10 (the workload's) elsewhere."""
import sys

import numpy as np
import pytest
import torch

from helpers import nn, orc, rel_err

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_nb(case):
    p = case.params
    st, rp, col, ml = O.build_neighbors(case.x, case.y, case.z, p.delta * (1.00 + 1e-5),
                                        p.dx * 10., p.dy * 10., p.dz * 10.)
    assert st == 0
    return rp, col, ml


def gpu_problem(case, rp=None, col=None, rtol=1e-13, fmt=None):
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    if fmt is not None:
        pb.set_format(fmt)
    if rp is None:
        pb.build_clouds_hash()
    else:
        pb.set_neighbors(rp, col)
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    pb.set_solver(rtol=rtol, maxit=100000)
    return pb


def field_errors(a, b):
    out = []
    for f in range(a.shape[0]):
        d = a[f] - b[f]
        out.append((float(np.abs(d).max() / max(np.abs(b[f]).max(), 1e-300)),
                    float(np.linalg.norm(d) / max(np.linalg.norm(b[f]), 1e-300))))
    return out


def box_crop(case, lo, hi):
    """Nodes with lo <= (x, y, z) < hi componentwise, as a case of its own."""
    m = np.ones(case.n, dtype=bool)
    for a, l, h in zip((case.x, case.y, case.z), lo, hi):
        m &= (a >= l) & (a < h)
    p = case.params.copy()
    p.n_nodes = int(m.sum())
    return nn.synth.Case(p, case.x[m].copy(), case.y[m].copy(), case.z[m].copy(), None,
                         case.bctypes[m].copy(), case.bcvalues[:, m].copy(),
                         case.icvalues[:, m].copy(), name=case.name + "-crop")


def test_c1_regular24_100_steps_fp64():
    say("case C1 regular 24^3, 100 FP64 steps")
    case = nn.synth.regular(24, S=11, seed=0)
    rp, col, _ = oracle_nb(case)
    assert rp[-1] == 1455112
    ref = O.run(case, rp, col, 100, rtol=1e-13)
    with gpu_problem(case) as pb:
        rp_g, col_g = pb.neighbors()
        assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
        st = pb.step(100, nn.FP64)
        U, ph = pb.get_state(), pb.get_phase()
        t = pb.get_time()
    assert st["status"] == 0 and st["steps"] == 100
    assert np.array_equal(ph, ref["phase"])
    assert t[0] == ref["step"] and abs(t[1] - ref["voln1rel"]) <= 1e-12 * abs(ref["voln1rel"])
    for f in range(12):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, (f, rel_err(U[f], ref["U"][f]))


@pytest.fixture(scope="module")
def case128():
    return nn.synth.regular(128, S=11, seed=0)


def test_c2_crop32_of_128_fp64(case128):
    say("case C2 32^3 crop of the 128^3 case, 10 FP64 steps")
    crop = box_crop(case128, (0.25, 0.25, 0.0), (0.5, 0.5, 0.25))
    assert crop.n == 32 ** 3
    assert (crop.bctypes == 0).any() and (crop.icvalues[0] > 0).any()  # a face and the sphere
    rp, col, _ = oracle_nb(crop)
    ref = O.run(crop, rp, col, 10, rtol=1e-13)
    with gpu_problem(crop) as pb:
        rp_g, col_g = pb.neighbors()
        assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
        st = pb.step(10, nn.FP64)
        U, ph = pb.get_state(), pb.get_phase()
    # a full box lattice: the z-marching forms of the flux and of the potential operator
    assert st["status"] == 0 and st["flux_stencil"] == 3 and st["pot_stencil"] == 4
    assert np.array_equal(ph, ref["phase"])
    for f in range(12):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, (f, rel_err(U[f], ref["U"][f]))


def test_c2_full128_stencil_vs_column_forms(case128):
    say("case C2 128^3: stencil forms against column forms")
    case = case128
    res = {}
    for name, fmt in (("auto", nn.FORMAT_AUTO), ("columns", nn.FORMAT_NO_STENCIL)):
        with gpu_problem(case, rtol=1e-10, fmt=fmt) as pb:
            x = np.random.default_rng(0).standard_normal(case.n)
            y = pb.potential_operator_apply(x, nn.FP64)
            st = pb.step(1, nn.FP64)
            res[name] = (y, pb.get_state(), pb.get_phase(), st)
            say(name, "flux_stencil", st["flux_stencil"], "pot_stencil", st["pot_stencil"],
                "cg", st["cg_iterations"])
    a, c = res["auto"], res["columns"]
    assert a[3]["flux_stencil"] == 3 and a[3]["pot_stencil"] == 4
    assert c[3]["flux_stencil"] == 0 and c[3]["pot_stencil"] == 0
    assert rel_err(a[0], c[0]) <= 1e-14  # one operator application, same bond sums
    assert np.array_equal(a[2], c[2])
    for f in range(12):
        assert rel_err(a[1][f], c[1][f]) < 1e-12, (f, rel_err(a[1][f], c[1][f]))


def test_c2_full128_implicit_jacobian_forms(case128):
    """128^3: one implicit step with the Jacobian apply in the z-marching form (the default on a
    full box lattice) and in the tile stencil form (NNPP_FORMAT_NO_ZMARCH). Both solve the same
    Newton systems with the bond sums added in another order, so with tight Newton and linear
    tolerances the states agree far below what a wrong bond or row would leave."""
    say("case C2 128^3: implicit step, z-marching against tile-form Jacobian apply")
    tight = dict(newton_tol=1e-11, newton_maxit=100, lin_rtol=1e-12, lin_maxit=5000)
    res = {}
    for name, fmt in (("zmarch", nn.FORMAT_AUTO), ("tile", nn.FORMAT_NO_ZMARCH)):
        with gpu_problem(case128, fmt=fmt) as pb:
            st = pb.step_implicit(1, **tight)
            res[name] = (pb.get_state(), pb.get_phase(), st)
            say(name, "newton", st["newton_iterations"], "lin", st["lin_iterations"],
                "cg", st["cg_iterations"])
    (Uz, pz, sz), (Ut, pt, stt) = res["zmarch"], res["tile"]
    assert sz["status"] == 0 and stt["status"] == 0
    assert np.array_equal(pz, pt)
    for f in range(12):
        assert rel_err(Uz[f], Ut[f]) < 1e-9, (f, rel_err(Uz[f], Ut[f]))


def test_c3_mixed_vs_fp64_1000_steps(case128):
    say("case C3 128^3 mixed vs FP64 after 1000 steps")
    case = case128
    with nn.Problem(case.params, case.x, case.y, case.z, None, device=0) as pb:
        pb.build_clouds_hash()
        rp, col = pb.neighbors()
    out = {}
    for name, prec in (("fp64", nn.FP64), ("mixed", nn.MIXED)):
        with gpu_problem(case, rp, col, rtol=1e-10) as pb:
            done = 0
            while done < 1000:
                st = pb.step(250, prec)
                assert st["status"] == 0
                done += st["steps"]
                say(name, done)
            out[name] = (pb.get_state(), pb.get_phase(), st)
    # the lattice's flux: the z-marching form in FP64, FP32 bond sums in mixed precision
    assert out["fp64"][2]["flux_stencil"] == 3 and out["mixed"][2]["flux_stencil"] == 2
    errs = field_errors(out["mixed"][0], out["fp64"][0])
    say("per-field (max, L2):", errs)
    assert np.array_equal(out["mixed"][1], out["fp64"][1])
    assert max(e[0] for e in errs) <= 1e-6 and max(e[1] for e in errs) <= 1e-6


def test_c4_screw128_neighbors_volumes_fp64_mixed():
    say("case C4 screw 128^3")
    case = nn.synth.screw(128, S=11, seed=0)
    assert 1.4e6 < case.n < 1.6e6
    rp, col, ml = oracle_nb(case)
    with nn.Problem(case.params, case.x, case.y, case.z, None, device=0) as pb:
        nnz, ml_g = pb.build_clouds_hash()
        rp_g, col_g = pb.neighbors()
        assert ml_g == ml and np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
        vol = pb.voronoi_volumes()
    deg = np.diff(rp)
    say("degree", deg.min(), deg.max(), "nnz", nnz)
    # Voronoi-approximate volumes: every probe of the box goes to one node within radius
    p = case.params
    assert np.all(vol > 0) and 0.5 * p.dx ** 3 < np.median(vol) < 1.5 * p.dx ** 3
    case.vol = vol
    out = {}
    for name, prec in (("fp64", nn.FP64), ("mixed", nn.MIXED)):
        with gpu_problem(case, rp, col, rtol=1e-10) as pb:
            st = pb.step(200, prec)
            assert st["status"] == 0 and st["steps"] == 200
            out[name] = (pb.get_state(), pb.get_phase(), st)
            say(name, "flips", st["flipped"], "cg", st["cg_iterations"])
    assert out["fp64"][2]["flipped"] > 0
    errs = field_errors(out["mixed"][0], out["fp64"][0])
    say("per-field (max, L2):", errs)
    assert np.array_equal(out["mixed"][1], out["fp64"][1])
    assert max(e[0] for e in errs) <= 1e-6


def _adjoint_setup(ng, seed=1):
    case = nn.synth.regular(ng, S=11, seed=seed)
    return case


def test_c5_adjoint64_checkpoints_and_fd():
    say("case C5 adjoint 64^3, 200 steps, checkpoints every 20")
    case = _adjoint_setup(64)
    S = case.params.S
    N = 200
    pb = gpu_problem(case, rtol=1e-13)
    try:
        start = pb.snapshot()
        Dl0 = torch.tensor([case.params.D_liquid[k] for k in range(S)], dtype=torch.float64)
        Ds0 = torch.tensor([case.params.D_solid[k] for k in range(S)], dtype=torch.float64)
        rd0 = torch.tensor(case.params.r_diss, dtype=torch.float64)
        sim = nn.autograd.simulate
        with torch.no_grad():
            target = sim(pb, Dl0 * 1.05, Ds0 * 1.05, rd0, N, start, 20).clone()

        def loss_of(U):
            return ((U[0] - target[0]) ** 2).sum() + ((U[S] - target[S]) ** 2).sum()

        grads, tape = {}, {}
        for every in (20, 1):
            Dl, Ds, rd = (t.clone().requires_grad_(True) for t in (Dl0, Ds0, rd0))
            loss = loss_of(sim(pb, Dl, Ds, rd, N, start, every))
            loss.backward()
            grads[every] = np.concatenate([Dl.grad.numpy(), Ds.grad.numpy(), [float(rd.grad)]])
            tape[every] = pb.last_adjoint_stats["tape_bytes"]
            say("every", every, "tape MiB", tape[every] / 2 ** 20)
        pb.tape_end()
        assert np.array_equal(grads[20], grads[1])
        assert tape[20] < tape[1] / 4
        g = grads[20]
        for name, k in (("D_liquid", 0), ("D_liquid", 5), ("D_solid", 0), ("D_solid", 7),
                        ("r_diss", None)):
            base = getattr(case.params, name) if k is None else getattr(case.params, name)[k]
            h = 1e-6 * abs(base)
            L = []
            for sgn in (1, -1):
                p = case.params.copy()
                if k is None:
                    setattr(p, name, base + sgn * h)
                else:
                    getattr(p, name)[k] = base + sgn * h
                pb.set_params(p)
                pb.restore(start)
                pb.step(N, nn.FP64)
                L.append(float(loss_of(torch.from_numpy(pb.get_state()))))
            fd = (L[0] - L[1]) / (2 * h)
            q = k if name == "D_liquid" else (S + k if name == "D_solid" else 2 * S)
            say(name, k, "adjoint", g[q], "fd", fd)
            assert abs(g[q] - fd) <= 1e-5 * abs(fd), (name, k, g[q], fd)
    finally:
        pb.close()


def test_c5_adjoint_vs_oracle_adjoint():
    say("case C5 GPU adjoint vs oracle adjoint, 16^3, 30 steps")
    case = _adjoint_setup(16)
    case.params.sub_steps = 4
    N = 30
    rp, col, _ = oracle_nb(case)
    rng = np.random.default_rng(11)
    lamN = rng.standard_normal(case.icvalues.shape)
    g_o, lam0_o, done, st = O.adjoint(case, rp, col, N, lamN, rtol=1e-13)
    assert st == 0 and done == N
    with gpu_problem(case, rp, col, rtol=1e-13) as pb:
        pb.tape_begin(5)
        pb.step(N, nn.FP64)
        g, lam0, ast = pb.tape_adjoint(lamN, want_lam0=True)
        pb.tape_end()
    assert ast["steps"] == N
    assert np.abs(g - g_o).max() <= 1e-9 * np.abs(g_o).max(), (g, g_o)
    assert rel_err(lam0, lam0_o) <= 1e-9
