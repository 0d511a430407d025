"""GPU adjoint (nnpp_tape_adjoint) against the oracle's reverse-mode adjoint, against central
finite differences of the GPU forward, and the torch.autograd binding.

Tolerances: GPU vs oracle gradients within 1e-9 of the largest entry (both FP64, solves at
rtol 1e-13); finite differences within 1e-5 relative (h = 1e-6 relative, the accuracy the
adjoint is specified to); checkpoint intervals bit-identical."""
import sys

import numpy as np
import pytest
import torch

from helpers import nn, orc

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_case(kind, seed=3, thr=1.0):
    if kind == "regular":
        case = nn.synth.regular(10, S=3, seed=seed)
    else:
        case = nn.synth.screw(12, S=3, seed=seed)
    case.params.sub_steps = 2  # several voln1rel (BV) updates
    if thr is not None:
        case.params.front_threshold = thr  # flips solid nodes on the way
    p = case.params
    st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    return case, rp, col


def gpu_problem(case, rp, col, rtol=1e-13):
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    pb.set_neighbors(rp, col)
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    pb.set_solver(rtol=rtol, maxit=100000)
    return pb


@pytest.mark.parametrize("kind", ["regular", "screw"])
def test_adjoint_matches_oracle(kind):
    say("case adjoint vs oracle", kind)
    case, rp, col = small_case(kind)
    N = 7
    rng = np.random.default_rng(1)
    lamN = rng.standard_normal(case.icvalues.shape)
    g_o, lam0_o, done, st = O.adjoint(case, rp, col, N, lamN, rtol=1e-13)
    assert st == 0 and done == N
    res = {}
    for every in (1, 3):
        with gpu_problem(case, rp, col) as pb:
            pb.tape_begin(every)
            stt = pb.step(N, nn.FP64)
            assert stt["status"] == 0 and stt["steps"] == N
            if every == 1:
                assert stt["flipped"] > 0
            U_end = pb.get_state()
            g, lam0, ast = pb.tape_adjoint(lamN, want_lam0=True)
            assert ast["steps"] == N
            assert ast["recomputed_steps"] == (0 if every == 1 else N - 3)
            assert np.array_equal(pb.get_state(), U_end)  # state left as it was
            g2, _, _ = pb.tape_adjoint(lamN)              # the tape can be reversed again
            assert np.array_equal(g, g2)
            pb.tape_end()
            res[every] = (g, lam0)
    g, lam0 = res[1]
    assert np.array_equal(res[3][0], g) and np.array_equal(res[3][1], lam0)
    scale = np.abs(g_o).max()
    assert np.abs(g - g_o).max() <= 1e-9 * scale, (g, g_o)
    assert np.abs(lam0 - lam0_o).max() <= 1e-9 * np.abs(lam0_o).max()


def test_adjoint_finite_differences():
    """dL/dtheta from the adjoint against central differences of the GPU forward, for every
    parameter (h = 1e-6 relative)."""
    say("case adjoint vs finite differences")
    case, rp, col = small_case("regular", seed=5)
    N = 6
    rng = np.random.default_rng(2)
    lamN = rng.standard_normal(case.icvalues.shape)
    S = case.params.S
    with gpu_problem(case, rp, col, rtol=1e-14) as pb:
        start = pb.snapshot()
        pb.tape_begin(2)
        pb.step(N, nn.FP64)
        g, _, _ = pb.tape_adjoint(lamN)
        pb.tape_end()
        names = [("D_liquid", k) for k in range(S)] + [("D_solid", k) for k in range(S)] + \
                [("r_diss", None)]
        fd = np.zeros_like(g)
        for q, (name, k) in enumerate(names):
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
                st = pb.step(N, nn.FP64)
                assert st["status"] == 0
                L.append(float((lamN * pb.get_state()).sum()))
            fd[q] = (L[0] - L[1]) / (2 * h)
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    say("adjoint", g, "fd", fd, "rel", rel)
    assert rel.max() <= 1e-5


def test_autograd_binding_mismatch_loss():
    """The torch.autograd.Function on the C5-style loss (final Mg and phi against a target made
    with diffusivities x1.05), checkpointed and full tape: same gradient, matches the adjoint
    on the same loss gradient through the oracle."""
    say("case autograd binding")
    case, rp, col = small_case("regular", seed=7, thr=None)
    S = case.params.S
    N = 8
    with gpu_problem(case, rp, col) as pb:
        start = pb.snapshot()
        Dl0 = torch.tensor([case.params.D_liquid[k] for k in range(S)], dtype=torch.float64)
        Ds0 = torch.tensor([case.params.D_solid[k] for k in range(S)], dtype=torch.float64)
        rd0 = torch.tensor(case.params.r_diss, dtype=torch.float64)
        with torch.no_grad():
            target = nn.autograd.simulate(pb, Dl0 * 1.05, Ds0 * 1.05, rd0, N, start, 1).clone()
        grads = {}
        for every in (1, 4):
            Dl = Dl0.clone().requires_grad_(True)
            Ds = Ds0.clone().requires_grad_(True)
            rd = rd0.clone().requires_grad_(True)
            U = nn.autograd.simulate(pb, Dl, Ds, rd, N, start, every)
            loss = ((U[0] - target[0]) ** 2).sum() + ((U[S] - target[S]) ** 2).sum()
            loss.backward()
            grads[every] = (Dl.grad.numpy().copy(), Ds.grad.numpy().copy(), float(rd.grad))
            assert pb.last_adjoint_stats["steps"] == N
        for a_, b_ in zip(grads[1], grads[4]):
            assert np.array_equal(np.asarray(a_), np.asarray(b_))
        pb.tape_end()
    # the same loss gradient through the oracle's adjoint
    U_o = O.run(case, rp, col, N, rtol=1e-13)["U"]
    lamN = np.zeros_like(U_o)
    lamN[0] = 2 * (U_o[0] - target[0].numpy())
    lamN[S] = 2 * (U_o[S] - target[S].numpy())
    g_o, _, _, st = O.adjoint(case, rp, col, N, lamN, rtol=1e-13)
    assert st == 0
    g = np.concatenate([grads[1][0], grads[1][1], [grads[1][2]]])
    assert np.abs(g - g_o).max() <= 1e-8 * np.abs(g_o).max(), (g, g_o)


def test_tape_guards():
    say("case tape guards")
    case, rp, col = small_case("regular", seed=2)
    with gpu_problem(case, rp, col) as pb:
        pb.tape_begin(1)
        with pytest.raises(nn.NnppError) as e:
            pb.set_state(case.icvalues)
        assert e.value.code == nn._abi.ERR_STATE
        with pytest.raises(nn.NnppError) as e:
            pb.step(1, nn.MIXED)
        assert e.value.code == nn._abi.ERR_UNSUPPORTED
        g, lam0, st = pb.tape_adjoint(np.ones(case.icvalues.shape), want_lam0=True)
        assert st["steps"] == 0 and np.all(g == 0) and np.all(lam0 == 1)  # empty tape
        pb.tape_end()
        pb.set_state(case.icvalues)  # allowed again
        with pytest.raises(nn.NnppError) as e:
            pb.tape_adjoint(np.ones(case.icvalues.shape))
        assert e.value.code == nn._abi.ERR_STATE


def test_autograd_stale_graph_raises():
    """A graph whose tape was replaced by a later simulate(), extended by further steps or
    ended gives an error in backward, never the gradient of another trajectory; a graph whose
    tape is intact still differentiates."""
    say("case autograd stale graph")
    case, rp, col = small_case("regular", seed=5, thr=None)
    S = case.params.S
    with gpu_problem(case, rp, col) as pb:
        start = pb.snapshot()
        Dl0 = torch.tensor([case.params.D_liquid[k] for k in range(S)], dtype=torch.float64)
        Ds0 = torch.tensor([case.params.D_solid[k] for k in range(S)], dtype=torch.float64)
        rd0 = torch.tensor(case.params.r_diss, dtype=torch.float64)

        def graph():
            Dl = Dl0.clone().requires_grad_(True)
            U = nn.autograd.simulate(pb, Dl, Ds0, rd0, 3, start, 1)
            return Dl, (U ** 2).sum()

        # (1) a second simulate() before backward replaces the first graph's tape
        _, l1 = graph()
        Dl2, l2 = graph()
        with pytest.raises(RuntimeError, match="replaced, extended or ended"):
            l1.backward()
        l2.backward()  # the current graph is intact
        assert torch.isfinite(Dl2.grad).all() and Dl2.grad.abs().max() > 0
        # (2) FP64 steps appended to the tape after forward
        _, l3 = graph()
        pb.step(1, nn.FP64)
        with pytest.raises(RuntimeError, match="replaced, extended or ended"):
            l3.backward()
        # (3) the tape ended
        _, l4 = graph()
        pb.tape_end()
        with pytest.raises(RuntimeError, match="replaced, extended or ended"):
            l4.backward()


def test_adjoint_leaves_state_on_noconv():
    """The adjoint restores the caller's state, phase and step scalars on every exit, also when
    its potential solves stop at the iteration limit (status NOCONV)."""
    say("case adjoint state restore")
    case, rp, col = small_case("regular", seed=4)
    for maxit in (100000, 2):
        with gpu_problem(case, rp, col) as pb:
            pb.set_solver(rtol=1e-13, maxit=maxit)  # frozen into the tape
            pb.tape_begin(2)
            st = pb.step(5, nn.FP64)
            assert st["status"] == (nn._abi.OK if maxit > 2 else nn._abi.ERR_NOCONV)
            U0, ph0, t0 = pb.get_state(), pb.get_phase(), pb.get_time()
            lam = np.ones(case.icvalues.shape)
            grad = np.zeros(2 * case.params.S + 1)
            code = pb.lib.nnpp_tape_adjoint(pb.ctx, nn._abi.dptr(lam), nn._abi.dptr(grad), None,
                                            None)
            assert code == (nn._abi.OK if maxit > 2 else nn._abi.ERR_NOCONV)
            assert np.array_equal(pb.get_state(), U0)
            assert np.array_equal(pb.get_phase(), ph0)
            assert pb.get_time() == t0
            pb.tape_end()
