"""torch.autograd binding of the explicit step and its hand-written adjoint.

`simulate(problem, D_liquid, D_solid, r_diss, nsteps, start, checkpoint_every)` returns the
state U^N (F, n) float64 after `nsteps` FP64 steps from the snapshot `start`
(Problem.snapshot()). Its backward calls nnpp_tape_adjoint, the hand-written adjoint of the
flux, potential and explicit-update kernels, recomputing each segment of `checkpoint_every`
steps from its checkpoint, and returns dL/dD_liquid, dL/dD_solid and dL/dr_diss (generic
chemistry). Only the parameters are differentiable; the start state is a constant.
The state is exchanged with the library through host arrays (the C ABI takes host pointers).

A Problem holds one tape, and the backward of a graph reverses that tape. The graph remembers
the tape's generation, the number of steps it recorded and the parameters; backward raises
RuntimeError instead of returning the gradient of another trajectory when a later simulate()
replaced the tape, when further steps were appended to it, or when the tape was ended.
"""
import numpy as np
import torch

from ._abi import FP64


class NnppSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D_liquid, D_solid, r_diss, problem, nsteps, start, checkpoint_every):
        S = problem.S
        p = problem.params.copy()
        dl = D_liquid.detach().to("cpu", torch.float64).numpy()
        ds = D_solid.detach().to("cpu", torch.float64).numpy()
        if dl.shape != (S,) or ds.shape != (S,):
            raise ValueError(f"D_liquid and D_solid must have shape ({S},)")
        for k in range(S):
            p.D_liquid[k] = float(dl[k])
            p.D_solid[k] = float(ds[k])
        p.r_diss = float(r_diss.detach().to("cpu", torch.float64))
        problem.tape_end()
        problem.set_params(p)
        problem.restore(start)
        problem.tape_begin(checkpoint_every)
        st = problem.step(nsteps, FP64)
        if st["status"] != 0:
            raise RuntimeError(f"forward steps ended with status {st['status']}")
        ctx.problem = problem
        ctx.device = D_liquid.device
        ctx.tape = problem.tape_info()
        ctx.params = (dl.copy(), ds.copy(), p.r_diss)
        return torch.from_numpy(problem.get_state()).to(D_liquid.device)

    @staticmethod
    def backward(ctx, gU):
        pb = ctx.problem
        S = pb.S
        if pb.tape_info() != ctx.tape:
            raise RuntimeError(
                "the differentiation tape of this graph was replaced, extended or ended "
                f"(recorded tape {ctx.tape}, problem now holds {pb.tape_info()}); run simulate() "
                "again before backward")
        cur = pb.params
        dl, ds, rd = ctx.params
        if (any(cur.D_liquid[k] != dl[k] or cur.D_solid[k] != ds[k] for k in range(S))
                or cur.r_diss != rd):
            raise RuntimeError("the problem's parameters changed since this graph's forward")
        grad, _, stats = pb.tape_adjoint(gU.detach().to("cpu", torch.float64).numpy())
        ctx.problem.last_adjoint_stats = stats
        g = torch.from_numpy(grad).to(ctx.device)
        return g[:S], g[S:2 * S], g[2 * S], None, None, None, None


def simulate(problem, D_liquid, D_solid, r_diss, nsteps, start, checkpoint_every=1):
    """U^N as a differentiable function of (D_liquid, D_solid, r_diss); see the module doc."""
    if not torch.is_tensor(r_diss):
        r_diss = torch.tensor(float(r_diss), dtype=torch.float64)
    return NnppSteps.apply(D_liquid, D_solid, r_diss, problem, int(nsteps), start,
                           int(checkpoint_every))
