"""Edge cases through the C ABI on the GPU: empty node sets, isolated nodes and ragged rows,
all-solid and all-Dirichlet potentials, a single-tile lattice, against the oracle."""
import sys

import numpy as np
import pytest

from helpers import nn, orc, rel_err

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_both(case, nsteps, prec=nn.FP64, rtol=1e-13, fmt=None, oracle=True):
    p = case.params
    st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    ref = O.run(case, rp, col, nsteps, rtol=rtol) if oracle else None
    with nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0) as pb:
        if fmt is not None:
            pb.set_format(fmt)
        nnz, _ = pb.build_clouds_hash()
        assert nnz == rp[-1]
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=rtol, maxit=100000)
        st = pb.step(nsteps, prec)
        return ref, pb.get_state(), pb.get_phase(), st


def test_empty_node_set():
    say("case empty")
    p = nn.synth.synthetic_params(8, S=3)
    p.n_nodes = 0
    e = np.zeros(0)
    with nn.Problem(p, e, e, e, None, device=0) as pb:
        nnz, ml = pb.build_clouds_hash()
        assert nnz == 0
        pb.set_bc(np.zeros(0, np.int32), np.zeros((4, 0)))
        pb.init_state(np.zeros((4, 0)))
        st = pb.step(2, nn.FP64)
        # no solid node: the run stops after one step, as main.cpp:924-932 does
        assert st["status"] == 0 and st["steps"] == 1
        st = pb.step(1, nn.MIXED)
        assert st["status"] == 0
        assert pb.get_state().shape == (4, 0)


def test_isolated_nodes_and_ragged_rows():
    """A regular block plus far-away isolated nodes (rows of length 0) and a sparse cluster."""
    say("case isolated / ragged")
    case = nn.synth.regular(8, S=3, seed=5)
    rng = np.random.default_rng(0)
    extra = 40
    ex = rng.uniform(3.0, 9.0, size=(3, extra))  # far from the unit cube, mostly isolated
    ex[:, :8] = 5.0 + rng.uniform(0, 0.2, size=(3, 8))  # a small sparse cluster
    case.x = np.concatenate([case.x, ex[0]])
    case.y = np.concatenate([case.y, ex[1]])
    case.z = np.concatenate([case.z, ex[2]])
    F = case.params.S + 1
    case.bctypes = np.concatenate([case.bctypes, np.ones(extra, np.int32)])
    case.bcvalues = np.concatenate([case.bcvalues, -np.ones((F, extra))], axis=1)
    ic = rng.uniform(0.9, 1.1, size=(F, extra))
    ic[case.params.S] = 0.0
    case.icvalues = np.concatenate([case.icvalues, ic], axis=1)
    case.params.n_nodes = case.x.size
    for prec in (nn.FP64, nn.MIXED):
        ref, U, ph, st = run_both(case, 5, prec)
        assert st["status"] == 0
        assert np.array_equal(ph, ref["phase"])
        tol = 1e-12 if prec == nn.FP64 else 1e-6
        for f in range(F):
            assert rel_err(U[f], ref["U"][f]) < tol, (prec, f)


def test_all_solid_and_all_dirichlet_potential():
    say("case all-solid / all-Dirichlet")
    case = nn.synth.regular(6, S=3, seed=2)
    case.icvalues[0] = 1.0  # every node solid: no liquid bond, no potential unknown
    ref, U, ph, st = run_both(case, 3)
    assert st["status"] == 0 and np.all(ph == ref["phase"])
    for f in range(4):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, f
    case = nn.synth.regular(6, S=3, seed=2)
    case.bctypes[:] = 0  # every node Dirichlet
    case.bcvalues[:, :] = np.where(case.bcvalues == -1, 1.0, case.bcvalues)
    ref, U, ph, st = run_both(case, 2)
    assert st["status"] == 0
    for f in range(4):
        assert rel_err(U[f], ref["U"][f]) < 1e-12, f


@pytest.mark.parametrize("ng", [5, 7, 9, 16])
def test_small_lattices_tile_edges(ng):
    """Lattices smaller than one 256-row tile or ending inside one (stencil staging bounds)."""
    say("case lattice", ng)
    case = nn.synth.regular(ng, S=3, seed=ng)
    case.params.front_threshold = 1.0
    for prec in (nn.FP64, nn.MIXED):
        ref, U, ph, st = run_both(case, 4, prec)
        assert st["status"] == 0
        assert np.array_equal(ph, ref["phase"])
        tol = 1e-12 if prec == nn.FP64 else 1e-6
        for f in range(4):
            assert rel_err(U[f], ref["U"][f]) < tol, (ng, prec, f)


def exact_lattice(nx, ny, nz, S=3, seed=0):
    """nx x ny x nz lattice with spacing 1/16: every bond vector is exact in binary, so each
    offset j - i has one weight and the stencil forms (potential operator and flux) apply when
    no two bond vectors share an offset (nx, ny >= 7 for delta = 3 spacings). A solid ball in
    the middle (interface bonds in both passes of the flux), Dirichlet faces."""
    h = 1.0 / 16
    p = nn.synth.synthetic_params(16, S=S, seed=seed)
    p.n_nodes = nx * ny * nz
    zz, yy, xx = np.meshgrid((np.arange(nz) + 0.5) * h, (np.arange(ny) + 0.5) * h,
                             (np.arange(nx) + 0.5) * h, indexing="ij")
    x, y, z = xx.ravel(), yy.ravel(), zz.ravel()
    n, F = x.size, S + 1
    rng = np.random.default_rng(seed + 7)
    ic = np.empty((F, n))
    ic[:S] = rng.uniform(0.9, 1.1, size=(S, n))
    ic[S] = 0.0
    c = 0.5 * h * np.array([nx, ny, nz])
    r = 0.3 * h * min(nx, ny, nz)
    solid = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
    ic[0, ~solid] = 0.0
    I = np.stack(np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij"),
                 0).reshape(3, -1)
    lim = np.array([nz, ny, nx])[:, None] - 1
    face = ((I == 0) | (I == lim)).any(axis=0)
    bct = np.where(face, 0, 1).astype(np.int32)
    bcv = np.full((F, n), -1.0)
    bcv[:S, face] = 1.0
    bcv[S, face] = 0.0
    return nn.synth.Case(p, x, y, z, None, bct, bcv, ic, name=f"exact{nx}x{ny}x{nz}")


def _exact_run(shape, stencil):
    case = exact_lattice(*shape, seed=sum(shape))
    case.params.front_threshold = 1.0
    assert (case.icvalues[0] > 0).any()
    for prec in (nn.FP64, nn.MIXED):
        ref, U, ph, st = run_both(case, 4, prec)
        assert st["status"] == 0
        # the lattices at least 7 wide hold the whole radius-3 ball of offsets (full boxes:
        # the z-marching SpMV and, in FP64, the z-marching flux); the stencil flux sums in FP32
        # in mixed precision
        fs = (3 if prec == nn.FP64 else 2) if stencil else 0
        assert st["flux_stencil"] == fs and st["pot_stencil"] == 4 * int(stencil), (shape, prec)
        assert np.array_equal(ph, ref["phase"])
        tol = 1e-12 if prec == nn.FP64 else 1e-6
        for f in range(case.params.S + 1):
            assert rel_err(U[f], ref["U"][f]) < tol, (shape, prec, f)
        if stencil:
            # the tile SpMV of the lattice ball against the oracle, and the table-driven loop
            # gives its trajectory bit for bit
            _, U1, ph1, st1 = run_both(case, 4, prec, fmt=nn.FORMAT_NO_ZMARCH, oracle=False)
            assert st1["pot_stencil"] == 3
            assert np.array_equal(ph1, ref["phase"])
            for f in range(case.params.S + 1):
                assert rel_err(U1[f], ref["U"][f]) < tol, (shape, prec, f)
            _, U2, ph2, st2 = run_both(case, 4, prec, fmt=nn.FORMAT_NO_BALL, oracle=False)
            assert st2["pot_stencil"] == 1
            assert np.array_equal(ph2, ph1), (shape, prec)
            if prec == nn.FP64:
                assert np.array_equal(U2, U1), shape
            else:  # the mixed flux's table-driven walk groups its FP32 sums otherwise
                for f in range(case.params.S + 1):
                    assert rel_err(U2[f], ref["U"][f]) < tol, (shape, prec, f)


@pytest.mark.parametrize("shape", [(7, 7, 5), (7, 7, 7), (9, 9, 9), (16, 9, 7), (13, 13, 13)])
def test_stencil_forms_small_lattices(shape):
    """The stencil forms on lattices inside one 256-row tile (7x7x5) or ending inside one:
    staging windows clipped at both ends, partial waves, interface nodes; against the oracle."""
    say("case exact lattice", shape)
    _exact_run(shape, True)


def test_stencil_table_loop_when_not_the_radius3_ball():
    """delta = 2 spacings: the offsets are the radius-2 ball (32 of them), so the stencil forms
    apply but the SpMV's lattice-ball loop (radius 3) does not; the table-driven loop runs and
    matches the oracle."""
    say("case radius-2 lattice")
    case = exact_lattice(9, 9, 9, seed=5)
    case.params.delta = 2.0 / 16
    case.params.front_threshold = 1.0
    for prec in (nn.FP64, nn.MIXED):
        ref, U, ph, st = run_both(case, 3, prec)
        assert st["status"] == 0
        assert st["pot_stencil"] == 1, prec
        assert st["flux_stencil"] == (1 if prec == nn.FP64 else 2), prec
        assert np.array_equal(ph, ref["phase"])
        tol = 1e-12 if prec == nn.FP64 else 1e-6
        for f in range(case.params.S + 1):
            assert rel_err(U[f], ref["U"][f]) < tol, (prec, f)


def test_stencil_refused_when_offsets_alias():
    """In a lattice 4 nodes wide, (3,0,0) and (-1,1,0) are both bonds with the offset j - i = 3
    and different weights: the per-offset weight check rejects the stencil forms, and the
    column forms run."""
    say("case aliasing lattice")
    _exact_run((4, 4, 4), False)


@pytest.mark.parametrize("shape", [(9, 9, 9), (13, 10, 8), (16, 9, 7), (20, 11, 12)])
def test_potential_operator_ring_form(shape):
    """The FP32 potential operator of the z-marching walk (bonds summed by rings of equal
    in-plane distance, differences from each plane's column value), with the staging in 16-byte
    quads (nx a multiple of 4) and in single values (other nx), partial column tiles: FP64
    against the oracle's K11 action (1e-12), FP32 within 4e-7 of the largest entry on a rough
    vector, and on a smooth one no worse than four times the FP32 error of the bond-by-bond tile
    form (NNPP_FORMAT_NO_ZMARCH)."""
    say("case ring operator", shape)
    case = exact_lattice(*shape, seed=sum(shape) + 1)
    p = case.params
    st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    U, ph, _, _ = O.init_state(case)
    S = p.S
    liquid = ph == 0
    rows = np.repeat(np.arange(case.n), np.diff(rp))
    has_ll = np.zeros(case.n, bool)
    has_ll[rows[liquid[rows] & liquid[col]]] = True
    dirphi = (case.bctypes == 0) & (case.bcvalues[S] != -1)
    free = liquid & has_ll & ~dirphi
    rng = np.random.default_rng(sum(shape))
    rough = np.where(free, rng.standard_normal(case.n), 0.0)
    smooth = np.where(free, np.sin(7.0 * case.x) * np.cos(5.0 * case.y) + case.z ** 2, 0.0)
    out = {}
    for fmt in (nn.FORMAT_AUTO, nn.FORMAT_NO_ZMARCH):
        with nn.Problem(p, case.x, case.y, case.z, case.vol, device=0) as pb:
            pb.set_format(fmt)
            pb.set_neighbors(rp, col)
            pb.set_bc(case.bctypes, case.bcvalues)
            pb.init_state(case.icvalues)
            out[fmt] = [(pb.potential_operator_apply(x, nn.FP64),
                         pb.potential_operator_apply(x, nn.MIXED)) for x in (rough, smooth)]
    for k, x in enumerate((rough, smooth)):
        yo = np.where(free, O.poisson_apply(case, rp, col, ph, x), 0.0)
        y64, y32 = out[nn.FORMAT_AUTO][k]
        t64, t32 = out[nn.FORMAT_NO_ZMARCH][k]
        assert rel_err(y64, yo) < 1e-12, k
        assert np.array_equal(y64, t64), k
        scale = np.abs(y64).max()
        err = np.abs(y32 - y64).max()
        if k == 0:
            assert err <= 4e-7 * scale, (shape, err / scale)
        else:
            assert err <= 4.0 * max(np.abs(t32 - y64).max(), 1e-7 * scale), (
                shape, err / scale, np.abs(t32 - y64).max() / scale)
