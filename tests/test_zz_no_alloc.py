"""nnpp_implant_classify_noalloc (the classification with its scratch in a static device array,
no device allocation) against nnpp_implant_classify: counts bit-identical and initial values
equal on two calls in a row, the second with fewer points and a different radius (the array
keeps the first call's contents), plus its size limit and its overflow status."""
import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, nn, reference_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cloud(seed, ng, npts):
    p = reference_params()
    dx = p.dx
    rng = np.random.default_rng(seed)
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    jit = rng.uniform(-0.3 * dx, 0.3 * dx, size=(3, ng ** 3))
    x, y, z = (a.ravel() + j for a, j in zip((xx, yy, zz), jit))
    L = ng * dx
    u = rng.standard_normal((3, npts))
    u /= np.linalg.norm(u, axis=0)
    rad = rng.uniform(0.2, 0.35, npts) * L
    b = 0.5 * L + u * rad
    p.n_nodes = x.size
    return p, x, y, z, b


def _noalloc(pb, b, r, max_count=1024, want_ic=True):
    counts = np.empty(pb.n, dtype=np.int32)
    ic = np.empty((pb.F, pb.n)) if want_ic else None
    bb = [np.ascontiguousarray(a, dtype=np.float64) for a in b]
    st = pb.lib.nnpp_implant_classify_noalloc(pb.ctx, bb[0].size, nn._abi.dptr(bb[0]),
                                              nn._abi.dptr(bb[1]), nn._abi.dptr(bb[2]), r,
                                              max_count, nn._abi.iptr(counts),
                                              nn._abi.dptr(ic) if want_ic else None)
    return st, counts, ic


def test_noalloc_matches_allocating_entry_point():
    say("case implant classify without allocation")
    p, x, y, z, b = _cloud(11, 24, 60000)
    with nn.Problem(p, x, y, z, None, device=0) as pb:
        r1 = 3.0 * 0.4 * p.dx * (1.0 + 1e-5)
        # first call: every point; second: a subset and a larger radius
        for pts, r in ((b, r1), (b[:, : 7000], 2.5 * r1)):
            ref, ic_ref = pb.implant_classify(pts[0], pts[1], pts[2], r, want_ic=True)
            st, counts, ic = _noalloc(pb, pts, r)
            assert st == 0, pb.lib.nnpp_last_error(pb.ctx)
            assert np.array_equal(counts, ref) and np.array_equal(ic, ic_ref)
            assert counts.max() > 0 and (counts == 0).any()
            say("points", pts.shape[1], "solid nodes", int((counts > 0).sum()))
        # no points: every node liquid
        e0 = np.zeros(0)
        st, counts, _ = _noalloc(pb, (e0, e0, e0), r1, want_ic=False)
        assert st == 0 and not counts.any()
        # the overflow status of the allocating entry point, counts still written
        ref, _ = pb.implant_classify(b[0], b[1], b[2], r1)
        st, counts, _ = _noalloc(pb, b, r1, max_count=int(ref.max()) - 1, want_ic=False)
        assert st == nn._abi.ERR_OVERFLOW and np.array_equal(counts, ref)
        # beyond the scratch array's size: a clear error, nothing run
        big = np.zeros((3, (1 << 19) + 1))
        st, _, _ = _noalloc(pb, big, r1, want_ic=False)
        assert st == nn._abi.ERR_ARG
        assert b"NNPP_NOALLOC_MAX_POINTS" in pb.lib.nnpp_last_error(pb.ctx)


def test_noalloc_reference_golden():
    say("case implant classify without allocation, golden")
    d = np.load(os.path.join(GOLDEN, "ref_implant.npz"))
    p = reference_params()
    p.n_nodes = d["x"].size
    with nn.Problem(p, d["x"], d["y"], d["z"], None, device=0) as pb:
        st, counts, _ = _noalloc(pb, (d["bx"], d["by"], d["bz"]), float(d["r"]), want_ic=False)
        assert st == 0 and np.array_equal(counts, d["counts"])
