"""CPU checks of the implant classification: the oracle's brute force against the reference's
own kd-tree build_clouds (kd_tree_interface.h:156-252, golden counts in
tests/golden/ref_implant.npz, whose implant points include points at distance r (1 + a few ulps)
of nodes, so that sqrt(s) <= r and s <= r^2 give different answers for 31 pairs), and the
initial values of main.cpp:422-457."""
import os

import numpy as np

from helpers import GOLDEN, nn, orc, reference_params

O = orc.Oracle(nn.Params)


def test_implant_counts_match_reference_kdtree():
    d = np.load(os.path.join(GOLDEN, "ref_implant.npz"))
    st, c = O.implant_classify(d["x"], d["y"], d["z"], d["bx"], d["by"], d["bz"], float(d["r"]))
    assert st == 0 and np.array_equal(c, d["counts"])
    # the squared-distance form would disagree on this fixture: the predicate is pinned
    dx = d["bx"][None, :] - d["x"][:, None]
    dy = d["by"][None, :] - d["y"][:, None]
    dz = d["bz"][None, :] - d["z"][:, None]
    sq = ((dx * dx + dy * dy + dz * dz) <= float(d["r"]) ** 2).sum(1)
    assert not np.array_equal(sq, d["counts"])


def test_implant_overflow_and_empty():
    d = np.load(os.path.join(GOLDEN, "ref_implant.npz"))
    st, c = O.implant_classify(d["x"], d["y"], d["z"], d["bx"], d["by"], d["bz"], float(d["r"]),
                               max_count=int(d["counts"].max()) - 1)
    assert st == nn._abi.ERR_OVERFLOW
    e = np.zeros(0)
    st, c = O.implant_classify(d["x"], d["y"], d["z"], e, e, e, float(d["r"]))
    assert st == 0 and not c.any()


def test_implant_initial_values_table():
    """main.cpp:422-457, value by value."""
    p = reference_params()
    CS = p.CSolid
    counts = np.array([0, 3, 0, 1], dtype=np.int32)
    ic = O.implant_icvalues(p, counts)
    liquid = [0.0, 1e-4 / CS, 1e-4 / CS, p.C0HCO3 / CS, 0.0, p.C0HPO4 / CS, 0.0, p.C0Ca / CS,
              p.C0Na / CS, p.C0Cl / CS, 0.0]
    solid = [CS / CS] + [0.0] * 10
    for i, c in enumerate(counts):
        assert list(ic[:, i]) == (liquid if c == 0 else solid)
    # the phase rule of nnpp_init_state (U_Mg > EPS) turns them into the reference's phases
    assert np.array_equal((ic[0] > p.phase_eps).astype(int), (counts > 0).astype(int))


def _lattice(ng, dx, jitter=0.0, seed=0):
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([xx, yy, zz]).reshape(3, -1)
    if jitter:
        P = P + np.random.default_rng(seed).uniform(-jitter * dx, jitter * dx, P.shape)
    return P[0].copy(), P[1].copy(), P[2].copy()


def test_voronoi_volumes_closed_forms():
    """A lattice with margin dx/2 and probes dx/3 apart: every node owns exactly its 27 probes
    (its cube, dx^3); a jittered cloud with a large radius assigns every probe once."""
    dx = 2e-5
    x, y, z = _lattice(6, dx)
    h = dx / 3
    v = O.voronoi_volumes(x, y, z, h, dx / 2, dx)
    assert np.all(v == 27 * h ** 3) and abs(v[0] - dx ** 3) <= 1e-12 * dx ** 3
    x, y, z = _lattice(7, dx, jitter=0.3, seed=2)
    v = O.voronoi_volumes(x, y, z, h, dx / 2, 10 * dx)
    ext = [(a.max() + dx / 2) - (a.min() - dx / 2) for a in (x, y, z)]
    nprobe = np.prod([round(e / h) for e in ext])
    assert abs(v.sum() / h ** 3 - nprobe) < 1e-6 and np.all(v > 0)
