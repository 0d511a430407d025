"""GPU implant classification (nnpp_implant_classify) against the reference's kd-tree counts
(golden) and the oracle's brute force: counts bit-exact, initial values equal."""
import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, nn, orc, reference_params

pytestmark = pytest.mark.gpu

O = orc.Oracle(nn.Params)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(x, y, z):
    p = reference_params()
    p.n_nodes = x.size
    return nn.Problem(p, x, y, z, None, device=0), p


def test_implant_counts_vs_reference_golden():
    say("case implant golden")
    d = np.load(os.path.join(GOLDEN, "ref_implant.npz"))
    pb, p = _problem(d["x"], d["y"], d["z"])
    with pb:
        counts, ic = pb.implant_classify(d["bx"], d["by"], d["bz"], float(d["r"]), want_ic=True)
        assert np.array_equal(counts, d["counts"])
        assert np.array_equal(ic, O.implant_icvalues(p, counts))
        with pytest.raises(nn.NnppError) as e:
            pb.implant_classify(d["bx"], d["by"], d["bz"], float(d["r"]),
                                max_count=int(d["counts"].max()) - 1)
        assert e.value.code == nn._abi.ERR_OVERFLOW
        e0 = np.zeros(0)
        c0, _ = pb.implant_classify(e0, e0, e0, float(d["r"]))
        assert not c0.any()
        # the initial values feed nnpp_init_state: the reference's phases and solid count
        pb.init_state(ic)
        assert np.array_equal(pb.get_phase(), (counts > 0).astype(np.int32))


def test_implant_counts_vs_oracle_random_cloud():
    say("case implant random cloud")
    rng = np.random.default_rng(3)
    p = reference_params()
    dx = p.dx
    ng = 20
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    jit = rng.uniform(-0.3 * dx, 0.3 * dx, size=(3, ng ** 3))
    x, y, z = (a.ravel() + j for a, j in zip((xx, yy, zz), jit))
    L = ng * dx
    # a shell-shaped implant partly outside the node box, and a far cluster
    u = rng.standard_normal((3, 40000))
    u /= np.linalg.norm(u, axis=0)
    rad = 0.35 * L + rng.uniform(-0.05, 0.05, 40000) * L
    P = np.concatenate([L / 2 + u * rad, 10 * L + rng.uniform(0, L, (3, 500))], axis=1)
    bx, by, bz = P
    r = 3. * 0.4 * dx * (1.00 + 1e-5)
    st, co = O.implant_classify(x, y, z, bx, by, bz, r)
    assert st == 0
    pb, _ = _problem(x, y, z)
    with pb:
        counts, _ = pb.implant_classify(bx, by, bz)  # default r: main.cpp:317
    assert np.array_equal(counts, co) and (co > 0).sum() > 100


def test_cpp_driver_bunny(tmp_path):
    """nnpp_main --bunny: initial values from the implant points (main.cpp:236-240, 316-325,
    422-457), then explicit steps, against the oracle from the same initial values."""
    import subprocess
    say("case nnpp_main --bunny")
    case = nn.synth.regular(10, S=10, seed=4)
    p = reference_params()
    case.params = p
    s = p.dx * 10
    case.x, case.y, case.z = case.x * s, case.y * s, case.z * s
    p.n_nodes = case.n
    p.n_steps = 4
    p.dt = 1.0
    p.sub_steps = 2  # outputs after steps 2 and 4
    case.bcvalues[:10][case.bcvalues[:10] != -1] = 1e-3 * p.CSolid
    rng = np.random.default_rng(8)
    u = rng.standard_normal((3, 3000))
    u /= np.linalg.norm(u, axis=0)
    bx, by, bz = s / 2 + u * (0.3 * s * rng.uniform(0, 1, 3000) ** (1 / 3))
    job = str(tmp_path / "NNPP")
    nn.formats.write_job(job, case)
    d1 = open(job + ".data1").read().split()
    d1[3] = str(bx.size)
    open(job + ".data1", "w").write("\n".join(d1) + "\n")
    with open(job + ".bunnyNodes", "w") as f:
        for a, b, c in zip(bx, by, bz):
            f.write("%.17g %.17g %.17g\n" % (a, b, c))
    exe = os.path.join(os.path.dirname(nn._abi.LIB_PATH), "bin", "nnpp_main")
    r = subprocess.run([exe, "-01", job, "0", "--bunny", "--rtol", "1e-13", "--steps", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Building bunnyNodes took" in r.stdout
    st, counts = O.implant_classify(case.x, case.y, case.z, bx, by, bz,
                                    3. * 0.4 * p.dx * (1.00 + 1e-5))
    assert st == 0 and (counts > 0).any() and (counts == 0).any()
    case.icvalues = O.implant_icvalues(p, counts)
    rp, col, _ = _oracle_nb(case)
    ref = O.run(case, rp, col, 4, rtol=1e-13)
    U = np.zeros((11, case.n))
    ph = np.zeros(case.n, dtype=np.int32)
    assert nn.lib.nnpp_read_start_csv((job + "-01-4.output.csv").encode(), case.n, 10,
                                      nn._abi.dptr(U), nn._abi.iptr(ph)) == 0
    assert np.array_equal(ph, ref["phase"])
    for f in range(11):
        assert np.abs(U[f] - ref["U"][f]).max() <= 1e-10 * max(np.abs(ref["U"][f]).max(), 1e-300)


def _oracle_nb(case):
    p = case.params
    st, rp, col, ml = O.build_neighbors(case.x, case.y, case.z, p.delta * (1.00 + 1e-5),
                                        p.dx * 10., p.dy * 10., p.dz * 10.)
    assert st == 0
    return rp, col, ml


@pytest.mark.parametrize("jitter,res", [(0.0, 3), (0.3, 3), (0.45, 2)])
def test_voronoi_volumes_vs_oracle(jitter, res):
    say("case voronoi", jitter, res)
    p = reference_params()
    dx = p.dx
    ng = 9
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([xx, yy, zz]).reshape(3, -1)
    P = P + np.random.default_rng(5).uniform(-jitter * dx, jitter * dx, P.shape)
    x, y, z = (a.copy() for a in P)
    h = dx / res
    vo = O.voronoi_volumes(x, y, z, h, dx / 2, 2 * dx)
    pb, _ = _problem(x, y, z)
    with pb:
        v = pb.voronoi_volumes(h=h, margin=dx / 2, radius=2 * dx)
    assert np.array_equal(v, vo)
    if jitter == 0.0:
        assert np.all(v == res ** 3 * h ** 3)
