"""CPU checks that pin the oracle: reference golden vectors (from the reference's own grid.h /
utils.h, tests/golden/*.npz), brute force, closed forms and invariants. No GPU needed."""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import GOLDEN, ell_to_sorted_csr, nn, orc, ref_chem_case, reference_params, rel_err

O = orc.Oracle(nn.Params)


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    p = reference_params()
    n = d["x"].size
    p.n_nodes = n
    case = nn.synth.Case(p, d["x"], d["y"], d["z"], None, d["bctypes"], d["bcvalues"],
                         d["Un"].copy())
    return d, case


# ---------------------------------------------------------------- Beta (utils.h:17-25) ------
def test_beta_golden_and_closed_form():
    d, case = load_golden("reg7")
    p = case.params
    got = np.array([O.beta(v, p.delta, p.dx) for v in d["beta_xi"]])
    assert np.array_equal(got, d["beta"])
    # lattice values at delta = 3 dx (independent closed form)
    dx, delta = 1.0, 3.0
    assert O.beta(math.sqrt(6.0), delta, dx) == 1.0
    assert abs(O.beta(math.sqrt(8.0), delta, dx) - (3.5 - math.sqrt(8.0))) < 1e-15
    assert abs(O.beta(math.sqrt(8.0), delta, dx) - 0.671573) < 1e-6
    assert O.beta(3.0, delta, dx) == 0.5
    assert O.beta(3.6, delta, dx) == 0.0


# ---------------------------------------------------------------- neighbours (grid.h) -------
@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_neighbors_match_reference(name):
    d, case = load_golden(name)
    p = case.params
    st, rp, col, ml = O.build_neighbors(case.x, case.y, case.z, p.delta * (1.00 + 1e-5),
                                        p.dx * 10., p.dy * 10., p.dz * 10.)
    assert st == 0
    assert np.array_equal(rp, d["row_ptr"]) and np.array_equal(col, d["col"])
    assert ml == int(d["max_len"])


def brute_force(x, y, z, r):
    n = x.size
    rows = []
    for i in range(n):
        d2 = (x[i] - x) ** 2 + (y[i] - y) ** 2 + (z[i] - z) ** 2  # same FP64 expression, no FMA
        m = d2 <= r * r
        m[i] = False
        rows.append(np.nonzero(m)[0])
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r_) for r_ in rows])
    return rp, np.concatenate(rows).astype(np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbors_brute_force_random_cloud(seed):
    rng = np.random.default_rng(seed)
    n = 600
    x, y, z = rng.uniform(0, 1, (3, n))
    r = 0.17
    st, rp, col, _ = O.build_neighbors(x, y, z, r, 0.5, 0.4, 0.3)
    assert st == 0
    rb, cb = brute_force(x, y, z, r)
    assert np.array_equal(rp, rb) and np.array_equal(col, cb)


def test_neighbors_symmetric_and_lattice_counts():
    for ng, nnz in ((24, 1_455_112), (32, 3_581_672)):
        c = nn.synth.regular(ng, S=2)
        p = c.params
        st, rp, col, ml = O.build_neighbors(c.x, c.y, c.z, p.delta * (1.00 + 1e-5), p.dx * 10.,
                                            p.dy * 10., p.dz * 10.)
        assert st == 0 and int(rp[-1]) == nnz and ml == 123
        if ng == 24:
            A = sp.csr_matrix((np.ones(col.size), col, rp), shape=(c.n, c.n))
            assert (A != A.T).nnz == 0


def test_neighbors_overflow_and_empty():
    c = nn.synth.regular(8, S=1)
    p = c.params
    st, *_ = O.build_neighbors(c.x, c.y, c.z, p.delta * (1.00 + 1e-5), p.dx * 10., p.dy * 10.,
                               p.dz * 10., max_len=100)
    assert st == nn._abi.ERR_OVERFLOW
    st, rp, col, ml = O.build_neighbors(np.zeros(0), np.zeros(0), np.zeros(0), 0.1, 1, 1, 1)
    assert st == 0 and rp.tolist() == [0] and ml == 0
    one = np.array([0.5])
    st, rp, col, ml = O.build_neighbors(one, one, one, 0.1, 1, 1, 1)
    assert st == 0 and rp.tolist() == [0, 0] and ml == 1


def test_ell_adapter_matches_reference_layout():
    d, case = load_golden("jit7")
    n = case.n
    ml = int(d["max_len"])
    ell = np.zeros(n * ml, dtype=np.int32)
    ln = np.zeros(n, dtype=np.int32)
    rp, col = d["row_ptr"], d["col"]
    st = nn.lib.nnpp_neighbors_to_ell(n, nn._abi.lptr(rp), nn._abi.iptr(col), ml,
                                      nn._abi.iptr(ell), nn._abi.iptr(ln))
    assert st == 0
    assert np.array_equal(ln, d["lengths"])
    E = ell.reshape(ml, n).T  # column-major n x max_len
    assert np.array_equal(E[:, 0], np.arange(n))
    rp2, col2 = ell_to_sorted_csr(E, ln)
    assert np.array_equal(rp2, rp) and np.array_equal(col2, col)
    assert nn.lib.nnpp_neighbors_to_ell(n, nn._abi.lptr(rp), nn._abi.iptr(col), ml - 1,
                                        nn._abi.iptr(ell), nn._abi.iptr(ln)) == nn._abi.ERR_OVERFLOW


def _from_ell(E, ln):
    n, ml = E.shape
    flat = np.ascontiguousarray(E.T).ravel()  # column-major n x max_len
    rp = np.zeros(n + 1, dtype=np.int64)
    st = nn.lib.nnpp_neighbors_from_ell(n, ml, nn._abi.iptr(flat), nn._abi.iptr(ln),
                                        nn._abi.lptr(rp), None)
    if st:
        return st, None, None
    col = np.zeros(max(int(rp[-1]), 1), dtype=np.int32)
    st = nn.lib.nnpp_neighbors_from_ell(n, ml, nn._abi.iptr(flat), nn._abi.iptr(ln),
                                        nn._abi.lptr(rp), nn._abi.iptr(col))
    return st, rp, col[: int(rp[-1])]


def test_ell_to_csr_adapter():
    """nnpp_neighbors_from_ell: the reference's horizons layout (self first, neighbours in
    emission order, -1 padding) back to id-sorted CSR."""
    d, case = load_golden("jit7")
    n, ml = case.n, int(d["max_len"])
    rp, col = d["row_ptr"], d["col"]
    rng = np.random.default_rng(3)
    E = np.full((n, ml), -1, dtype=np.int32)
    E[:, 0] = np.arange(n)
    for i in range(n):  # neighbours in an arbitrary (emission) order
        E[i, 1:d["lengths"][i]] = rng.permutation(col[rp[i]:rp[i + 1]])
    st, rp2, col2 = _from_ell(E, d["lengths"].astype(np.int32))
    assert st == 0 and np.array_equal(rp2, rp) and np.array_equal(col2, col)
    if orc.RefHarness.available():  # the reference's own emission order
        p = case.params
        st, ell, ln, _ = orc.RefHarness().build_clouds_hash(case.x, case.y, case.z,
                                                            p.delta * (1.00 + 1e-5), p.dx * 10,
                                                            p.dy * 10, p.dz * 10)
        st, rp3, col3 = _from_ell(ell, ln)
        assert st == 0 and np.array_equal(rp3, rp) and np.array_equal(col3, col)
    bad = E.copy()
    bad[5, 0] = 6  # self not in column 0
    assert _from_ell(bad, d["lengths"].astype(np.int32))[0] == nn._abi.ERR_ARG
    bad = E.copy()
    bad[5, 2] = bad[5, 1]  # repeated id
    assert _from_ell(bad, d["lengths"].astype(np.int32))[0] == nn._abi.ERR_ARG


# ---------------------------------------------------------------- residuals (utils.h) -------
@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_residuals_match_reference(name):
    d, case = load_golden(name)
    F = O.residuals(case, d["row_ptr"], d["col"], d["phase"], d["U1"], d["Un"],
                    float(d["voln1rel"]))
    for f in range(11):
        assert rel_err(F[f], d["F"][f]) < 1e-13, f


@pytest.mark.parametrize("name", ["reg7", "jit7"])
def test_poisson_operator_matches_reference_K11(name):
    d, case = load_golden(name)
    n = case.n
    K = sp.csr_matrix((d["K11_vals"], (d["K11_rows"], d["K11_cols"])), shape=(n, n))
    rng = np.random.default_rng(7)
    v = rng.standard_normal(n)
    y = O.poisson_apply(case, d["row_ptr"], d["col"], d["phase"], v)
    phase = d["phase"]
    rp, col = d["row_ptr"], d["col"]
    S = 10
    dirphi = (case.bctypes == 0) & (case.bcvalues[S] != -1)
    hasll = np.array([phase[i] == 0 and (phase[col[rp[i]:rp[i + 1]]] == 0).any() for i in range(n)])
    rows = hasll & ~dirphi  # K11 rows that are not identity rows (utils.h:1802-1805,1838-1843)
    Kv = K @ v
    assert rel_err(y[rows], Kv[rows]) < 1e-13
    assert np.allclose(Kv[~rows], v[~rows], rtol=0, atol=0)


def test_residual_bc_override():
    d, case = load_golden("reg7")
    F = d["F"]
    Fb = O.bc_override(case, d["U1"], F)
    S = 10
    for f in range(11):
        m = (case.bctypes == 0) & (case.bcvalues[f] != -1)
        exp = d["U1"][f] - (case.bcvalues[f] / case.params.CSolid if f < S else case.bcvalues[f])
        assert np.array_equal(Fb[f][m], exp[m])
        assert np.array_equal(Fb[f][~m], F[f][~m])


def _lattice_case(ng=10, S=3, liquid=True):
    c = nn.synth.regular(ng, S=S, sphere_radius=0.0 if liquid else 0.3, r_diss=0.0)
    p = c.params
    st, rp, col, _ = O.build_neighbors(c.x, c.y, c.z, p.delta * (1.00 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    return c, rp, col


def test_residual_closed_forms():
    c, rp, col = _lattice_case()
    p = c.params
    S, n = p.S, c.n
    phase = np.zeros(n, dtype=np.int32)
    # constant fields: every bond sum vanishes exactly
    U = np.ones((S + 1, n))
    F = O.residuals(c, rp, col, phase, U, U)
    assert np.abs(F[:S]).max() == 0.0
    # linear concentration, constant phi: zero on nodes whose horizon is complete
    h = p.dx
    g = np.round(np.stack([c.x, c.y, c.z]) / h - 0.5).astype(int)
    ng = 10
    interior = ((g >= 3) & (g <= ng - 4)).all(0)
    U[0] = 2.0 + 0.3 * c.x - 0.7 * c.y + 0.2 * c.z
    F = O.residuals(c, rp, col, phase, U, U)
    scale = np.abs(p.K1[0] * p.D_liquid[0] * 0.7 * h)
    assert np.abs(F[0][interior]).max() < 1e-9 * scale * 1e3
    # quadratic concentration: exact lattice sum  -K1 D sum_j beta V (xj-xi)^2 / xi^3
    U[0] = c.x ** 2
    F = O.residuals(c, rp, col, phase, U, U)
    offs = [(a, b, e) for a in range(-3, 4) for b in range(-3, 4) for e in range(-3, 4)
            if 0 < a * a + b * b + e * e <= 9]
    r = 3.0 * (1.0 + 1e-5)
    lat = 0.0
    for a, b, e in offs:
        xi = math.sqrt(a * a + b * b + e * e)
        if xi > r:
            continue
        be = 1.0 if xi <= 2.5 else (3.5 - xi)
        lat += be * (a * h) ** 2 / (xi * h) ** 3
    expect = -p.K1[0] * p.D_liquid[0] * lat * h ** 3
    assert rel_err(F[0][interior], np.full(interior.sum(), expect)) < 1e-12
    # and K = 3/(pi delta^2) makes the operator approximate -D * laplacian (= -2 D here)
    assert abs(expect / (-2.0 * p.D_liquid[0]) - 1.0) < 0.15
    # migration with constant concentration and linear phi vanishes on interior nodes
    U[0] = 1.0
    U[S] = 0.5 * c.x + 0.1 * c.z
    F = O.residuals(c, rp, col, phase, U, U)
    m_scale = abs(p.K2[0] * p.D_liquid[0] * p.z[0]) * 0.5 * h
    assert np.abs(F[0][interior]).max() < 1e-9 * m_scale * 1e3


def test_explicit_step_conserves_mass_without_sources():
    c, rp, col = _lattice_case(ng=9, S=2)
    p = c.params
    n = c.n
    c.bctypes[:] = 1  # no Dirichlet nodes
    rng = np.random.default_rng(3)
    U = rng.uniform(0.5, 1.5, size=(p.S + 1, n))
    U[p.S] = 0.0  # phi const -> no migration, no charge solve effect on species
    st = {"U": U.copy(), "phase": np.zeros(n, np.int32), "step": 1, "voln1rel": 0.0,
          "vol0rel": 1.0, "solid": 0}
    p.clamp_floor = -1e300
    out = O.run(c, rp, col, 1, state=st)
    m0 = U[:p.S].sum(1)
    m1 = out["U"][:p.S].sum(1)
    assert rel_err(m1, m0) < 1e-13


def test_poisson_solve_matches_direct_solve():
    d, case = load_golden("jit7")
    n = case.n
    rp, col, phase = d["row_ptr"], d["col"], d["phase"]
    U = d["Un"]
    phi, it, rr, st = O.poisson_solve(case, rp, col, phase, U, np.zeros(n), rtol=1e-14)
    assert st == 0 and rr <= 1e-14
    # direct: K11 (reference matrix) phi = rhs with rhs from the residual at phi = 0
    K = sp.csr_matrix((d["K11_vals"], (d["K11_rows"], d["K11_cols"])), shape=(n, n)).tocsc()
    U0 = U.copy()
    U0[10] = 0.0
    F0 = O.residuals(case, rp, col, phase, U0, U)
    F0 = O.bc_override(case, U0, F0)
    x = spla.spsolve(K, -F0[10])
    assert rel_err(phi, x) < 1e-10


def test_reference_params_read():
    p = reference_params()
    assert p.n_nodes == 4393605 and p.n_steps == 4480
    assert p.dt == 10800.0 and p.dx == 0.00002 and p.CSolid == 66082.4
    assert p.K1[0] == 2.6525823848649216e8 and p.z[0] == 2.0 and p.z[9] == -1.0
    assert p.sigma1 == 6.1e6 and p.vol_red == 4.79859938096492
    assert p.front_threshold == p.CSat / p.CSolid
    assert p.clamp[0] == nn._abi.CLAMP_ALWAYS and p.clamp[9] == nn._abi.CLAMP_LIQUID


def test_init_state_rules():
    c = nn.synth.regular(10, S=3)
    U, ph, sc, v0 = O.init_state(c)
    p = c.params
    assert np.array_equal(ph, (c.icvalues[0] > p.phase_eps).astype(np.int32))
    m = c.bctypes == 0
    assert np.all(U[:p.S, m] == 1.0) and np.all(U[p.S, m] == 0.0)
    # main.cpp:476-479 sums the initial values, before any Dirichlet value is written
    assert abs(v0 - (c.icvalues[0][ph == 1] - p.CSat / p.CSolid).sum()) < 1e-12 * abs(v0)
    c2 = nn.synth.regular(6, S=2)
    c2.icvalues[0][:] = 3.0  # every node solid, the Dirichlet faces included
    c2.bcvalues[0][c2.bctypes == 0] = 0.5 * c2.params.CSolid
    U2, ph2, sc2, v02 = O.init_state(c2)
    p2 = c2.params
    assert sc2 == c2.n and abs(v02 - c2.n * (3.0 - p2.CSat / p2.CSolid)) < 1e-12 * abs(v02)
    assert np.all(U2[0][c2.bctypes == 0] == 0.5)


def _fd_case(ng, S, seed, thr=None):
    case = nn.synth.regular(ng, S=S, seed=seed)
    case.params.sub_steps = 2  # several voln1rel (BV) updates inside the run
    if thr is not None:
        case.params.front_threshold = thr
    p = case.params
    st, rp, col, _ = O.build_neighbors(case.x, case.y, case.z, p.delta * (1 + 1e-5), p.dx * 10.,
                                       p.dy * 10., p.dz * 10.)
    assert st == 0
    return case, rp, col


@pytest.mark.parametrize("nsteps,thr", [(6, None), (9, 1.0)])
def test_adjoint_matches_central_differences(nsteps, thr):
    """or_adjoint against central differences of or_step, for every parameter and one random
    direction of the starting state; thr = 1.0 flips about half the solid nodes on the way."""
    import copy
    case, rp, col = _fd_case(8, 3, 3, thr)
    S = case.params.S
    rng = np.random.default_rng(0)
    lamN = rng.standard_normal(case.icvalues.shape)

    def loss(c, state=None):
        return float((lamN * O.run(c, rp, col, nsteps, rtol=1e-14, state=state)["U"]).sum())

    grad, lam0, done, st = O.adjoint(case, rp, col, nsteps, lamN, rtol=1e-14)
    assert st == 0 and done == nsteps
    names = [("D_liquid", k) for k in range(S)] + [("D_solid", k) for k in range(S)] + [("r_diss", None)]
    fd = np.zeros_like(grad)
    for q, (name, k) in enumerate(names):
        base = getattr(case.params, name) if k is None else getattr(case.params, name)[k]
        h = 1e-5 * abs(base)
        vals = []
        for sgn in (1, -1):
            c = copy.deepcopy(case)
            if k is None:
                setattr(c.params, name, base + sgn * h)
            else:
                getattr(c.params, name)[k] = base + sgn * h
            vals.append(loss(c))
        fd[q] = (vals[0] - vals[1]) / (2 * h)
    scale = np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    assert np.all(np.abs(grad - fd) <= 2e-6 * scale), (grad, fd)
    U0, ph0, sc0, v0 = O.init_state(case)
    d = rng.standard_normal(U0.shape)
    h = 1e-6
    Ls = [loss(case, {"U": U0 + s * h * d, "phase": ph0, "step": 1, "voln1rel": 0.0,
                      "vol0rel": v0, "solid": sc0}) for s in (1, -1)]
    fdd = (Ls[0] - Ls[1]) / (2 * h)
    assert abs(float((lam0 * d).sum()) - fdd) <= 1e-7 * abs(fdd)


def test_adjoint_rejects_reference_chemistry():
    case, rp, col = _fd_case(6, 10, 1)
    case.params.chemistry = nn.CHEM_REFERENCE
    _, _, done, st = O.adjoint(case, rp, col, 1, np.zeros_like(case.icvalues))
    assert st == nn._abi.ERR_UNSUPPORTED and done == 0
