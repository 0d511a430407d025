"""Write golden vectors from the reference's own code (oracle/_ref, compiled from the
reference's grid.h/utils.h) for small synthetic cases. Run in the build container, where the
reference sources exist; the .npz fixtures are committed so the checks also run elsewhere.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import ell_to_sorted_csr, orc, ref_chem_case, reference_params  # noqa: E402


def main():
    R = orc.RefHarness()
    for name, ng, jit, seed in (("reg7", 7, 0.0, 3), ("jit7", 7, 0.3, 4)):
        case, phase, U1, Un = ref_chem_case(ng=ng, seed=seed, jitter=jit)
        p = case.params
        r = p.delta * (1.00 + 1e-5)
        st, ell, ln, ml = R.build_clouds_hash(case.x, case.y, case.z, r, p.dx * 10, p.dy * 10,
                                              p.dz * 10)
        assert st == 0
        rp, col = ell_to_sorted_csr(ell, ln)
        voln1rel = 0.37
        F = R.residuals(case, ell, ln, phase, U1, Un, voln1rel=voln1rel)
        kr, kc, kv = R.newton_matrix(case, ell, ln, phase, U1, Un, 11, voln1rel=voln1rel)
        beta_xi = np.linspace(0.0, 4.0 * p.dx, 41)
        beta = np.array([R.lib.ref_beta(v, p.delta, p.dx) for v in beta_xi])
        # every block K_1..K_11 of NewtonMatrices applied to two seeded random vectors, and its
        # diagonal (the full matrices stay out of the fixture)
        rng = np.random.default_rng(100 + seed)
        jac_v = rng.uniform(-1.0, 1.0, size=(2, 11, case.n))
        jac_kv = np.zeros((11, 2, case.n))
        jac_diag = np.zeros((11, case.n))
        for f in range(11):
            r_, c_, v_ = R.newton_matrix(case, ell, ln, phase, U1, Un, f + 1, voln1rel=voln1rel)
            for m in range(2):
                np.add.at(jac_kv[f, m], r_, v_ * jac_v[m, f, c_])
            d = r_ == c_
            np.add.at(jac_diag[f], r_[d], v_[d])
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), x=case.x, y=case.y, z=case.z,
                            bctypes=case.bctypes, bcvalues=case.bcvalues, phase=phase, U1=U1,
                            Un=Un, voln1rel=voln1rel, row_ptr=rp, col=col, max_len=ml,
                            lengths=ln, F=F, K11_rows=kr, K11_cols=kc, K11_vals=kv,
                            beta_xi=beta_xi, beta=beta, jac_v=jac_v, jac_kv=jac_kv,
                            jac_diag=jac_diag)
        print(name, case.n, int(rp[-1]), ml)


def implant_case(seed=11, ng=12):
    """Lattice nodes with the production spacing and an implant point cloud: points inside a
    sphere plus points at distance r (1 + e) of nodes for e within a few ulps of 0, some exactly
    on an axis, so the boundary of the predicate is exercised."""
    p = reference_params()
    rng = np.random.default_rng(seed)
    dx = p.dx
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = (a.ravel().copy() for a in (xx, yy, zz))
    L = ng * dx
    r = 3. * 0.4 * dx * (1.00 + 1e-5)  # main.cpp:317
    u = rng.standard_normal((3, 3000))
    u /= np.linalg.norm(u, axis=0)
    rad = 0.3 * L * rng.uniform(0, 1, 3000) ** (1 / 3)
    pts = [L / 2 + u * rad]
    pick = rng.choice(x.size, 400, replace=False)
    v = rng.standard_normal((3, 400))
    v /= np.linalg.norm(v, axis=0)
    e = rng.choice([-4e-16, -2e-16, 0.0, 2e-16, 4e-16, -1e-12, 1e-12], 400)
    node = np.stack([x[pick], y[pick], z[pick]])
    pts.append(node + v * (r * (1 + e)))
    ax = np.zeros((3, 60))
    for j in range(60):
        ax[j % 3, j] = r if j % 2 == 0 else -r
    pts.append(np.stack([x[pick[:60]], y[pick[:60]], z[pick[:60]]]) + ax)
    P = np.concatenate(pts, axis=1)
    return p, x, y, z, P[0].copy(), P[1].copy(), P[2].copy(), r


def main_implant():
    R = orc.RefHarness()
    p, x, y, z, bx, by, bz, r = implant_case()
    st, counts = R.build_clouds_implant(x, y, z, bx, by, bz, r)
    assert st == 0
    np.savez_compressed(os.path.join(HERE, "ref_implant.npz"), x=x, y=y, z=z, bx=bx, by=by,
                        bz=bz, r=r, counts=counts)
    print("implant", x.size, bx.size, int((counts > 0).sum()), int(counts.max()))


if __name__ == "__main__":
    main()
    main_implant()
