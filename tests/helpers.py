"""Shared test helpers: package/oracle loading, reference-chemistry cases, comparisons."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

nn = ge.load_package()
orc = ge.load_oracle()
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_DATA = os.path.join(GOLDEN, "reference_data", "NNPP")


def reference_params():
    """The reference's production parameters (data/NNPP.data1/.data2/.datac1..11), S = 10."""
    import ctypes as C
    p = nn.Params()
    st = nn.lib.nnpp_read_params(REF_DATA.encode(), 10, nn.CHEM_REFERENCE, C.byref(p))
    assert st == 0, st
    return p


def ref_chem_case(ng=8, seed=0, jitter=0.0, solid_radius=0.3):
    """Reference-chemistry case on an ng^3 lattice with the production spacing dx = 2e-5.

    States are log-uniform in [1e-12, 1] so that every max(.,0) term of the reactions is hit
    in both branches; phi ~ U(-0.01, 0.01); a centred sphere is solid; the outer layer carries
    Dirichlet values for some fields and -1 for others."""
    p = reference_params()
    rng = np.random.default_rng(seed)
    dx = p.dx
    g = (np.arange(ng) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([xx, yy, zz]).reshape(3, -1)
    if jitter:
        P = P + rng.uniform(-jitter * dx, jitter * dx, size=P.shape)
    x, y, z = P
    n = x.size
    p.n_nodes = n
    L = ng * dx
    solid = ((x - L / 2) ** 2 + (y - L / 2) ** 2 + (z - L / 2) ** 2) <= (solid_radius * L) ** 2
    U1 = 10.0 ** rng.uniform(-12, 0, size=(11, n))
    Un = 10.0 ** rng.uniform(-12, 0, size=(11, n))
    U1[10] = rng.uniform(-0.01, 0.01, size=n)
    Un[10] = rng.uniform(-0.01, 0.01, size=n)
    phase = solid.astype(np.int32)
    idx = np.arange(ng)
    I = np.stack(np.meshgrid(idx, idx, idx, indexing="ij")).reshape(3, -1)
    face = ((I == 0) | (I == ng - 1)).any(0)
    bt = np.where(face, 0, 1).astype(np.int32)
    bcv = np.full((11, n), -1.0)
    for f in range(11):
        if f % 3 != 2:
            bcv[f, face] = rng.uniform(0, 1) * (p.CSolid if f < 10 else 1.0)
    case = nn.synth.Case(p, x.copy(), y.copy(), z.copy(), None, bt, bcv, Un.copy(),
                         name=f"refchem{ng}")
    return case, phase, U1, Un


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / max(s, 1e-300))


def ell_to_sorted_csr(ell, lengths):
    """Reference ELL (row-major n x max_len, self at 0) -> id-sorted CSR without self."""
    n = ell.shape[0]
    rows = [np.sort(ell[i, 1:lengths[i]]) for i in range(n)]
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([r.size for r in rows])
    col = np.concatenate(rows).astype(np.int32) if n else np.zeros(0, np.int32)
    return rp, col
