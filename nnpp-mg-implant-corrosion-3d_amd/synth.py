"""Synthetic NNPP cases in the reference's input layout (numpy, seeded).

A case holds what the reference reads from <job>.data1/.data2/.datac*/.nodes/.bctypes/
.bcvalues*/.icvalues* (main.cpp:210-289): parameters, coordinates, optional per-node
volumes, BC types, raw BC values (species in units of CSolid, -1 = free) and raw IC values.

Regular case (`regular`): nodes at (i+1/2) dx in the unit cube, dx = 1/ng, delta = 3 dx,
S species (default 11) with D_liquid ~ U[0.1, 1], D_solid = 0.1 D_liquid, valences from
{-2,-1,1,2}, concentrations ~ U[0.9, 1.1]; a centred Mg sphere of radius 0.3 is solid (Mg IC
= U[0.9,1.1] inside, 0 outside, so the reference's rule "solid iff U_Mg > EPS" selects it);
outer faces carry Dirichlet c = 1 (bcvalues = CSolid) and phi = 0; dt = 0.2 dx^2 / max D;
generic chemistry with an interface Mg sink of rate r_diss = 0.05.

Irregular case (`screw`): the ng^3 lattice jittered by U(+-0.25 dx), restricted to a fluid
cylinder of radius 0.48 around the z axis through (0.5, 0.5); a screw-like solid (core radius
0.2, triangular helical thread of pitch 0.1 and depth 0.03, z in [0.1, 0.9]) inside it;
per-node volumes from half the spacing to the lattice neighbours on each axis; Dirichlet
c = 1, phi = 0 on the fluid's outer layer; front threshold 0.5.
"""
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._abi import CHEM_GENERIC, Params, load


@dataclass
class Case:
    params: Params
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vol: Optional[np.ndarray]
    bctypes: np.ndarray  # (n,) int32, 0 = Dirichlet node
    bcvalues: np.ndarray  # (F, n) raw values, -1 = not prescribed
    icvalues: np.ndarray  # (F, n) raw initial values
    name: str = ""

    @property
    def n(self):
        return int(self.x.shape[0])

    @property
    def S(self):
        return self.params.S


def synthetic_params(ng, S=11, seed=0, r_diss=0.05, vol_red=1.0, sigma=2.0, dt_factor=0.2,
                     d_solid_ratio=0.1, csat=0.5):
    """Parameters of the synthetic generic-chemistry cases (units with CSolid = F = R = T = 1)."""
    rng = np.random.default_rng(seed)
    p = Params()
    dx = 1.0 / ng
    delta = 3.0 * dx
    Dl = rng.uniform(0.1, 1.0, size=S)
    z = rng.choice(np.array([-2.0, -1.0, 1.0, 2.0]), size=S)
    Kp = 3.0 / (math.pi * delta * delta)  # K = 3/(pi delta^2) for s = 0, as data/NNPP.datac1
    p.n_nodes = ng ** 3
    p.n_steps = 100
    p.t0 = 0.0
    p.dt = dt_factor * dx * dx / float(Dl.max())
    p.t1 = p.dt * p.n_steps
    p.dx = p.dy = p.dz = dx
    p.delta = delta
    p.CSolid = 1.0
    p.CSat = csat
    p.T = 1.0
    p.R = 1.0
    p.F = 1.0
    p.vol_red = vol_red
    for k in range(S):
        p.z[k] = z[k]
        p.D_liquid[k] = Dl[k]
        p.D_solid[k] = d_solid_ratio * Dl[k]
        p.s1[k] = 0.0
        p.s2[k] = 0.0
        p.K1[k] = Kp
        p.K2[k] = Kp
    p.sigma0 = sigma
    p.sigma1 = sigma
    p.s_phi = 0.0
    p.K_phi = Kp
    st = load().nnpp_params_defaults(p, S, CHEM_GENERIC)
    if st != 0:
        raise ValueError(f"nnpp_params_defaults failed ({st})")
    p.r_diss = r_diss
    return p


def regular(ng, S=11, seed=0, sphere_radius=0.3, **kw):
    """Regular ng^3 lattice case (the 24^3 / 32^3 / 64^3 / 128^3 workloads)."""
    p = synthetic_params(ng, S=S, seed=seed, **kw)
    dx = 1.0 / ng
    g = (np.arange(ng, dtype=np.float64) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")  # node id = ix + ng*(iy + ng*iz)
    x, y, z = xx.ravel(), yy.ravel(), zz.ravel()
    n = x.size
    F = S + 1
    rng = np.random.default_rng(seed + 1000)
    ic = np.empty((F, n), dtype=np.float64)
    ic[:S] = rng.uniform(0.9, 1.1, size=(S, n))
    ic[S] = 0.0
    solid = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 <= sphere_radius ** 2
    ic[0, ~solid] = 0.0
    idx = np.arange(ng)
    I = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), 0).reshape(3, -1)
    face = ((I == 0) | (I == ng - 1)).any(axis=0)
    bctypes = np.where(face, 0, 1).astype(np.int32)
    bcv = np.full((F, n), -1.0)
    bcv[:S, face] = 1.0 * p.CSolid
    bcv[S, face] = 0.0
    return Case(p, x, y, z, None, bctypes, bcv, ic, name=f"regular{ng}")


def screw(ng, S=11, seed=0, jitter=0.25, r_fluid=0.48, r_core=0.2, pitch=0.1, depth=0.03,
          **kw):
    """Irregular jittered lattice in a fluid cylinder around a screw-like solid."""
    p = synthetic_params(ng, S=S, seed=seed, **kw)
    dx = 1.0 / ng
    rng = np.random.default_rng(seed + 2000)
    g = (np.arange(ng, dtype=np.float64) + 0.5) * dx
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([xx, yy, zz], 0) + rng.uniform(-jitter * dx, jitter * dx, size=(3, ng, ng, ng))
    # per-node volume: product over axes of half the distance between the lattice neighbours
    vol = np.ones((ng, ng, ng))
    for a, ax in enumerate((2, 1, 0)):  # x varies along array axis 2, y along 1, z along 0
        c = P[a]
        fwd = np.diff(c, axis=ax)
        h = np.zeros_like(c)
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(0, -1)
        sl_hi[ax] = slice(1, None)
        h[tuple(sl_lo)] += 0.5 * fwd
        h[tuple(sl_hi)] += 0.5 * fwd
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[ax] = slice(0, 1)
        last[ax] = slice(-1, None)
        h[tuple(first)] *= 2.0
        h[tuple(last)] *= 2.0
        vol *= h
    x, y, z = P[0].ravel(), P[1].ravel(), P[2].ravel()
    vol = vol.ravel()
    rho = np.hypot(x - 0.5, y - 0.5)
    keep = rho <= r_fluid
    x, y, z, vol, rho = x[keep], y[keep], z[keep], vol[keep], rho[keep]
    theta = np.arctan2(y - 0.5, x - 0.5)
    u = np.mod(z / pitch - theta / (2 * math.pi), 1.0)
    r_screw = r_core + depth * (1.0 - np.abs(2.0 * u - 1.0))
    solid = (rho <= r_screw) & (z >= 0.1) & (z <= 0.9)
    n = x.size
    F = S + 1
    ic = np.empty((F, n))
    ic[:S] = rng.uniform(0.9, 1.1, size=(S, n))
    ic[S] = 0.0
    ic[0, ~solid] = 0.0
    outer = (rho > r_fluid - dx) | (z < dx) | (z > 1.0 - dx)
    bc = outer & ~solid
    bctypes = np.where(bc, 0, 1).astype(np.int32)
    bcv = np.full((F, n), -1.0)
    bcv[:S, bc] = 1.0 * p.CSolid
    bcv[S, bc] = 0.0
    p.n_nodes = n
    return Case(p, x, y, z, vol, bctypes, bcv, ic, name=f"screw{ng}")


def crop(case, lo, hi):
    """Sub-box [lo, hi)^3 of a case (node coordinates), e.g. the 32^3 crop of the 128^3 case."""
    m = ((case.x >= lo) & (case.x < hi) & (case.y >= lo) & (case.y < hi) & (case.z >= lo)
         & (case.z < hi))
    p = case.params.copy()
    p.n_nodes = int(m.sum())
    return Case(p, case.x[m].copy(), case.y[m].copy(), case.z[m].copy(),
                None if case.vol is None else case.vol[m].copy(), case.bctypes[m].copy(),
                case.bcvalues[:, m].copy(), case.icvalues[:, m].copy(), name=case.name + "-crop")
