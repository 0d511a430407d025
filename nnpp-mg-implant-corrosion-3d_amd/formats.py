"""Writers/readers of the reference's job files (main.cpp:210-289, utils.h:1851-1887,2334-2484).

<job>.data1   nodesNo stepsNo bcNodesNo bunnyNodesNo
<job>.data2   51 values (utils::parameter order)
<job>.datac1 .. <job>.datac<S>   z DSolid DLiquid s1 s2 Kprop1 Kprop2
<job>.datac<S+1>                 sigma0 sigma1 s111 K11prop1
<job>.nodes   n lines "x y z"
<job>.bctypes n ints;  <job>.bcvalues<f>, <job>.icvalues<f>  n doubles, f = 1..S+1
Reading goes through the C ABI readers (nnpp_read_*), the same code a C++ caller uses.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import Params, dptr, iptr
from .synth import Case


def _fmt(v):
    return repr(float(v))


def write_job(job, case):
    """Write a Case as reference job files with the given path prefix."""
    p = case.params
    S = p.S
    n = case.n
    os.makedirs(os.path.dirname(os.path.abspath(job)), exist_ok=True)
    with open(job + ".data1", "w") as f:
        nbc = int((case.bctypes == 0).sum())
        f.write(f"{n}\n{int(p.n_steps)}\n{nbc}\n0\n")
    with open(job + ".data2", "w") as f:
        f.write("\n".join(_fmt(v) for v in p.data2()) + "\n")
    for k in range(S):
        with open(job + f".datac{k + 1}", "w") as f:
            f.write("\n".join(_fmt(v) for v in p.datac(k)) + "\n")
    with open(job + f".datac{S + 1}", "w") as f:
        f.write("\n".join(_fmt(v) for v in p.datac_phi()) + "\n")
    np.savetxt(job + ".nodes", np.stack([case.x, case.y, case.z], 1), fmt="%.17g")
    np.savetxt(job + ".bctypes", case.bctypes, fmt="%d")
    for f_ in range(S + 1):
        np.savetxt(job + f".bcvalues{f_ + 1}", case.bcvalues[f_], fmt="%.17g")
        np.savetxt(job + f".icvalues{f_ + 1}", case.icvalues[f_], fmt="%.17g")
    if case.vol is not None:
        np.savetxt(job + ".volumes", case.vol, fmt="%.17g")
    with open(job + ".nnpp", "w") as f:  # extension fields the reference files cannot hold
        f.write(f"r_diss = {_fmt(p.r_diss)}\nfront_threshold = {_fmt(p.front_threshold)}\n"
                f"clamp_floor = {_fmt(p.clamp_floor)}\nphase_eps = {_fmt(p.phase_eps)}\n"
                f"sub_steps = {int(p.sub_steps)}\n"
                f"solid_bonds = {' '.join(str(int(v)) for v in p.solid_bonds[:S])}\n"
                f"clamp = {' '.join(str(int(v)) for v in p.clamp[:S])}\n")


def read_job(job, n_species, chemistry=_abi.CHEM_GENERIC):
    """Read reference job files into a Case via the library's C readers."""
    lib = _abi.load()
    p = Params()
    st = lib.nnpp_read_params(job.encode(), n_species, chemistry, C.byref(p))
    if st != _abi.OK:
        raise IOError(f"nnpp_read_params({job}) failed with {st}")
    n = int(p.n_nodes)
    x, y, z = (np.empty(n) for _ in range(3))
    if lib.nnpp_read_nodes((job + ".nodes").encode(), n, dptr(x), dptr(y), dptr(z)) != 0:
        raise IOError(job + ".nodes")
    bt = np.empty(n, dtype=np.int32)
    if lib.nnpp_read_ivector((job + ".bctypes").encode(), n, iptr(bt)) != 0:
        raise IOError(job + ".bctypes")
    F = n_species + 1
    bcv = np.empty((F, n))
    ic = np.empty((F, n))
    for f in range(F):
        if lib.nnpp_read_vector((job + f".bcvalues{f + 1}").encode(), n, dptr(bcv[f])) != 0:
            raise IOError(job + f".bcvalues{f + 1}")
        if lib.nnpp_read_vector((job + f".icvalues{f + 1}").encode(), n, dptr(ic[f])) != 0:
            raise IOError(job + f".icvalues{f + 1}")
    vol = None
    if os.path.exists(job + ".volumes"):
        vol = np.empty(n)
        lib.nnpp_read_vector((job + ".volumes").encode(), n, dptr(vol))
    return Case(p, x, y, z, vol, bt, bcv, ic, name=os.path.basename(job))
