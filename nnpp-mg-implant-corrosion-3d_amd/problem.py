"""Host-side mirror of the reference's step-path interface over the C ABI.

One `Problem` owns one device context (nnpp_ctx). Method names follow the reference:
`build_clouds_hash` (grid.h:43-204), `residuals` (utils::Residuals, utils.h:27-842),
`step` (the time loop of main.cpp:547-1025 with an explicit update), `front_update`
(main.cpp:915-922). Field arrays are numpy (F, n) float64, field f = species f+1 of the
reference, field S = phi.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import FP64, MIXED, SLOT_N, SLOT_N1, NnppError, dptr, iptr, lptr


class Problem:
    def __init__(self, params, x, y, z, volumes=None, device=0):
        self.lib = _abi.load()
        self.params = params.copy()
        self.S = params.S
        self.F = self.S + 1
        ctx = C.c_void_p()
        st = self.lib.nnpp_create(C.byref(ctx), device)
        if st != _abi.OK:
            raise NnppError(st, "nnpp_create failed (no HIP device?)")
        self.ctx = ctx
        self._check(self.lib.nnpp_set_params(self.ctx, C.byref(self.params)))
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        self.n = int(self.x.shape[0])
        self.vol = None if volumes is None else np.ascontiguousarray(volumes, dtype=np.float64)
        self._check(self.lib.nnpp_set_nodes(self.ctx, self.n, dptr(self.x), dptr(self.y),
                                            dptr(self.z), dptr(self.vol) if self.vol is not None else None))
        self.nnz = None
        # differentiation tape bookkeeping: a generation number per tape_begin and the steps
        # recorded since, so a graph built on one tape can tell that it was replaced or extended
        self._tape_gen = 0
        self._tape_active = False
        self._tape_steps = 0

    # -- plumbing ------------------------------------------------------------------------
    def _check(self, st):
        if st != _abi.OK:
            msg = self.lib.nnpp_last_error(self.ctx)
            raise NnppError(st, msg.decode() if msg else "")
        return st

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.nnpp_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- neighbourhoods ------------------------------------------------------------------
    def build_clouds_hash(self, max_node_no=1024, r=None, dx=None, dy=None, dz=None):
        """grid.h:43-204 as called at main.cpp:304-305: r = delta*(1+1e-5), cells 10*dx.

        Returns (nnz, max_cloud_length); max_cloud_length counts self as the reference does.
        Raises NnppError(ERR_OVERFLOW) where the reference throws std::runtime_error."""
        p = self.params
        r = p.delta * (1.00 + 1e-5) if r is None else r
        dx = p.dx * 10. if dx is None else dx
        dy = p.dy * 10. if dy is None else dy
        dz = p.dz * 10. if dz is None else dz
        nnz = C.c_int64()
        ml = C.c_int32()
        self._check(self.lib.nnpp_build_neighbors(self.ctx, r, dx, dy, dz, max_node_no,
                                                  C.byref(nnz), C.byref(ml)))
        self.nnz = int(nnz.value)
        return self.nnz, int(ml.value)

    def implant_classify(self, bx, by, bz, r=None, max_count=1024, want_ic=False):
        """kd_tree_interface.h:156-252 as main.cpp:316-320 calls it: implant points within
        r = 3*0.4*dx*(1+1e-5) of every node; returns (counts, initial values of main.cpp:422-457
        or None). Raises NnppError(ERR_OVERFLOW) where the reference throws."""
        p = self.params
        r = 3. * 0.4 * p.dx * (1.00 + 1e-5) if r is None else r
        b = [np.ascontiguousarray(a, dtype=np.float64) for a in (bx, by, bz)]
        counts = np.empty(self.n, dtype=np.int32)
        ic = np.empty((self.F, self.n)) if want_ic else None
        self._check(self.lib.nnpp_implant_classify(self.ctx, b[0].size, dptr(b[0]), dptr(b[1]),
                                                   dptr(b[2]), r, max_count, iptr(counts),
                                                   dptr(ic) if want_ic else None))
        return counts, ic

    def voronoi_volumes(self, h=None, margin=None, radius=None):
        """Voronoi-approximate node volumes (probe lattice of spacing h, default dx/4; margin
        default dx/2; radius default 2 delta)."""
        p = self.params
        d = min(p.dx, p.dy, p.dz)
        h = d / 4 if h is None else h
        margin = d / 2 if margin is None else margin
        radius = 2 * p.delta if radius is None else radius
        out = np.empty(self.n)
        self._check(self.lib.nnpp_voronoi_volumes(self.ctx, h, margin, radius, dptr(out)))
        return out

    def neighbors(self):
        rp = np.empty(self.n + 1, dtype=np.int64)
        col = np.empty(max(self.nnz or 0, 1), dtype=np.int32)
        self._check(self.lib.nnpp_get_neighbors(self.ctx, lptr(rp), iptr(col)))
        return rp, col[: self.nnz]

    def set_neighbors(self, row_ptr, col):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        self._check(self.lib.nnpp_set_neighbors(self.ctx, lptr(rp), iptr(cl) if cl.size else None))
        self.nnz = int(rp[-1])

    # -- inputs / state ------------------------------------------------------------------
    def set_bc(self, bctypes, bcvalues):
        bt = np.ascontiguousarray(bctypes, dtype=np.int32)
        bv = np.ascontiguousarray(bcvalues, dtype=np.float64).reshape(self.F, self.n)
        self._check(self.lib.nnpp_set_bc(self.ctx, iptr(bt), dptr(bv)))

    def init_state(self, icvalues):
        ic = np.ascontiguousarray(icvalues, dtype=np.float64).reshape(self.F, self.n)
        self._check(self.lib.nnpp_init_state(self.ctx, dptr(ic)))

    def set_state(self, U, slot=SLOT_N):
        a = np.ascontiguousarray(U, dtype=np.float64).reshape(self.F, self.n)
        self._check(self.lib.nnpp_set_state(self.ctx, slot, dptr(a)))

    def get_state(self, slot=SLOT_N):
        a = np.empty((self.F, self.n), dtype=np.float64)
        self._check(self.lib.nnpp_get_state(self.ctx, slot, dptr(a)))
        return a

    def set_phase(self, phase):
        a = np.ascontiguousarray(phase, dtype=np.int32)
        self._check(self.lib.nnpp_set_phase(self.ctx, iptr(a)))

    def get_phase(self):
        a = np.empty(self.n, dtype=np.int32)
        self._check(self.lib.nnpp_get_phase(self.ctx, iptr(a)))
        return a

    def set_time(self, step, voln1rel, vol0rel):
        self._check(self.lib.nnpp_set_time(self.ctx, step, voln1rel, vol0rel))

    def get_time(self):
        """(step counter, voln1rel, vol0rel)."""
        st, vr, v0 = C.c_int64(), C.c_double(), C.c_double()
        self._check(self.lib.nnpp_get_time(self.ctx, C.byref(st), C.byref(vr), C.byref(v0)))
        return int(st.value), float(vr.value), float(v0.value)

    def set_params(self, params):
        self.params = params.copy()
        self._check(self.lib.nnpp_set_params(self.ctx, C.byref(self.params)))

    def snapshot(self):
        """Host copy of everything a step depends on besides parameters and geometry."""
        return {"U": self.get_state(), "phase": self.get_phase(), "time": self.get_time()}

    def restore(self, snap):
        self.set_state(snap["U"])
        self.set_phase(snap["phase"])
        self.set_time(*snap["time"])

    def potential_operator_apply(self, x, precision=FP64):
        """One application of the solver's (scaled, free-row) potential operator."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        self._check(self.lib.nnpp_potential_operator_apply(self.ctx, precision, dptr(x), dptr(y)))
        return y

    def set_solver(self, rtol=1e-10, maxit=20000, inner_rtol=1e-3):
        self._check(self.lib.nnpp_set_solver(self.ctx, rtol, maxit, inner_rtol))

    def set_format(self, flags=0):
        """Operator storage format (FORMAT_* flags); all formats give bit-identical results."""
        self._check(self.lib.nnpp_set_format(self.ctx, int(flags)))

    # -- operators -----------------------------------------------------------------------
    def residuals(self, voln1rel=0.0, apply_bc=False, precision=FP64):
        """utils::Residuals on (slot N1 = U^{n+1}, slot N = U^n); (F, n) array."""
        out = np.empty((self.F, self.n), dtype=np.float64)
        self._check(self.lib.nnpp_residuals(self.ctx, voln1rel, int(apply_bc), precision, dptr(out)))
        return out

    def poisson_apply(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty(self.n, dtype=np.float64)
        self._check(self.lib.nnpp_poisson_apply(self.ctx, dptr(a), dptr(out)))
        return out

    def poisson_solve(self, precision=FP64):
        it = C.c_int32()
        rr = C.c_double()
        st = self.lib.nnpp_poisson_solve(self.ctx, precision, C.byref(it), C.byref(rr))
        if st not in (_abi.OK, _abi.ERR_NOCONV):
            self._check(st)
        return int(it.value), float(rr.value), st

    def step(self, nsteps=1, precision=FP64, check_finite=True):
        s = _abi.StepStats()
        st = self.lib.nnpp_step(self.ctx, nsteps, precision, int(check_finite), C.byref(s))
        d = s.as_dict()
        d["status"] = st
        if self._tape_active:
            self._tape_steps += int(d["steps"])
        if st not in (_abi.OK, _abi.ERR_NOCONV):
            self._check(st)
        return d

    def newton_opts(self, **kw):
        """The reference's Newton settings (main.cpp:35-36), with keyword overrides."""
        o = _abi.NewtonOpts()
        self._check(self.lib.nnpp_newton_defaults(C.byref(o)))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def step_implicit(self, nsteps=1, precision=FP64, check_finite=True, **opts):
        """The reference's implicit step (main.cpp:547-1025): modified Newton with the frozen
        Jacobian of utils.h:844-1849; keyword options as nnpp_newton_opts."""
        o = self.newton_opts(**opts)
        s = _abi.NewtonStats()
        st = self.lib.nnpp_step_implicit(self.ctx, nsteps, precision, int(check_finite),
                                          C.byref(o), C.byref(s))
        d = s.as_dict()
        d["status"] = st
        if st not in (_abi.OK, _abi.ERR_NOCONV):
            self._check(st)
        return d

    def jacobian_apply(self, v, voln1rel=0.0, want_diag=False):
        """Frozen Jacobian at slot N (utils.h:844-1849) applied block by block to v (F, n);
        returns (K v, diagonals or None)."""
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self.F, self.n)
        out = np.empty((self.F, self.n))
        dg = np.empty((self.F, self.n)) if want_diag else None
        self._check(self.lib.nnpp_jacobian_apply(self.ctx, voln1rel, dptr(a), dptr(out),
                                                 dptr(dg) if want_diag else None))
        return out, dg

    def front_update(self, threshold=None):
        thr = self.params.front_threshold if threshold is None else threshold
        fl = C.c_int32()
        self._check(self.lib.nnpp_front_update(self.ctx, thr, C.byref(fl)))
        return int(fl.value)

    def state_device_ptr(self, field, slot=SLOT_N):
        p = C.c_void_p()
        self._check(self.lib.nnpp_state_device_ptr(self.ctx, slot, field, C.byref(p)))
        return p.value

    def stream(self):
        p = C.c_void_p()
        self._check(self.lib.nnpp_stream(self.ctx, C.byref(p)))
        return p.value

    # -- differentiation -------------------------------------------------------------------
    def tape_begin(self, checkpoint_every=1):
        """Record the following FP64 steps for tape_adjoint (state checkpoint every k steps)."""
        self._check(self.lib.nnpp_tape_begin(self.ctx, int(checkpoint_every)))
        self._tape_gen += 1
        self._tape_active = True
        self._tape_steps = 0

    def tape_info(self):
        """(generation of the current tape, recording, steps recorded on it)."""
        return self._tape_gen, self._tape_active, self._tape_steps

    def tape_adjoint(self, lam, want_lam0=False):
        """Reverse the recorded steps for dL/dU = lam (F, n) at the current state. Returns
        (grad, lam0, stats): grad = [dL/dD_liquid (S), dL/dD_solid (S), dL/dr_diss]."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        if lam.shape != (self.F, self.n):
            raise ValueError(f"lam must have shape {(self.F, self.n)}")
        grad = np.zeros(2 * self.S + 1)
        lam0 = np.empty((self.F, self.n)) if want_lam0 else None
        st = _abi.AdjointStats()
        self._check(self.lib.nnpp_tape_adjoint(self.ctx, dptr(lam), dptr(grad),
                                               dptr(lam0) if want_lam0 else None, C.byref(st)))
        return grad, lam0, st.as_dict()

    def tape_end(self):
        self._check(self.lib.nnpp_tape_end(self.ctx))
        self._tape_active = False

    def time_kernel(self, kernel, precision=FP64, reps=20):
        """Mean duration (ms) of one launch of a hot kernel, HIP events on the context stream."""
        ms = C.c_double()
        self._check(self.lib.nnpp_time_kernel(self.ctx, kernel, precision, reps, C.byref(ms)))
        return float(ms.value)
