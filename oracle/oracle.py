"""ctypes binding of the CPU oracle (liboracle.so) and of the reference harness (_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnnpp_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_i32, _i64, _d = C.c_int32, C.c_int64, C.c_double


def _d_(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i_(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _l_(a):
    return None if a is None else a.ctypes.data_as(_lp)


class Oracle:
    """The C restatement of the reference path (oracle/nnpp_oracle.c)."""

    def __init__(self, params_cls):
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        self.lib = C.CDLL(ORACLE_SO)
        P = C.POINTER(params_cls)
        L = self.lib
        L.or_beta.restype = _d
        L.or_beta.argtypes = [_d, _d, _d]
        L.or_sphere_cube_intersects.argtypes = [_d] * 10
        L.or_build_neighbors.argtypes = [_i32, _dp, _dp, _dp, _d, _d, _d, _d, _i32, _lp, _ip, _ip]
        L.or_residuals.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _dp, _dp, _d, _dp]
        L.or_bc_override.argtypes = [_i32, P, _ip, _dp, _dp, _dp]
        L.or_poisson_apply.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _dp, _dp]
        L.or_poisson_solve.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _ip, _dp, _dp,
                                       _dp, _d, _i32, _ip, _dp]
        L.or_init_state.argtypes = [_i32, P, _ip, _dp, _dp, _dp, _ip, _lp, _dp]
        L.or_step.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _dp, _dp, _ip, _lp, _dp,
                              _d, _lp, _d, _i32, _ip, _dp, _ip]
        L.or_poisson_solve_rhs.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _ip, _dp,
                                           _dp, _dp, _d, _i32, _ip, _dp]
        L.or_adjoint.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _dp, _dp, _ip,
                                 C.c_int64, _d, _d, _i32, _d, _i32, _dp, _dp, _dp, _ip]
        L.or_implant_classify.argtypes = [_i32, _dp, _dp, _dp, C.c_int64, _dp, _dp, _dp, _d,
                                          _i32, _ip]
        L.or_implant_icvalues.argtypes = [_i32, P, _ip, _dp]
        L.or_voronoi_volumes.argtypes = [_i32, _dp, _dp, _dp, _d, _d, _d, _dp]
        L.or_jacobian_apply.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _ip, _dp,
                                        _dp, _d, _dp, _dp, _dp]
        L.or_step_implicit.argtypes = [_i32, P, _dp, _dp, _dp, _dp, _lp, _ip, _ip, _dp, _dp, _dp,
                                       _dp, _i32, _ip, _lp, _dp, _d, _lp, _d, _i32, _d, _i32,
                                       _ip, _ip, _ip, _dp, _ip]

    def beta(self, xi, delta, dx):
        return self.lib.or_beta(xi, delta, dx)

    def build_neighbors(self, x, y, z, r, cdx, cdy, cdz, max_len=1024):
        n = x.size
        rp = np.zeros(n + 1, dtype=np.int64)
        ml = C.c_int32()
        st = self.lib.or_build_neighbors(n, _d_(x), _d_(y), _d_(z), r, cdx, cdy, cdz, max_len,
                                         _l_(rp), None, C.byref(ml))
        if st != 0:
            return st, None, None, None
        col = np.zeros(max(int(rp[-1]), 1), dtype=np.int32)
        st = self.lib.or_build_neighbors(n, _d_(x), _d_(y), _d_(z), r, cdx, cdy, cdz, max_len,
                                         _l_(rp), _i_(col), C.byref(ml))
        return st, rp, col[: int(rp[-1])], int(ml.value)

    def residuals(self, case, rp, col, phase, U1, Un, voln1rel=0.0):
        F = np.zeros_like(U1)
        st = self.lib.or_residuals(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                   _d_(case.z), _d_(case.vol), _l_(rp), _i_(col),
                                   _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                   _d_(np.ascontiguousarray(U1)), _d_(np.ascontiguousarray(Un)),
                                   voln1rel, _d_(F))
        assert st == 0, st
        return F

    def bc_override(self, case, U1, F):
        F = F.copy()
        self.lib.or_bc_override(case.n, C.byref(case.params), _i_(case.bctypes),
                                _d_(case.bcvalues), _d_(np.ascontiguousarray(U1)), _d_(F))
        return F

    def poisson_apply(self, case, rp, col, phase, v):
        out = np.zeros(case.n)
        self.lib.or_poisson_apply(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                  _d_(case.z), _d_(case.vol), _l_(rp), _i_(col),
                                  _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                  _d_(np.ascontiguousarray(v, dtype=np.float64)), _d_(out))
        return out

    def poisson_solve(self, case, rp, col, phase, U, phi, rtol=1e-10, maxit=100000):
        phi = np.array(phi, dtype=np.float64, copy=True)
        it = C.c_int32()
        rr = C.c_double()
        st = self.lib.or_poisson_solve(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                       _d_(case.z), _d_(case.vol), _l_(rp), _i_(col),
                                       _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                       _i_(case.bctypes), _d_(case.bcvalues),
                                       _d_(np.ascontiguousarray(U)), _d_(phi), rtol, maxit,
                                       C.byref(it), C.byref(rr))
        return phi, int(it.value), float(rr.value), st

    def init_state(self, case):
        U = np.zeros_like(case.icvalues)
        ph = np.zeros(case.n, dtype=np.int32)
        sc = C.c_int64()
        v0 = C.c_double()
        self.lib.or_init_state(case.n, C.byref(case.params), _i_(case.bctypes),
                               _d_(case.bcvalues), _d_(case.icvalues), _d_(U), _i_(ph),
                               C.byref(sc), C.byref(v0))
        return U, ph, int(sc.value), float(v0.value)

    def adjoint(self, case, rp, col, nsteps, lamN, state=None, rtol=1e-13, maxit=100000):
        """Reverse-mode gradient of nsteps explicit steps from the initial state (or `state`):
        returns (grad [D_liquid(S), D_solid(S), r_diss], lam0 = dL/dU^0, steps, status) for
        L with dL/dU^N = lamN."""
        if state is None:
            U, ph, sc, v0 = self.init_state(case)
            state = {"U": U, "phase": ph, "step": 1, "voln1rel": 0.0, "vol0rel": v0, "solid": sc}
        S = case.params.S
        grad = np.zeros(2 * S + 1)
        lam0 = np.zeros_like(state["U"])
        done = C.c_int32()
        st = self.lib.or_adjoint(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                 _d_(case.z), _d_(case.vol), _l_(rp), _i_(col), _i_(case.bctypes),
                                 _d_(case.bcvalues), _d_(np.ascontiguousarray(state["U"])),
                                 _i_(np.ascontiguousarray(state["phase"], dtype=np.int32)),
                                 state["step"], state["voln1rel"], state["vol0rel"], nsteps, rtol,
                                 maxit, _d_(np.ascontiguousarray(lamN, dtype=np.float64)),
                                 _d_(grad), _d_(lam0), C.byref(done))
        return grad, lam0, int(done.value), st

    def implant_classify(self, x, y, z, bx, by, bz, r, max_count=1024):
        """Implant points within r of every node (brute force); (status, counts)."""
        counts = np.zeros(x.size, dtype=np.int32)
        st = self.lib.or_implant_classify(x.size, _d_(x), _d_(y), _d_(z), bx.size, _d_(bx),
                                          _d_(by), _d_(bz), r, max_count, _i_(counts))
        return st, counts

    def voronoi_volumes(self, x, y, z, h, margin, radius):
        vol = np.zeros(x.size)
        self.lib.or_voronoi_volumes(x.size, _d_(x), _d_(y), _d_(z), h, margin, radius, _d_(vol))
        return vol

    def implant_icvalues(self, params, counts):
        ic = np.zeros((11, counts.size))
        st = self.lib.or_implant_icvalues(counts.size, C.byref(params),
                                          _i_(np.ascontiguousarray(counts, dtype=np.int32)),
                                          _d_(ic))
        assert st == 0, st
        return ic

    def jacobian_apply(self, case, rp, col, phase, Un, v, voln1rel=0.0, want_diag=False):
        """Frozen Jacobian of utils.h:844-1849 at Un applied field by field to v (F, n));
        returns (K v, diagonals or None)."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(v)
        dg = np.zeros_like(v) if want_diag else None
        st = self.lib.or_jacobian_apply(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                        _d_(case.z), _d_(case.vol), _l_(rp), _i_(col),
                                        _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                        _i_(case.bctypes), _d_(case.bcvalues),
                                        _d_(np.ascontiguousarray(Un)), voln1rel, _d_(v),
                                        _d_(out), _d_(dg))
        assert st == 0, st
        return out, dg

    def run_implicit(self, case, rp, col, nsteps, state=None, newton_tol=1e-5, newton_maxit=200,
                     lin_rtol=1e-5, lin_maxit=200):
        """nsteps implicit (reference Newton) steps from the initial state (or `state`, which
        then also carries the rotation buffers "Uprev", "Uit" and the flag "first")."""
        if state is None:
            U, ph, sc, v0 = self.init_state(case)
            st_ = {"U": U, "phase": ph, "step": 1, "voln1rel": 0.0, "vol0rel": v0, "solid": sc}
        else:
            st_ = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in state.items()}
        st_.setdefault("Uprev", np.zeros_like(st_["U"]))
        st_.setdefault("Uit", np.zeros_like(st_["U"]))
        st_.setdefault("first", 1)
        step = C.c_int64(st_["step"])
        vr = C.c_double(st_["voln1rel"])
        sc = C.c_int64(st_["solid"])
        nit, lit, conv, fl = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        rm = C.c_double()
        log = []
        for _ in range(nsteps):
            s = self.lib.or_step_implicit(
                case.n, C.byref(case.params), _d_(case.x), _d_(case.y), _d_(case.z),
                _d_(case.vol), _l_(rp), _i_(col), _i_(case.bctypes), _d_(case.bcvalues),
                _d_(st_["U"]), _d_(st_["Uprev"]), _d_(st_["Uit"]), int(st_["first"]),
                _i_(st_["phase"]), C.byref(step), C.byref(vr), st_["vol0rel"], C.byref(sc),
                newton_tol, newton_maxit, lin_rtol, lin_maxit, C.byref(nit), C.byref(lit),
                C.byref(conv), C.byref(rm), C.byref(fl))
            assert s in (0, 4), s
            st_["first"] = 0
            log.append({"newton": nit.value, "lin": lit.value, "converged": conv.value,
                        "resmax": rm.value, "flips": fl.value})
            if sc.value <= 0:
                break
        st_["step"] = step.value
        st_["voln1rel"] = vr.value
        st_["solid"] = sc.value
        st_["log"] = log
        return st_

    def run(self, case, rp, col, nsteps, rtol=1e-10, maxit=100000, state=None):
        """nsteps explicit steps from the initial state (or from `state`)."""
        if state is None:
            U, ph, sc, v0 = self.init_state(case)
            st_ = {"U": U, "phase": ph, "step": 1, "voln1rel": 0.0, "vol0rel": v0, "solid": sc}
        else:
            st_ = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in state.items()}
        step = C.c_int64(st_["step"])
        vr = C.c_double(st_["voln1rel"])
        sc = C.c_int64(st_["solid"])
        its = C.c_int32()
        rr = C.c_double()
        fl = C.c_int32()
        total_its = 0
        for _ in range(nsteps):
            s = self.lib.or_step(case.n, C.byref(case.params), _d_(case.x), _d_(case.y),
                                 _d_(case.z), _d_(case.vol), _l_(rp), _i_(col), _i_(case.bctypes),
                                 _d_(case.bcvalues), _d_(st_["U"]), _i_(st_["phase"]),
                                 C.byref(step), C.byref(vr), st_["vol0rel"], C.byref(sc), rtol,
                                 maxit, C.byref(its), C.byref(rr), C.byref(fl))
            assert s in (0, 4), s
            total_its += its.value
            if sc.value <= 0:
                break
        st_["step"] = step.value
        st_["voln1rel"] = vr.value
        st_["solid"] = sc.value
        st_["cg_iterations"] = total_its
        st_["relres"] = rr.value
        return st_


class RefHarness:
    """The reference's own grid.h / utils.h compiled from source (oracle/_ref)."""

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def __init__(self):
        self.lib = C.CDLL(REF_SO)
        L = self.lib
        L.ref_beta.restype = _d
        L.ref_beta.argtypes = [_d, _d, _d]
        L.ref_build_clouds_hash.argtypes = [C.c_int, _dp, _dp, _dp, _d, _d, _d, _d, C.c_int, _ip,
                                            _ip, _ip]
        L.ref_residuals.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _ip,
                                    _dp, _dp, _dp, _dp, _ip, _ip, _d, _dp]
        L.ref_residuals_timed.restype = _d
        L.ref_residuals_timed.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                          _ip, _ip, _dp, _dp, _dp, _ip, _ip, _d, C.c_int, _dp]
        L.ref_build_clouds_implant.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                               _d, C.c_int, _ip]
        L.ref_newton_matrix.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip,
                                        _ip, _dp, _dp, _dp, _ip, _ip, _d, C.c_int, C.c_int, _ip,
                                        _ip, _dp, _ip]

    def build_clouds_hash(self, x, y, z, r, cdx, cdy, cdz, max_len=1024):
        n = x.size
        ell = np.zeros((n, max_len), dtype=np.int32)
        ln = np.zeros(n, dtype=np.int32)
        ml = C.c_int32()
        st = self.lib.ref_build_clouds_hash(n, _d_(x), _d_(y), _d_(z), r, cdx, cdy, cdz, max_len,
                                            _i_(ell), _i_(ln), C.byref(ml))
        return st, ell, ln, int(ml.value)

    @staticmethod
    def _param_blocks(p):
        d1 = np.array([p.n_nodes, p.n_steps, 0, 0], dtype=np.float64)
        d2 = p.data2()
        dc = np.concatenate([p.datac(k) for k in range(10)])
        dphi = p.datac_phi()
        return d1, d2, dc, dphi

    def residuals(self, case, ell, lengths, phase, U1, Un, Un0=None, voln1rel=0.0):
        p = case.params
        assert p.S == 10
        d1, d2, dc, dphi = self._param_blocks(p)
        n = case.n
        F = np.zeros((11, n))
        Un0 = np.zeros_like(Un) if Un0 is None else Un0
        self.lib.ref_residuals(n, _d_(d1), _d_(d2), _d_(dc), _d_(dphi), _d_(case.x), _d_(case.y),
                               _d_(case.z), ell.shape[1], _i_(np.ascontiguousarray(ell)),
                               _i_(lengths), _d_(case.bcvalues), _d_(np.ascontiguousarray(U1)),
                               _d_(np.ascontiguousarray(Un)), _d_(np.ascontiguousarray(Un0)),
                               _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                               _i_(case.bctypes), voln1rel, _d_(F))
        return F

    def residuals_timed(self, case, ell, lengths, phase, U1, Un, reps=1, voln1rel=0.0):
        """(F, seconds of the `reps` utils::Residuals calls alone)."""
        p = case.params
        assert p.S == 10
        d1, d2, dc, dphi = self._param_blocks(p)
        n = case.n
        F = np.zeros((11, n))
        t = self.lib.ref_residuals_timed(n, _d_(d1), _d_(d2), _d_(dc), _d_(dphi), _d_(case.x),
                                         _d_(case.y), _d_(case.z), ell.shape[1],
                                         _i_(np.ascontiguousarray(ell)), _i_(lengths),
                                         _d_(case.bcvalues), _d_(np.ascontiguousarray(U1)),
                                         _d_(np.ascontiguousarray(Un)),
                                         _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                         _i_(case.bctypes), voln1rel, int(reps), _d_(F))
        return F, float(t)

    def build_clouds_implant(self, x, y, z, bx, by, bz, r, max_len=1024):
        lengths = np.zeros(x.size, dtype=np.int32)
        st = self.lib.ref_build_clouds_implant(x.size, _d_(x), _d_(y), _d_(z), bx.size, _d_(bx),
                                               _d_(by), _d_(bz), r, max_len, _i_(lengths))
        return st, lengths

    def newton_matrix(self, case, ell, lengths, phase, U1, Un, which, voln1rel=0.0):
        p = case.params
        d1, d2, dc, dphi = self._param_blocks(p)
        n = case.n
        cap = int(lengths.sum()) + n + 16
        rows = np.zeros(cap, dtype=np.int32)
        cols = np.zeros(cap, dtype=np.int32)
        vals = np.zeros(cap)
        nnz = C.c_int32()
        st = self.lib.ref_newton_matrix(n, _d_(d1), _d_(d2), _d_(dc), _d_(dphi), _d_(case.x),
                                        _d_(case.y), _d_(case.z), ell.shape[1],
                                        _i_(np.ascontiguousarray(ell)), _i_(lengths),
                                        _d_(case.bcvalues), _d_(np.ascontiguousarray(U1)),
                                        _d_(np.ascontiguousarray(Un)),
                                        _i_(np.ascontiguousarray(phase, dtype=np.int32)),
                                        _i_(case.bctypes), voln1rel, which, cap, _i_(rows),
                                        _i_(cols), _d_(vals), C.byref(nnz))
        assert st == 0
        k = nnz.value
        return rows[:k], cols[:k], vals[:k]
