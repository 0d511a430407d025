"""ctypes declarations of libnnpp's C ABI (include/nnpp.h).

The library is loaded from this package directory. There is no fallback: if libnnpp.so is
missing, importing the package raises, so nothing can silently run on another path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnnpp.so")

MAX_SPECIES = 16
OK, ERR_ARG, ERR_OVERFLOW, ERR_NONFINITE, ERR_NOCONV, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_IO = range(9)
CHEM_GENERIC, CHEM_REFERENCE = 0, 1
FP64, MIXED = 0, 1
CLAMP_NONE, CLAMP_ALWAYS, CLAMP_LIQUID = 0, 1, 2
SLOT_N, SLOT_N1 = 0, 1
KERNEL_FLUX, KERNEL_POT_SPMV, KERNEL_CG_ITER = 0, 1, 2
FORMAT_AUTO, FORMAT_WIDE_COLUMNS, FORMAT_NO_DICTIONARY, FORMAT_NO_STENCIL = 0, 1, 2, 4
FORMAT_NO_BALL = 8
FORMAT_NO_ZMARCH = 16

# .data2 field names in file order (utils.h:2348-2398); pKH2O is the reference's pKH20
DATA2_FIELDS = [
    "t0", "t1", "dt", "dx", "dy", "dz", "delta", "CSolid", "CSat", "C0Na", "C0Ca", "C0HCO3",
    "C0HPO4", "C0Cl", "pKMgOH2", "pKMgCO3", "pKMgHCO3", "pKMgPO4", "pKCaCO3", "pKCa5OHPO4",
    "pKH2O", "pKHCO3", "pKHPO4", "phim", "phimse", "beta", "alphaA", "alphaC", "k1f", "k1b",
    "k2f", "k2b", "k3f", "k3b", "k4f", "k4b", "k5f", "k5b", "k6f", "k6b", "k7f", "k7b", "k8f",
    "k8b", "k9f", "k9b", "JHp", "T", "R", "F", "vol_red",
]
DATAC_FIELDS = ["z", "D_solid", "D_liquid", "s1", "s2", "K1", "K2"]
DATAC_PHI_FIELDS = ["sigma0", "sigma1", "s_phi", "K_phi"]

_D16 = C.c_double * MAX_SPECIES
_I16 = C.c_int32 * MAX_SPECIES


class Params(C.Structure):
    """Mirror of nnpp_params (utils::parameter + extensions)."""

    _fields_ = (
        [("n_nodes", C.c_int32), ("n_steps", C.c_int32), ("n_species", C.c_int32),
         ("chemistry", C.c_int32)]
        + [(f, C.c_double) for f in DATA2_FIELDS]
        + [(f, _D16) for f in DATAC_FIELDS]
        + [(f, C.c_double) for f in DATAC_PHI_FIELDS]
        + [("solid_bonds", _I16), ("clamp", _I16), ("clamp_floor", C.c_double),
           ("front_threshold", C.c_double), ("phase_eps", C.c_double), ("r_diss", C.c_double),
           ("sub_steps", C.c_int32), ("reserved", C.c_int32)]
    )

    @property
    def S(self):
        return int(self.n_species)

    def data2(self):
        return np.array([getattr(self, f) for f in DATA2_FIELDS], dtype=np.float64)

    def datac(self, k):
        return np.array([getattr(self, f)[k] for f in DATAC_FIELDS], dtype=np.float64)

    def datac_phi(self):
        return np.array([getattr(self, f) for f in DATAC_PHI_FIELDS], dtype=np.float64)

    def copy(self):
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


class StepStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("cg_iterations", C.c_int64), ("cg_spmv_fp64", C.c_int64),
        ("cg_spmv_fp32", C.c_int64), ("last_relres", C.c_double), ("flipped", C.c_int64),
        ("solid_count", C.c_int64), ("voln1rel", C.c_double), ("nnz_flux", C.c_int64),
        ("nnz_ll", C.c_int64), ("pot_col_bytes", C.c_int32), ("pot_weight_bytes", C.c_int32),
        ("pot_stencil", C.c_int32), ("flux_stencil", C.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AdjointStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("recomputed_steps", C.c_int64), ("cg_iterations", C.c_int64),
        ("tape_bytes", C.c_int64), ("last_relres", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


NEWTON_START_ROTATE = 0
NEWTON_START_STATE = 1


class NewtonOpts(C.Structure):
    _fields_ = [
        ("newton_tol", C.c_double), ("lin_rtol", C.c_double), ("newton_maxit", C.c_int32),
        ("lin_maxit", C.c_int32), ("start", C.c_int32), ("reserved", C.c_int32),
    ]


class NewtonStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("newton_iterations", C.c_int32),
        ("unconverged_steps", C.c_int32), ("lin_iterations", C.c_int32),
        ("cg_iterations", C.c_int32), ("flipped", C.c_int32), ("solid_count", C.c_int64),
        ("voln1rel", C.c_double), ("last_resmax", C.c_double), ("lin_relres_max", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NnppError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nnpp error {code}: {msg}")
        self.code = code


_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_i32, _i64, _d = C.c_int32, C.c_int64, C.c_double

# name -> (restype, argtypes); every symbol include/nnpp.h declares
SIGNATURES = {
    "nnpp_abi_params_size": (C.c_int64, []),
    "nnpp_abi_stats_size": (C.c_int64, []),
    "nnpp_params_defaults": (C.c_int, [C.POINTER(Params), _i32, _i32]),
    "nnpp_read_params": (C.c_int, [C.c_char_p, _i32, _i32, C.POINTER(Params)]),
    "nnpp_read_vector": (C.c_int, [C.c_char_p, _i64, _dp]),
    "nnpp_read_ivector": (C.c_int, [C.c_char_p, _i64, _ip]),
    "nnpp_read_nodes": (C.c_int, [C.c_char_p, _i64, _dp, _dp, _dp]),
    "nnpp_neighbors_to_ell": (C.c_int, [_i32, _lp, _ip, _i32, _ip, _ip]),
    "nnpp_neighbors_from_ell": (C.c_int, [_i32, _i32, _ip, _ip, _lp, _ip]),
    "nnpp_write_output_csv": (C.c_int, [C.c_char_p, _i32, _i32, _dp, _dp, _dp, _dp, _ip]),
    "nnpp_read_start_csv": (C.c_int, [C.c_char_p, _i32, _i32, _dp, _ip]),
    "nnpp_create": (C.c_int, [C.POINTER(_p), _i32]),
    "nnpp_destroy": (C.c_int, [_p]),
    "nnpp_last_error": (C.c_char_p, [_p]),
    "nnpp_set_params": (C.c_int, [_p, C.POINTER(Params)]),
    "nnpp_set_nodes": (C.c_int, [_p, _i32, _dp, _dp, _dp, _dp]),
    "nnpp_build_neighbors": (C.c_int, [_p, _d, _d, _d, _d, _i32, _lp, _ip]),
    "nnpp_get_neighbors": (C.c_int, [_p, _lp, _ip]),
    "nnpp_set_neighbors": (C.c_int, [_p, _lp, _ip]),
    "nnpp_set_bc": (C.c_int, [_p, _ip, _dp]),
    "nnpp_set_state": (C.c_int, [_p, _i32, _dp]),
    "nnpp_get_state": (C.c_int, [_p, _i32, _dp]),
    "nnpp_set_phase": (C.c_int, [_p, _ip]),
    "nnpp_get_phase": (C.c_int, [_p, _ip]),
    "nnpp_init_state": (C.c_int, [_p, _dp]),
    "nnpp_set_time": (C.c_int, [_p, _i64, _d, _d]),
    "nnpp_get_time": (C.c_int, [_p, _lp, _dp, _dp]),
    "nnpp_residuals": (C.c_int, [_p, _d, _i32, _i32, _dp]),
    "nnpp_poisson_apply": (C.c_int, [_p, _dp, _dp]),
    "nnpp_potential_operator_apply": (C.c_int, [_p, _i32, _dp, _dp]),
    "nnpp_set_solver": (C.c_int, [_p, _d, _i32, _d]),
    "nnpp_set_format": (C.c_int, [_p, _i32]),
    "nnpp_poisson_solve": (C.c_int, [_p, _i32, _ip, _dp]),
    "nnpp_step": (C.c_int, [_p, _i32, _i32, _i32, C.POINTER(StepStats)]),
    "nnpp_front_update": (C.c_int, [_p, _d, _ip]),
    "nnpp_state_device_ptr": (C.c_int, [_p, _i32, _i32, C.POINTER(_p)]),
    "nnpp_stream": (C.c_int, [_p, C.POINTER(_p)]),
    "nnpp_time_kernel": (C.c_int, [_p, _i32, _i32, _i32, _dp]),
    "nnpp_tape_begin": (C.c_int, [_p, _i32]),
    "nnpp_tape_adjoint": (C.c_int, [_p, _dp, _dp, _dp, C.POINTER(AdjointStats)]),
    "nnpp_tape_end": (C.c_int, [_p]),
    "nnpp_abi_adjoint_stats_size": (C.c_int64, []),
    "nnpp_implant_classify": (C.c_int, [_p, _i64, _dp, _dp, _dp, _d, _i32, _ip, _dp]),
    "nnpp_implant_classify_noalloc": (C.c_int, [_p, _i64, _dp, _dp, _dp, _d, _i32, _ip, _dp]),
    "nnpp_voronoi_volumes": (C.c_int, [_p, _d, _d, _d, _dp]),
    "nnpp_newton_defaults": (C.c_int, [C.POINTER(NewtonOpts)]),
    "nnpp_step_implicit": (C.c_int, [_p, _i32, _i32, _i32, C.POINTER(NewtonOpts),
                                     C.POINTER(NewtonStats)]),
    "nnpp_jacobian_apply": (C.c_int, [_p, _d, _dp, _dp, _dp]),
    "nnpp_abi_newton_stats_size": (C.c_int64, []),
}

_lib = None


def load():
    """Load libnnpp.so from the package directory (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def lptr(a):
    return a.ctypes.data_as(_lp)
