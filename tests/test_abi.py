"""CPU checks of the C-ABI library: it loads, exports every symbol include/nnpp.h declares,
its struct layouts match the Python mirrors, and the host-only entry points work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from helpers import ROOT, nn, reference_params


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "nnpp.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(nnpp_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    so = nn._abi.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (nnpp_\w+)", out))
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    assert set(decl) == set(nn._abi.SIGNATURES), "ctypes table out of sync with the header"


def test_struct_layouts_match():
    assert nn.lib.nnpp_abi_params_size() == C.sizeof(nn.Params)
    assert nn.lib.nnpp_abi_stats_size() == C.sizeof(nn.StepStats)
    assert nn.lib.nnpp_abi_adjoint_stats_size() == C.sizeof(nn._abi.AdjointStats)
    assert nn.lib.nnpp_abi_newton_stats_size() == C.sizeof(nn._abi.NewtonStats)


def test_newton_defaults_are_the_references():
    o = nn._abi.NewtonOpts()
    assert nn.lib.nnpp_newton_defaults(C.byref(o)) == 0
    # main.cpp:35-36: EPS 10E-6, MAX_ITER 200 (Newton and linear solves)
    assert o.newton_tol == 10E-6 and o.lin_rtol == 10E-6
    assert o.newton_maxit == 200 and o.lin_maxit == 200
    assert o.start == nn._abi.NEWTON_START_ROTATE
    assert nn.lib.nnpp_newton_defaults(None) == nn._abi.ERR_ARG


def test_params_defaults_and_errors():
    p = nn.Params()
    p.CSolid, p.CSat = 2.0, 0.5
    assert nn.lib.nnpp_params_defaults(C.byref(p), 10, nn.CHEM_REFERENCE) == 0
    assert p.clamp_floor == 1e-8 / 2.0 and p.front_threshold == 0.25 and p.sub_steps == 8
    assert list(p.solid_bonds[:3]) == [1, 0, 0]
    assert nn.lib.nnpp_params_defaults(C.byref(p), 11, nn.CHEM_REFERENCE) == nn._abi.ERR_UNSUPPORTED
    assert nn.lib.nnpp_params_defaults(C.byref(p), 17, nn.CHEM_GENERIC) == nn._abi.ERR_ARG


def test_job_files_roundtrip(tmp_path):
    case = nn.synth.regular(6, S=4, seed=2)
    job = str(tmp_path / "job" / "NNPP")
    nn.formats.write_job(job, case)
    back = nn.formats.read_job(job, 4)
    assert back.n == case.n
    for a in ("x", "y", "z", "bctypes", "bcvalues", "icvalues"):
        assert np.array_equal(getattr(back, a), getattr(case, a)), a
    for f in nn._abi.DATA2_FIELDS + ["r_diss", "front_threshold", "clamp_floor", "phase_eps",
                                     "sub_steps"]:
        assert getattr(back.params, f) == getattr(case.params, f), f
    assert list(back.params.clamp) == list(case.params.clamp)
    assert list(back.params.solid_bonds[:4]) == list(case.params.solid_bonds[:4])
    for k in range(4):
        assert np.array_equal(back.params.datac(k), case.params.datac(k))
    assert np.array_equal(back.params.datac_phi(), case.params.datac_phi())


def test_missing_file_is_an_error(tmp_path):
    p = nn.Params()
    assert nn.lib.nnpp_read_params(str(tmp_path / "nope").encode(), 10, 1, C.byref(p)) == nn._abi.ERR_IO
    v = np.zeros(3)
    assert nn.lib.nnpp_read_vector(str(tmp_path / "nope").encode(), 3, nn._abi.dptr(v)) == nn._abi.ERR_IO


def test_reference_data_files_parse():
    p = reference_params()
    assert p.k4b == 1.0798e-2 and p.JHp == -0.5 and p.T == 298.15 and p.F == 96485.0


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    ctx = C.c_void_p()
    assert nn.lib.nnpp_create(C.byref(ctx), 0) == nn._abi.ERR_HIP


def test_output_csv_roundtrip_and_header(tmp_path):
    rng = np.random.default_rng(0)
    for S in (10, 3):
        n = 17
        x, y, z = rng.uniform(size=(3, n))
        U = rng.uniform(-1, 1, size=(S + 1, n))
        ph = rng.integers(0, 2, n).astype(np.int32)
        path = str(tmp_path / f"o{S}.csv")
        assert nn.lib.nnpp_write_output_csv(path.encode(), n, S, nn._abi.dptr(x), nn._abi.dptr(y),
                                            nn._abi.dptr(z), nn._abi.dptr(U), nn._abi.iptr(ph)) == 0
        head = open(path).readline().strip()
        if S == 10:  # utils.h:2083-2086 column names
            assert head == "X,Y,Z,Mg++,H+,OH-,HCO3-,CO32-,HPO42-,PO43-,Ca++,Na+,Cl-,Phi,Phase"
        U2 = np.zeros_like(U)
        ph2 = np.zeros_like(ph)
        assert nn.lib.nnpp_read_start_csv(path.encode(), n, S, nn._abi.dptr(U2),
                                          nn._abi.iptr(ph2)) == 0
        assert np.array_equal(ph2, ph)
        assert np.abs(U2 - U).max() <= 1e-15 * np.abs(U).max()  # 16 significant digits


def test_driver_usage_and_missing_job(tmp_path):
    exe = os.path.join(ROOT, "nnpp-mg-implant-corrosion-3d_amd", "bin", "nnpp_main")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([exe, "-01", str(tmp_path / "none"), "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "reading parameters" in r.stderr
