"""NNPP step benchmark on one MI355X.

Workload: the 128^3 regular synthetic case (2,097,152 nodes, 249,027,944 directed bonds at
delta = 3 dx), 11 species + phi, generic chemistry, Mg sphere, Dirichlet faces; each step =
fused 11-species flux + explicit update + potential solve to relative residual 1e-10
(warm-started CG) + front update. A "bond-update" is one directed bond processed once in one
operator application (the flux counts it once for all species; each potential-operator
application counts its liquid-liquid bonds).

Reported value: bond-updates/s of the mixed-precision step (inputs resident in HBM), over the
timed steps after the warm-up steps (default 2; the potential's extrapolated warm start uses the
two previous potentials and reaches its steady iteration count after a few steps, so a longer
warm-up gives shorter steps with fewer CG iterations each). The FP64 step is timed from the same state; the mixed/FP64 ratio and
the per-field error of the mixed run against the FP64 run after 1000 steps are reported beside
it. `roofline` is for the dominant kernel (on the box lattice the fused FP32 CG iteration,
k_zm_m_cgs), timed with HIP events on the library's stream, with the HBM traffic per launch from
the counter summary in profiles/ when it was taken on the same sources; `cpu_baseline` times the
reference's own utils::Residuals (compiled from its sources into oracle/_ref) once on the same
lattice, or the oracle's port of it where that library is absent.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def field_errors(a, b):
    out = []
    for f in range(a.shape[0]):
        d = a[f] - b[f]
        s = max(np.abs(b[f]).max(), 1e-300)
        out.append({"max": float(np.abs(d).max() / s),
                    "l2": float(np.linalg.norm(d) / max(np.linalg.norm(b[f]), 1e-300))})
    return out


def run_mode(nn, case, rp, col, precision, warmup, steps, rtol, inner_rtol=1e-3):
    import torch
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    pb.set_neighbors(rp, col)
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    pb.set_solver(rtol=rtol, maxit=50000, inner_rtol=inner_rtol)
    if warmup:
        pb.step(warmup, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = pb.step(steps, precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return pb, st, dt


def cpu_baseline(nn, case, rp, col, budget_s=20.0):
    """The reference's own CPU path when oracle/_ref holds it (utils::Residuals of
    src/helpers/utils.h:27-842 compiled from the reference sources, OpenMP on the host cores,
    its column-major ELL neighbour layout and per-field passes), timed for one evaluation on the
    same lattice and neighbour lists (the reference's fixed 10 species + phi, generic values);
    otherwise the oracle's port of the same loops (CSR layout, friendlier to the caches than
    the reference's, so it flatters the CPU)."""
    orc = ge.load_oracle()
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nnz = int(rp[-1])
    n = case.n
    if orc.RefHarness.available():
        p10 = case.params.copy()
        p10.n_species = 10
        ml = int(np.diff(rp).max()) + 1
        ell = np.empty(n * ml, dtype=np.int32)
        ln = np.empty(n, dtype=np.int32)
        st = nn.lib.nnpp_neighbors_to_ell(n, nn._abi.lptr(rp), nn._abi.iptr(col), ml,
                                          nn._abi.iptr(ell), nn._abi.iptr(ln))
        if st != 0:
            raise RuntimeError(f"nnpp_neighbors_to_ell: {st}")
        # the C ABI writes the reference's column-major n x max_len; the harness takes row-major
        ell = np.ascontiguousarray(ell.reshape(ml, n).T)
        U = np.ascontiguousarray(case.icvalues[:11])
        c10 = nn.synth.Case(p10, case.x, case.y, case.z, case.vol, case.bctypes,
                            np.ascontiguousarray(case.bcvalues[:11]), U)
        ph = (U[0] > p10.phase_eps).astype(np.int32)
        R = orc.RefHarness()
        _, secs = R.residuals_timed(c10, ell, ln, ph, U, U, reps=1)
        del ell
        return {"value": nnz / secs, "unit": "bond-updates/s", "cores": nthreads,
                "kind": "reference",
                "sample": f"the reference's utils::Residuals (compiled from its sources, OpenMP,"
                          f" column-major ELL, 11 per-field passes of its 10 species + phi) once on"
                          f" the same {n}-node lattice and {nnz} bonds: {secs:.2f} s",
                "residuals_s": secs}
    O = orc.Oracle(nn.Params)
    U, ph, _, _ = O.init_state(case)
    t0 = time.perf_counter()
    O.residuals(case, rp, col, ph, U, U)
    t_flux = time.perf_counter() - t0
    return {"value": nnz / t_flux, "unit": "bond-updates/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle/nnpp_oracle.c (the oracle's CSR port of the reference loops, "
                      f"OpenMP) once on the same {n}-node case: {nnz} bonds, {case.S + 1} "
                      f"per-field passes, {t_flux:.2f} s",
            "residuals_s": t_flux}


def kernels_only(nn, case, reps):
    """Time the hot kernels alone (for rocprofv3 runs): no solve runs first, so every traced
    launch of these kernels is a full one (their cost does not depend on the vector values)."""
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    nnz, _ = pb.build_clouds_hash()
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    out = {"nodes": case.n, "bonds": nnz}
    for name, k, prec in (("pot_spmv_fp32", nn._abi.KERNEL_POT_SPMV, nn.MIXED),
                          ("pot_spmv_fp64", nn._abi.KERNEL_POT_SPMV, nn.FP64),
                          ("cg_iter_fp32", nn._abi.KERNEL_CG_ITER, nn.MIXED),
                          ("flux_fp64", nn._abi.KERNEL_FLUX, nn.FP64),
                          ("flux_mixed", nn._abi.KERNEL_FLUX, nn.MIXED)):
        out[name + "_ms"] = pb.time_kernel(k, prec, reps)
    pb.close()
    print(json.dumps(out))


def validate(nn, kind, ng, nsteps, rtol):
    """Mixed against FP64 after nsteps steps from the same initial state: max and L2 relative
    error per field (species 1..S, then phi)."""
    case = (nn.synth.regular if kind == "regular" else nn.synth.screw)(ng, S=11, seed=0)
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    pb.build_clouds_hash()
    rp, col = pb.neighbors()
    pb.close()
    states, info = {}, {}
    for name, prec in (("fp64", nn.FP64), ("mixed", nn.MIXED)):
        pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
        pb.set_neighbors(rp, col)
        pb.set_bc(case.bctypes, case.bcvalues)
        pb.init_state(case.icvalues)
        pb.set_solver(rtol=rtol, maxit=50000)
        t0 = time.perf_counter()
        its = flips = 0
        done = 0
        while done < nsteps:
            k = min(100, nsteps - done)
            st = pb.step(k, prec)
            if st["status"] != 0:
                raise RuntimeError(f"{name}: step status {st['status']}")
            done += k
            its += st["cg_iterations"]
            flips += st["flipped"]
            log(f"[validate] {name}: {done}/{nsteps} steps, {time.perf_counter() - t0:.1f} s")
        info[name] = {"s": time.perf_counter() - t0, "cg_iterations": its, "flipped": flips,
                      "voln1rel": st["voln1rel"]}
        states[name] = (pb.get_state(), pb.get_phase())
        pb.close()
    errs = field_errors(states["mixed"][0], states["fp64"][0])
    print(json.dumps({"validate": f"{kind} {ng}^3", "steps": nsteps, "rtol": rtol,
                      "max_rel": max(e["max"] for e in errs),
                      "max_l2": max(e["l2"] for e in errs), "fields": errs,
                      "phase_equal": bool(np.array_equal(states["mixed"][1], states["fp64"][1])),
                      "runs": info}))


def adjoint_workload(nn, ng, nsteps, every=20):
    """Differentiation workload: regular ng^3, 11 species, nsteps FP64 steps. Loss = L2 mismatch
    of the final Mg2+ and phi fields against a target made with every diffusivity x1.05 (case
    seed 1). Gradients with respect to D_liquid, D_solid and r_diss through the torch binding,
    checkpointed (every `every` steps) and with the full tape; central differences (h = 1e-6
    relative) for a subset of the parameters; device bytes of each tape."""
    import torch
    case = nn.synth.regular(ng, S=11, seed=1)
    S = case.params.S
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    pb.build_clouds_hash()
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    pb.set_solver(rtol=1e-13, maxit=100000)
    start = pb.snapshot()
    Dl0 = torch.tensor([case.params.D_liquid[k] for k in range(S)], dtype=torch.float64)
    Ds0 = torch.tensor([case.params.D_solid[k] for k in range(S)], dtype=torch.float64)
    rd0 = torch.tensor(case.params.r_diss, dtype=torch.float64)
    sim = nn.autograd.simulate
    with torch.no_grad():
        target = sim(pb, Dl0 * 1.05, Ds0 * 1.05, rd0, nsteps, start, every).clone()

    def loss_of(U):
        return ((U[0] - target[0]) ** 2).sum() + ((U[S] - target[S]) ** 2).sum()

    runs = {}
    for k_ in (every, 1):
        Dl, Ds, rd = (t.clone().requires_grad_(True) for t in (Dl0, Ds0, rd0))
        t0 = time.perf_counter()
        U = sim(pb, Dl, Ds, rd, nsteps, start, k_)
        t1 = time.perf_counter()
        loss = loss_of(U)
        loss.backward()
        t2 = time.perf_counter()
        st = pb.last_adjoint_stats
        runs[k_] = {"forward_s": t1 - t0, "backward_s": t2 - t1, "loss": float(loss),
                    "tape_bytes": st["tape_bytes"], "recomputed_steps": st["recomputed_steps"],
                    "adjoint_cg_iterations": st["cg_iterations"],
                    "grad": np.concatenate([Dl.grad.numpy(), Ds.grad.numpy(),
                                            [float(rd.grad)]]).tolist()}
        log(f"[adjoint] checkpoint every {k_}: forward {t1 - t0:.2f} s, backward {t2 - t1:.2f} s,"
            f" tape {st['tape_bytes'] / 2**20:.0f} MiB")
    pb.tape_end()
    g = np.array(runs[every]["grad"])
    same = bool(np.array_equal(g, np.array(runs[1]["grad"])))
    # central differences of the forward for a subset of the parameters
    names = [("D_liquid", 0), ("D_liquid", 5), ("D_solid", 0), ("D_solid", 7), ("r_diss", None)]
    fd = []
    for name, k in names:
        base = getattr(case.params, name) if k is None else getattr(case.params, name)[k]
        h = 1e-6 * abs(base)
        L = []
        for sgn in (1, -1):
            p = case.params.copy()
            if k is None:
                setattr(p, name, base + sgn * h)
            else:
                getattr(p, name)[k] = base + sgn * h
            pb.set_params(p)
            pb.restore(start)
            pb.step(nsteps, nn.FP64)
            L.append(float(loss_of(torch.from_numpy(pb.get_state()))))
        q = k if name == "D_liquid" else (S + k if name == "D_solid" else 2 * S)
        fdv = (L[0] - L[1]) / (2 * h)
        fd.append({"param": f"{name}[{k}]" if k is not None else name, "adjoint": float(g[q]),
                   "fd": fdv, "rel": abs(g[q] - fdv) / max(abs(fdv), 1e-300)})
        log(f"[adjoint] fd {fd[-1]}")
    pb.close()
    print(json.dumps({"adjoint": f"regular {ng}^3, {S} species, {nsteps} FP64 steps",
                      "loss": runs[every]["loss"], "checkpointed": runs[every],
                      "full_tape": runs[1], "checkpoint_every": every,
                      "checkpointed_equals_full_tape": same, "finite_differences": fd,
                      "max_fd_rel": max(f["rel"] for f in fd)}))


def copy_probe(mib):
    """A plain device copy of known size, to scale FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(mib * 2 ** 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4,
                      "launches": 3}))


KERNEL_SOURCES = ["nnpp-mg-implant-corrosion-3d_amd/csrc/k_potential.hip",
                  "nnpp-mg-implant-corrosion-3d_amd/csrc/k_flux.hip",
                  "nnpp-mg-implant-corrosion-3d_amd/csrc/nnpp_internal.hpp"]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_regular128.json")


def sources_digest():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def kernel_short_name(full):
    n = full.replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    return n.split("(")[0]


def _db_rows(db, sql):
    import sqlite3
    con = sqlite3.connect(db)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def summarize_rocprof(trace_db, fetch_db=None, write_db=None, copyf_db=None, copyw_db=None,
                      copy_bytes=None, out_json=None):
    """Text summary of rocprofv3 results (kernel trace + FETCH_SIZE/WRITE_SIZE counter runs).

    Counter runs are scaled by a plain copy of known size: scale = bytes moved / counter KiB."""
    lines = ["| kernel | calls | total ms | avg us |", "|---|---|---|---|"]  # top_kernels: us
    rows = _db_rows(trace_db, "select name, total_calls, total_duration, average from top_kernels "
                              "order by total_duration desc limit 16")
    for name, calls, tot, avg in rows:
        lines.append(f"| `{kernel_short_name(name)[:70]}` | {calls} | {tot / 1e3:.3f} | {avg:.1f} |")
    out = {"digest": sources_digest(), "kernels": {}}
    if fetch_db and write_db and copyf_db and copyw_db and copy_bytes:
        def per_kernel(db):
            r = _db_rows(db, "select kernel_name, avg(value), count(*), avg(duration) from "
                             "counters_collection group by kernel_name")
            return {kernel_short_name(k): (v, c, d) for k, v, c, d in r}
        cf = max(v for v, c, d in per_kernel(copyf_db).values())
        cw = max(v for v, c, d in per_kernel(copyw_db).values())
        fs, ws = copy_bytes / cf, copy_bytes / cw  # bytes per counter unit
        out.update({"fetch_scale": fs, "write_scale": ws})
        fk, wk = per_kernel(fetch_db), per_kernel(write_db)
        lines += ["", f"Counter scale from a {copy_bytes / 2**30:.0f} GiB copy: FETCH_SIZE x{fs:.1f}"
                      f" B/unit, WRITE_SIZE x{ws:.1f} B/unit.", "",
                  "| kernel | dispatches | FETCH bytes/dispatch | WRITE bytes/dispatch | avg us |",
                  "|---|---|---|---|---|"]
        for k in sorted(fk, key=lambda k: -fk[k][0] * fk[k][1])[:10]:
            f_ = fk[k][0] * fs
            w_ = wk.get(k, (0, 0, 0))[0] * ws
            out["kernels"][k] = {"fetch": f_, "write": w_, "avg_ns": fk[k][2]}
            lines.append(f"| `{k[:70]}` | {fk[k][1]} | {f_:.4g} | {w_:.4g} | {fk[k][2] / 1e3:.1f} |")
    if out_json:
        with open(out_json, "w") as fh:
            json.dump(out, fh, indent=1)
    print("\n".join(lines))


def roofline_entry(t_ms, fmt_bytes, csr_bytes, stencil, cb, wb, zmarch=False):
    """Roofline of the dominant kernel: `achieved` = the bytes it must move per launch
    (compulsory traffic: the stored operator form's per-row words and the vectors, each value
    counted once) / the measured launch time. On a full box lattice that kernel is the fused
    FP32 CG iteration (SpMV, the previous iteration's vector updates and the dot products in
    one launch), elsewhere the FP32 potential SpMV. The survey's CSR figure (a 4-byte column
    per liquid bond) is kept beside it as `csr_model_GBps`: the stencil forms never read those
    bytes, so that ratio is not a bandwidth."""
    kname = "k_zm_m_cgs" if zmarch else ("k_m_spmv_st" if stencil else "k_m_spmv")
    label = ("FP32 single-reduction CG iteration: M u, vector updates, gamma/delta/r.r"
             if zmarch else "FP32 potential SpMV + p.q")
    traffic = measured_traffic(kname)
    ach = fmt_bytes / (t_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
           "kernel": f"{kname} ({label})", "avg_launch_ms": t_ms,
           "bytes_per_launch": fmt_bytes,
           "csr_model_bytes_per_launch": csr_bytes,
           "csr_model_GBps": csr_bytes / (t_ms * 1e-3) / 1e9,
           "operator_format": ({"stencil": True, "zmarch": True, "flag_bytes_per_row": 1}
                               if zmarch else
                               {"stencil": True, "mask_bytes_per_row": 16} if stencil else
                               {"stencil": False, "col_bytes": cb, "weight_bytes": wb})}
    if traffic:
        out["hbm_measured_GBps"] = traffic / (t_ms * 1e-3) / 1e9
    assert out["frac"] <= 1.0, f"roofline fraction above 1: {out}"
    return out


def measured_traffic(kernel_prefix):
    """HBM bytes per dispatch from the committed counter summary, if it matches these sources."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("digest") != sources_digest():
        return None
    for k, v in t.get("kernels", {}).items():
        if kernel_prefix in k:
            return v["fetch"] + v["write"]
    return None


def implicit_workload(nn, kind, ng, steps, warmup, lin_rtol, newton_tol, precision="fp64"):
    """The reference's implicit step (modified Newton with the frozen Jacobian, main.cpp's
    tolerances unless overridden) on the same synthetic case: per-step time and the iteration
    counts behind it. Not the headline metric (that is the explicit step)."""
    import torch
    case = (nn.synth.regular if kind == "regular" else nn.synth.screw)(ng, S=11, seed=0)
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    nnz, _ = pb.build_clouds_hash()
    pb.set_bc(case.bctypes, case.bcvalues)
    pb.init_state(case.icvalues)
    opts = dict(lin_rtol=lin_rtol, newton_tol=newton_tol)
    prec = nn.MIXED if precision == "mixed" else nn.FP64  # the potential solves' precision
    if warmup:
        pb.step_implicit(warmup, prec, **opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = pb.step_implicit(steps, prec, **opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"workload": f"implicit (reference Newton) step, {kind} {ng}^3, 11 species + phi",
           "potential_precision": precision,
           "nodes": case.n, "bonds": nnz, "steps": st["steps"], "warmup": warmup,
           "ms_per_step": dt / max(st["steps"], 1) * 1e3, "newton_tol": newton_tol,
           "lin_rtol": lin_rtol, "stats": st}
    nj = 2 * st["lin_iterations"] + st["newton_iterations"] * 1  # Jacobian applies + residuals
    out["bond_passes_per_step"] = nj / max(st["steps"], 1)
    pb.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--case", choices=["regular", "screw"], default="regular",
                    help="regular lattice with a Mg sphere (default) or the jittered screw "
                         "geometry with per-node volumes")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--inner-rtol", type=float, default=1e-3,
                    help="mixed mode: residual-norm reduction of each FP32 refinement pass")
    ap.add_argument("--err-steps", type=int, default=1000,
                    help="steps after which the mixed fields are compared with the FP64 ones")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the hot kernels (for counter runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter scale)")
    ap.add_argument("--summarize-rocprof", nargs="+", metavar="DB",
                    help="trace.db [fetch.db write.db copyfetch.db copywrite.db copy_MiB]")
    ap.add_argument("--traffic-out", default=None)
    ap.add_argument("--validate", type=int, default=0, metavar="STEPS",
                    help="only run the FP64 and mixed steps STEPS times from the same state and "
                         "report the per-field error of mixed against FP64")
    ap.add_argument("--adjoint", action="store_true",
                    help="only run the differentiation workload (64^3, 200 FP64 steps, "
                         "checkpoints every 20; loss against a +5%% diffusivity target)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-grid", type=int, default=64)
    ap.add_argument("--implicit-precision", choices=["fp64", "mixed"], default="fp64",
                    help="--implicit: precision of the potential solves (species stay FP64)")
    ap.add_argument("--implicit", action="store_true",
                    help="only time the reference's implicit (Newton) step on the case")
    ap.add_argument("--lin-rtol", type=float, default=1e-5)
    ap.add_argument("--newton-tol", type=float, default=1e-5)
    a = ap.parse_args()
    if a.summarize_rocprof:
        d = a.summarize_rocprof
        if len(d) >= 6:
            summarize_rocprof(d[0], d[1], d[2], d[3], d[4], int(d[5]) * 2 ** 20, a.traffic_out)
        else:
            summarize_rocprof(d[0])
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    import torch
    nn = ge.load_package()
    ng = a.grid
    if a.kernels_only:
        kernels_only(nn, nn.synth.regular(ng, S=11, seed=0), 20)
        return
    if a.validate:
        validate(nn, a.case, ng, a.validate, a.rtol)
        return
    if a.adjoint:
        adjoint_workload(nn, a.adjoint_grid, a.adjoint_steps)
        return
    if a.implicit:
        implicit_workload(nn, a.case, ng, a.steps, a.warmup, a.lin_rtol, a.newton_tol,
                          a.implicit_precision)
        return
    log(f"[bench] building {a.case} {ng}^3 case")
    case = (nn.synth.regular if a.case == "regular" else nn.synth.screw)(ng, S=11, seed=0)
    # neighbour list built on the GPU once, shared by both runs
    pb = nn.Problem(case.params, case.x, case.y, case.z, case.vol, device=0)
    t0 = time.perf_counter()
    nnz, ml = pb.build_clouds_hash()
    torch.cuda.synchronize()
    t_nb = time.perf_counter() - t0
    rp, col = pb.neighbors()
    pb.close()
    log(f"[bench] neighbours: nnz={nnz} max_len={ml} in {t_nb:.3f}s")

    res = {}
    for name, prec in (("fp64", nn.FP64), ("mixed", nn.MIXED)):
        log(f"[bench] {name}: warmup {a.warmup}, timed {a.steps}")
        pbm, st, dt = run_mode(nn, case, rp, col, prec, a.warmup, a.steps, a.rtol, a.inner_rtol)
        bu_flux = st["steps"] * st["nnz_flux"]
        bu_cg = (st["cg_spmv_fp64"] + st["cg_spmv_fp32"]) * st["nnz_ll"]
        res[name] = {"time_s": dt, "stats": st, "bu": bu_flux + bu_cg, "bu_flux": bu_flux,
                     "bu_cg": bu_cg, "U": pbm.get_state(), "pb": pbm}
        log(f"[bench] {name}: {dt / a.steps * 1e3:.2f} ms/step, cg iters {st['cg_iterations']}, "
            f"relres {st['last_relres']:.2e}")

    # precision gate of the metric: both runs continue from where they are to err_steps steps
    # from the same initial state, and the mixed fields are compared with the FP64 ones
    for name, prec in (("fp64", nn.FP64), ("mixed", nn.MIXED)):
        r_ = res[name]
        done = a.warmup + a.steps
        while done < a.err_steps:
            st_e = r_["pb"].step(min(250, a.err_steps - done), prec)
            if st_e["status"] != 0:
                raise RuntimeError(f"{name}: step status {st_e['status']}")
            done += st_e["steps"]
        log(f"[bench] {name}: {done} steps for the error check")
        r_["U"] = r_["pb"].get_state()
        r_["phase"] = r_["pb"].get_phase()
        r_["err_steps"] = done
    mixed = res["mixed"]
    fp64 = res["fp64"]
    errs = field_errors(mixed["U"], fp64["U"])
    max_err = max(e["max"] for e in errs)
    max_l2 = max(e["l2"] for e in errs)
    value = mixed["bu"] / mixed["time_s"]
    value64 = fp64["bu"] / fp64["time_s"]

    # roofline of the dominant kernel: potential SpMV (FP32 variant in the mixed run)
    pbm = mixed["pb"]
    n = case.n
    nll = mixed["stats"]["nnz_ll"]
    reps = 50
    t_spmv32 = pbm.time_kernel(nn._abi.KERNEL_POT_SPMV, nn.MIXED, reps)
    t_spmv64 = pbm.time_kernel(nn._abi.KERNEL_POT_SPMV, nn.FP64, reps)
    t_flux64 = pbm.time_kernel(nn._abi.KERNEL_FLUX, nn.FP64, 10)
    t_flux32 = pbm.time_kernel(nn._abi.KERNEL_FLUX, nn.MIXED, 10)
    # Algorithmic bytes of one potential-SpMV launch, the per-unit figure of the CSR operator:
    # a 4-byte column per liquid bond, 8 bytes of row pointer per row, x in and y out (FP64).
    alg_bytes = 4 * nll + 8 * (n + 1) + 16 * n
    # Bytes the stored form itself must move per launch (the operator as held on the device):
    # stencil form: a 16-byte mask per row; column form: per liquid bond the column (2-byte
    # step or 4 bytes) and the weight (1-byte class, or the value), plus the first column per
    # row; both: p_i (each gathered p_j counted once), dd_i, q_i.
    st_ = mixed["stats"]
    stencil = bool(st_["pot_stencil"])
    cb, wb32 = st_["pot_col_bytes"], st_["pot_weight_bytes"]
    wb64 = 1 if wb32 == 1 else (0 if stencil else 8)
    zmarch = st_["pot_stencil"] == 4
    if zmarch:  # no per-row words: a free flag per node
        bytes32 = n * (1 + 3 * 4)
        bytes64 = n * (1 + 3 * 8)
        # the fused FP32 CG iteration: free flag, diagonal, r w s D^-1 in, p x in, s r w p x out
        bytes_cgit = n * (1 + 4 + 4 * 4 + 2 * 4 + 5 * 4)
    elif stencil:
        bytes32 = n * (16 + 3 * 4)
        bytes64 = n * (16 + 3 * 8)
    else:
        bytes32 = nll * (cb + wb32) + n * (4 + 3 * 4)
        bytes64 = nll * (cb + wb64) + n * (4 + 3 * 8)
    S = case.S
    # flux bytes per launch: column form, per bond a column and a weight; tile stencil form
    # (lattices), a 16-byte bond mask per row; z-marching form (full box lattices), nothing per
    # bond or row; all: the fields read once, phase, Dirichlet mask and the updated fields
    fform = st_["flux_stencil"]  # the mixed run's form
    fform64 = fp64["stats"]["flux_stencil"]  # the FP64 run's form
    fbytes = lambda form, wb: ((0 if form == 3 else n * 16 if form else nnz * (4 + wb))  # noqa: E731
                               + n * (8 * (S + 1) + 1 + 2 + 8 * S))
    t_cgit = pbm.time_kernel(nn._abi.KERNEL_CG_ITER, nn.MIXED, reps) if zmarch else None
    kern = {
        "pot_spmv_fp32": {"ms": t_spmv32, "alg_GBps": alg_bytes / (t_spmv32 * 1e-3) / 1e9,
                          "format_GBps": bytes32 / (t_spmv32 * 1e-3) / 1e9},
        "pot_spmv_fp64": {"ms": t_spmv64, "alg_GBps": alg_bytes / (t_spmv64 * 1e-3) / 1e9,
                          "format_GBps": bytes64 / (t_spmv64 * 1e-3) / 1e9},
        "flux_fp64": {"ms": t_flux64,
                      "form": {3: "z-marching", 1: "stencil"}.get(fform64, "columns"),
                      "GBps": fbytes(fform64, 8) / (t_flux64 * 1e-3) / 1e9},
        "flux_mixed": {"ms": t_flux32,
                       "form": {2: "stencil (fp32 bond sums)", 1: "stencil (fp64)"}.get(
                           fform, "columns (fp32)"),
                       "GBps": fbytes(fform, 4) / (t_flux32 * 1e-3) / 1e9},
    }
    if zmarch:
        kern["cg_iter_fp32"] = {"ms": t_cgit, "GBps": bytes_cgit / (t_cgit * 1e-3) / 1e9,
                                "form": "fused single-reduction CG iteration"}
    for r_ in res.values():
        r_["pb"].close()

    cpu = None
    if not a.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(nn, case, rp, col, a.cpu_budget)

    out = {
        "metric": "bond-updates/s per NNPP step, mixed vs FP64 @ rel err<=1e-6; % of HBM peak",
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mixed["time_s"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 CG inner iterations; f32 flux bond sums "
                  + ("(FP32 differences per bond, partial sums per ball plane added in f64)"
                     if fform == 2 else "(f64 stencil form)" if fform == 1 else "(Kahan)")
                  + "; f64 state, dots, residual checks, reactions and update"),
        "data": "synthetic (seeded lattice case in the reference's input layout)",
        "config": {"workload": (f"regular {ng}^3 lattice, 11 species + phi, delta=3dx, Mg sphere "
                                f"r=0.3, Dirichlet faces, CG rtol {a.rtol:g}" if a.case == "regular"
                                else f"jittered {ng}^3 lattice in a fluid cylinder around a screw-"
                                     f"like solid, per-node volumes, 11 species + phi, "
                                     f"CG rtol {a.rtol:g}"),
                   "nodes": n, "bonds": nnz, "liquid_bonds": nll, "species": S},
        "fp64": {"value": value64, "ms_per_step": fp64["time_s"] / a.steps * 1e3,
                 "cg_iterations": fp64["stats"]["cg_iterations"],
                 "bu_flux": fp64["bu_flux"], "bu_cg": fp64["bu_cg"]},
        "mixed": {"cg_inner_iterations": mixed["stats"]["cg_iterations"],
                  "spmv_fp64": mixed["stats"]["cg_spmv_fp64"],
                  "spmv_fp32": mixed["stats"]["cg_spmv_fp32"],
                  "bu_flux": mixed["bu_flux"], "bu_cg": mixed["bu_cg"]},
        "mixed_over_fp64": (fp64["time_s"] / mixed["time_s"]),
        "mixed_vs_fp64_err": {"max_rel": max_err, "max_l2": max_l2, "fields": errs,
                              "after_steps": mixed["err_steps"],
                              "phase_equal": bool(np.array_equal(mixed["phase"], fp64["phase"])),
                              "within_1e-6": bool(max_err <= 1e-6 and max_l2 <= 1e-6)},
        "neighbor_build_s": t_nb,
        "kernels": kern,
        "roofline": (roofline_entry(t_cgit, bytes_cgit, alg_bytes, stencil, cb, wb32, zmarch)
                     if zmarch else
                     roofline_entry(t_spmv32, bytes32, alg_bytes, stencil, cb, wb32, zmarch)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
