"""Throughput of the DSEM RHS + RK4(5) step on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Workload: Taylor-Green vortex on the periodic cube, 22^3 hex8 elements at
p = 8 (9^3 Gauss points per element, 38,811,960 DOF), FP64, state resident in
HBM.  A timed step is one full Carpenter-Kennedy RK4(5) step: five RHS
evaluations with their register updates.  The metric is GDOF/s per RHS
evaluation = DOF * 5 * steps / time / 1e9.

Prints one JSON line with the throughput, a roofline entry for the dominant
kernel (per-kernel HIP-event durations on the library's stream) and a CPU
baseline measured with the oracle on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dsem_amd import cases  # noqa: E402
from dsem_amd import mesh as M  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per launch measured with rocprofv3 on this workload (22^3, p=8,
# FP64, per-point metrics): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors
# that make the counters read 1 GiB for a 1 GiB device copy.  Source:
# profiles/tgv22p8_fp64.md.
MEASURED_TRAFFIC = {(22, 8, "fp64", "per_point"): {"k_volume": 2.449e9, "k_update": 2.015e9, "k_face_inv": 0.541e9,
                                                   "k_face_visc": 0.477e9, "k_traces": 0.583e9}}


def kernel_work(n_elem: int, nmort: int, p: int) -> dict:
    """Algorithmic bytes / flops per launch for every kernel (FP64).

    Bytes count each array the kernel must read or write once (per-point
    metrics, the parity mode the bench runs).  `flops` of the volume kernel
    are the tensor-product contraction flops alone, 180 nc ncg^3 per element
    (Int_fluxes 30, Gauss_Deriv 30, gradient prolongation 90 and the
    divergence 30, each times nc ncg^3): the figure its MFMA roofline
    fraction is quoted on.  `pointwise_flops` (an estimate of ~260 flops of
    Euler + viscous flux per staggered point) is reported beside it and is
    not part of the fraction.
    """
    ncg, nc = p + 1, p + 2
    npt, ns, nf = ncg ** 3, nc * ncg ** 2, ncg ** 2
    K, Mm = n_elem, nmort
    d = 8
    w = {}
    w["k_traces"] = dict(bytes=K * (5 * npt + 6 * 5 * nf) * d, flops=K * 6 * 5 * nf * 2 * ncg)
    # 192 B per face point: Q_L, Q_R, n, norm in; flux and averaged state out
    w["k_face_inv"] = dict(bytes=Mm * nf * (10 + 4 + 10) * d, flops=Mm * nf * 120)
    # Bytes: read Q, 1/J and the 9 gradient metrics at Gauss points, the 3
    # staggered metric rows of each grid, the face averages of the 6 faces
    # and the face norms; write FV and Qtp.
    w["k_volume"] = dict(bytes=K * (5 * npt + npt + 9 * npt + 3 * 3 * ns + 6 * 5 * nf + 6 * 5 * nf + 5 * npt) * d
                         + Mm * nf * d,
                         flops=K * 180 * nc * ncg ** 3,
                         pointwise_flops=K * 3 * ns * 260)
    # 128 B per face point: both viscous traces and the norm in, the average out.
    # The kernel also folds in send_flux_to_faces_add: it reads the Roe flux
    # and writes FI - average/Re in place of the average (fused: 168 B).
    w["k_face_visc"] = dict(bytes=Mm * nf * (10 + 1 + 5) * d, flops=Mm * nf * 5 * 6,
                            fused_bytes=Mm * nf * (10 + 1 + 5 + 5) * d)
    # 40 B per DOF: Qtp, G, JQ in; G, JQ out.  The fused kernel also reads the
    # six faces' total fluxes (face part of the divergence) and 1/J, and
    # writes the next stage's face traces; G is not read by the first stage
    # (one launch in five of an RK step).
    w["k_update"] = dict(bytes=K * npt * 5 * (1 + 2 + 2) * d, flops=K * npt * 5 * 16,
                         fused_bytes=K * (npt * 5 * (1 + 0.8 + 1 + 2) + npt + 2 * 6 * 5 * nf) * d)
    return w


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, JQ, args) -> dict:
    """Oracle RHS on the host cores over a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nproc = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = nproc
    cores = max(1, cores)
    g = O.Geometry(mesh)
    ph = O.phys(re=1600.0, pr=0.71)
    O.rhs(g, JQ, ph, nthreads=cores)              # first touch
    n_eval = 0
    t0 = time.perf_counter()
    while True:
        O.rhs(g, JQ, ph, nthreads=cores)
        n_eval += 1
        if time.perf_counter() - t0 > args.cpu_seconds or n_eval >= 8:
            break
    dt = time.perf_counter() - t0
    return {"value": round(mesh.dof * n_eval / dt / 1e9, 5), "unit": "GDOF/s per RHS", "cores": cores,
            "kind": "port", "nproc": nproc, "cpu_model": cpu_model(),
            "sample": f"{n_eval} full RHS evaluations of the same 22^3 p=8 state by the FP64 C oracle "
                      f"(OpenMP, {cores} threads), {dt:.1f} s"}


def mixed_modes(mesh, JQ, s64, dt, args, ms64) -> dict:
    """Throughput of the FP32 / BF16x3 contraction modes on the same input,
    their per-field relative L2 error of one RHS against the FP64 path, and
    of the state after ``args.mixed_steps`` steps from the clean TGV field."""
    from dsem_amd.mixed_study import field_errors
    from dsem_amd.solver import DSEM
    s64.set_state(JQ)
    r64 = s64.rhs()
    xg, jac = s64.gauss_points()
    clean = cases.tgv_state(xg, jac, "reference")
    s64.set_state(clean)
    s64.steps(args.mixed_steps, dt, check=False)
    q64 = s64.get_state()
    res = {}
    for mode in ("fp32", "bf16x3"):
        with DSEM(mesh, re=1600.0, pr=0.71, mach=1.0, precision=mode, metrics=args.metrics) as sm:
            sm.set_state(JQ)
            err = field_errors(sm.rhs(), r64)
            sm.set_state(clean)
            sm.steps(args.mixed_steps, dt, check=False)
            errs = field_errors(sm.get_state(), q64)
            sm.set_state(JQ)
            sm.steps(2, dt, check=False)
            sm.synchronize()
            t0 = time.perf_counter()
            sm.steps(args.steps, dt, check=False)
            sm.synchronize()
            el = time.perf_counter() - t0
        ms = el * 1e3 / args.steps
        res[mode] = {"value": round(mesh.dof * 5 * args.steps / el / 1e9, 4), "unit": "GDOF/s per RHS",
                     "ms_per_step": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3),
                     "rel_l2_max_over_fields": max(v["rel_l2"] for v in err.values()),
                     "rel_l2": {k: v["rel_l2"] for k, v in err.items()},
                     "rel_linf": {k: v["rel_linf"] for k, v in err.items()},
                     f"state_after_{args.mixed_steps}_steps": {
                         "rel_l2": {k: v["rel_l2"] for k, v in errs.items()},
                         "rel_linf": {k: v["rel_linf"] for k, v in errs.items()},
                         "input": "clean TGV field (reference IC), same dt as the timed run"}}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--elements", type=int, default=22, help="elements per direction")
    ap.add_argument("--p", type=int, default=8)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--metrics", default="per_point", choices=["auto", "per_point", "per_element"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32 / BF16x3 contraction modes")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost/gradient checks")
    ap.add_argument("--mixed-steps", type=int, default=1000, help="steps of the mixed-precision state comparison")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one device-to-device copy kernel of this many bytes first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1 (sharding across GPUs is "
                                   "designed, not built)"}))
        sys.exit(2)

    if args.copy_bytes > 0:
        # a plain copy kernel of known size, so that profiler byte counters
        # (FETCH_SIZE / WRITE_SIZE) can be scaled against it
        import torch
        n = args.copy_bytes // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    from dsem_amd.solver import DSEM
    mesh = M.periodic_box(args.elements, args.p)
    s = DSEM(mesh, re=1600.0, pr=0.71, mach=1.0, precision=args.precision, metrics=args.metrics)
    xg, jac = s.gauss_points()
    JQ = cases.tgv_state(xg, jac, "reference", perturb=1e-3, seed=1234)
    s.set_state(JQ)
    dt = s.dtcalc(cfl=0.5)

    for _ in range(args.warmup):
        s.steps(1, dt, check=False)
    s.synchronize()
    t0 = time.perf_counter()
    s.steps(args.steps, dt, check=False)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    resid = s.tstep(0.0, dt)          # untimed: NaN/Inf check of the run
    if not math.isfinite(resid):
        raise RuntimeError("non-finite residual")

    dof = mesh.dof
    ms_step = elapsed * 1e3 / args.steps
    value = dof * 5 * args.steps / elapsed / 1e9

    out = {"metric": "GDOF/s per RHS (FP64 vs mixed @ rel L2<=1e-6); adjoint/forward time ratio",
           "value": round(value, 4), "unit": "GDOF/s per RHS", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else args.precision,
           "data": "synthetic (TGV initial field with seeded 1e-3 relative perturbation, rng 1234)",
           "config": {"workload": f"Taylor-Green vortex, {args.elements}^3 hex8 elements, p={args.p}, "
                                  f"FP64 state, RK4(5) 2N step = 5 RHS", "elements": mesh.K, "p": args.p,
                      "dof": dof, "mortars": mesh.nmortar, "precision": args.precision, "metrics": args.metrics,
                      "gdof_per_s_per_rk_step": round(dof * args.steps / elapsed / 1e9, 4)}}

    if not args.no_roofline:
        s.timing(True)
        s.steps(1, dt, check=False)
        rep = s.timing_report()
        s.timing(False)
        work = kernel_work(mesh.K, mesh.nmortar, args.p)
        per = {}
        for name, (tot_ms, n) in rep.items():
            if n == 0 or name not in work:
                continue
            avg = tot_ms / n
            wk = work[name]
            per[name] = {"avg_ms": round(avg, 5), "launches": n,
                         "bound": "mfma" if name == "k_volume" else "hbm",
                         "algorithmic_GB": round(wk["bytes"] / 1e9, 4),
                         "GB/s": round(wk["bytes"] / (avg * 1e-3) / 1e9, 1),
                         "frac_of_hbm_peak": round(wk["bytes"] / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "TFLOP/s": round(wk["flops"] / (avg * 1e-3) / 1e12, 2)}
            if "pointwise_flops" in wk:
                per[name]["pointwise_TFLOP/s_estimate"] = round(wk["pointwise_flops"] / (avg * 1e-3) / 1e12, 2)
            if "fused_bytes" in wk:
                # every array the fused kernel must move once (the per-unit
                # figure above counts the unfused step it replaces)
                per[name]["fused_GB"] = round(wk["fused_bytes"] / 1e9, 4)
                per[name]["fused_frac_of_hbm_peak"] = round(
                    wk["fused_bytes"] / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        dom = max(per, key=lambda k: per[k]["avg_ms"])
        wk = work[dom]
        avg_s = per[dom]["avg_ms"] * 1e-3
        if dom == "k_volume":
            ach = wk["flops"] / avg_s / 1e12
            out["roofline"] = {"bound": "mfma", "kernel": dom, "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                               "note": "FP64 peak (matrix and vector); achieved = contraction flops per launch "
                                       "(180 nc ncg^3 per element) / HIP-event kernel time; pointwise flux "
                                       "flops are not counted (see kernels.k_volume); traffic = measured HBM "
                                       "bytes per launch"}
        else:
            ach = wk["bytes"] / avg_s / 1e9
            out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None}
        meas = MEASURED_TRAFFIC.get((args.elements, args.p, args.precision, args.metrics))
        if meas and dom in meas:
            out["roofline"]["traffic"] = meas[dom]
        for name in per:
            if meas and name in meas:
                per[name]["hbm_GB_measured"] = round(meas[name] / 1e9, 3)
                per[name]["hbm_GB/s_measured"] = round(meas[name] / (per[name]["avg_ms"] * 1e-3) / 1e9, 1)
        out["kernels"] = per

    if not args.no_mixed and args.precision == "fp64":
        out["mixed"] = mixed_modes(mesh, JQ, s, dt, args, ms_step)
    s.close()
    if not args.no_adjoint:
        from dsem_amd import adjoint_study
        r = adjoint_study.run(n=8, p=6, steps=50, every=10, ndir=4)
        out["adjoint"] = {"adjoint_over_forward": round(r["adjoint_over_forward"], 3),
                          "forward_ms": round(r["forward_ms"], 2), "backward_ms": round(r["backward_ms"], 2),
                          "grad_fd_max_rel_err": r["velocity_directions"]["max_rel_err"],
                          "nu_fd_rel_err": r["nu"]["rel_err"],
                          "config": "TGV 8^3 hex8, p=6, Re=1600, 50 steps, checkpoint every 10, "
                                    "J = int 1/2 rho|u|^2 dV; dJ/du0 on 4 seeded directions and dJ/dnu vs "
                                    "central differences"}
        # 10 steps at 22^3 p=8: a step-tape slot is 4.1 GB, so the whole run
        # fits on the tape (41 GB of the 288 GB) and the forward pass records
        # it: the reversal recomputes nothing
        r = adjoint_study.run(n=args.elements, p=args.p, steps=10, every=10, ndir=2)
        out["adjoint_production_size"] = {
            "adjoint_over_forward": round(r["adjoint_over_forward"], 3), "forward_ms": round(r["forward_ms"], 2),
            "backward_ms": round(r["backward_ms"], 2), "grad_fd_max_rel_err": r["velocity_directions"]["max_rel_err"],
            "nu_fd_rel_err": r["nu"]["rel_err"],
            "config": f"TGV {args.elements}^3 hex8, p={args.p}, Re=1600, 10 steps, all 10 on the step tape "
                      "(recorded by the forward pass), J = int 1/2 rho|u|^2 dV; dJ/du0 on 2 seeded directions and "
                      "dJ/dnu vs central differences"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, JQ, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
