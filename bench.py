"""Benchmark of the block-coupled implicit step on one MI355X.

Workload: the transonic swept wing (semi-span 1, 30 deg sweep, taper 0.56,
10 % analytic section) on the 2 027 520-cell C-H hex mesh at Mach 0.84,
AoA 3.06 deg, AUSM+up, Roe-Jacobian implicit step with local time stepping
(CFL ramp to 20), GMRES(30) + block-ILU(0), characteristic far field.

A "step" is one Single::iterate (face flux + residual, Jacobian assembly,
linear solve, primitive/BC update, CFL/LTS) with every array resident in
HBM.  The line reports the face-flux kernel throughput (Gfaces/s) as the
value, the per-step time of the FP64 and mixed-precision paths, the roofline
of the face-flux kernel (algorithmic bytes / HIP-event kernel time), the time
to a 1e-8 drop of |R|_2 of the mixed and the FP64 path with the exact-Jacobian
(Newton) operator (Roe-ILU(0) preconditioner, CFL ramp from 2 to 1e4, GMRES(30)
to relTol 1e-2 in at most 4 cycles per step), and the CPU oracle timed on a
bounded sample of the same workload.  --turbulence SARC2 runs the viscous Spalart-Allmaras form of the
step (throughput only).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--small] [--skip-converge]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def k1_bytes(n_cells: int, n_int: int, n_bnd: int) -> float:
    """Algorithmic bytes of one face-flux + residual call, first-order
    inviscid: every array read once, R written once; per internal face the
    two int32 cell indices and Sf, per boundary face the cell index, Sf and
    the boundary state, per cell W, 1/V and R."""
    return n_int * (8 + 24) + n_bnd * (4 + 24 + 40) + n_cells * (40 + 8 + 40)


def k1_bytes_viscous(n_cells: int, n_int: int, n_bnd: int) -> float:
    """The same for the viscous face loop (first order, laminar or SA): per
    internal face also the weight, deltaCoeff and the non-orthogonal
    correction vector; per boundary face also muEff_b and the corrected
    boundary gradient of U; per cell also grad U and muEff."""
    return n_int * (8 + 24 + 8 + 8 + 24) + n_bnd * (4 + 24 + 40 + 16 + 72) + n_cells * (40 + 8 + 72 + 16 + 40)


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(kernel: str):
    """HBM bytes per launch of the face-flux kernel on the default workload,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that
    kernel (the commands that produce it are in the file); None when it is
    absent."""
    try:
        with open(os.path.join(PROFILES, f"pmc_{kernel}_swept_wing_2M.md")) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_path(spec, steps: int, warmup: int, mesh=None):
    from rhodst_amd import DeviceCase
    dc = DeviceCase(spec, mesh)
    if warmup:
        dc.iterate(warmup)
    t0 = time.perf_counter()
    res, perfs = dc.iterate(steps, with_perf=True)
    t1 = time.perf_counter()
    return dc, (t1 - t0) * 1e3 / steps, res


def wing_coefficients(dc, mesh, mach=0.84, aoa_deg=3.06):
    """Cd, Cl of the wing patch (forceCoeffs: pRef = pInf, q = rhoInf |UInf|^2 / 2,
    Aref = the semi-span planform area 0.78)."""
    gamma, R, Tinf, pinf = 1.4, 287.0, 288.15, 101325.0
    a = np.deg2rad(aoa_deg)
    U = mach * np.sqrt(gamma * R * Tinf)
    rho = pinf / (R * Tinf)
    out = dc.forces(mesh.patch_index("wing"), (np.cos(a), np.sin(a), 0.0), (-np.sin(a), np.cos(a), 0.0), pinf, rho,
                    U, 0.5 * (1.0 + 0.56) * 1.0)
    return float(out[0]), float(out[1])


def converge_path(name, spec, mesh, max_steps, dmesh=None, chunk=10):
    """Iterate one path until |R|_2 / |R|_2(0) <= 1e-8 (timeDST.C:328-331 norm);
    the time is taken at the end of the chunk of steps that reaches it."""
    from rhodst_amd import DeviceCase
    dc = DeviceCase(spec, dmesh)
    r0 = None
    hist = []
    t0 = time.perf_counter()
    steps = 0
    reached = False
    next_log = 50
    while steps < max_steps:
        res = dc.iterate(chunk)
        if r0 is None:
            r0 = res[0]
        steps += chunk
        hist.extend(res.tolist())
        drop = res[-1] / r0
        if steps >= next_log or not np.isfinite(drop):
            log(f"{name}: step {steps} drop {drop:.3e} t {time.perf_counter() - t0:.1f}s")
            next_log += 50
        if not np.isfinite(drop):
            break
        if np.min(np.array(hist) / r0) <= 1e-8:
            reached = True
            break
    t1 = time.perf_counter()
    h = np.array(hist) / r0
    k = int(np.argmax(h <= 1e-8)) + 1 if reached else None
    cd, cl = wing_coefficients(dc, mesh)
    out = {"reached": reached, "steps": k if reached else steps, "seconds": round(t1 - t0, 2),
           "final_drop": float(h[-1]), "min_drop": float(h.min()), "Cd": cd, "Cl": cl}
    log(f"{name}: {'reached 1e-8 at step ' + str(k) if reached else 'not reached'} in {t1 - t0:.1f}s, Cd {cd:.12f}")
    return out, dc.mesh


def converge(args, mesh, spec64, spec32) -> int:
    """Time to a 1e-8 drop of |R|_2 for the mixed and the FP64 path; prints
    one JSON line."""
    out = {}
    dmesh = None
    specs = {"mixed": spec32, "fp64": spec64}
    for name in args.paths.split(","):
        out[name], dmesh = converge_path(name, specs[name], mesh, args.max_steps, dmesh)
    print(json.dumps({"converge": out, "cells": mesh.n_cells}), flush=True)
    return 0


def flux_only(args, spec) -> int:
    """A 256 MiB device-to-device copy (the counter calibration: its
    FETCH_SIZE and WRITE_SIZE are 256 MiB each) and args.flux_calls face-flux
    calls on the workload."""
    import torch
    from rhodst_amd import DeviceCase
    dc = DeviceCase(spec)
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    ms = dc.time_flux(args.flux_calls)
    print(json.dumps({"flux_ms": ms, "copy_bytes": a.numel() * 8, "calls": args.flux_calls}), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--small", action="store_true", help="a 0.26 M-cell mesh instead of 2.03 M")
    ap.add_argument("--flux-calls", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--converge", action="store_true",
                    help="run the mixed and the FP64 paths (partitioned ILU, classical GS) to a 1e-8 drop of |R|_2")
    ap.add_argument("--max-steps", type=int, default=4000)
    ap.add_argument("--max-co", type=float, default=None, help="CFL ramp ceiling (default: the workload's 20)")
    ap.add_argument("--gmres-cycles", type=int, default=None, help="GMRES restart cycles per step (default 1)")
    ap.add_argument("--krylov-dirs", type=int, default=None, help="GMRES restart length (default 30)")
    ap.add_argument("--turbulence", default=None, help="SARC2 or SAneg: the viscous SA form of the workload")
    ap.add_argument("--jacobian", default="Roe", choices=["Roe", "exact"],
                    help="implicit operator: the reference's Roe Jacobian, or the exact linearisation of the "
                         "first-order inviscid residual (Newton)")
    ap.add_argument("--rel-tol", type=float, default=None, help="GMRES relTol (default: the workload's 1e-2)")
    ap.add_argument("--initial-co", type=float, default=None, help="initial CFL of the ramp (default 1)")
    ap.add_argument("--skip-converge", action="store_true",
                    help="leave out the time-to-1e-8 runs of the default mode")
    ap.add_argument("--converge-steps", type=int, default=1200, help="step limit of the time-to-1e-8 runs")
    ap.add_argument("--paths", default="mixed,fp64", help="converge mode: paths to run, in order")
    ap.add_argument("--flux-only", action="store_true",
                    help="only the face-flux kernel (n calls) and a 256 MiB device copy of known size: the "
                         "short program the FETCH_SIZE / WRITE_SIZE counter runs profile")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: every path here runs on a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    from rhodst_amd import cases, meshgen
    mesh_kw = dict(n_wake=24, n_surf=64, n_normal=64, n_span_wing=24, n_span_out=8) if args.small else {}
    t = time.perf_counter()
    mesh = meshgen.swept_wing(**mesh_kw)
    log(f"mesh {mesh.n_cells} cells, {mesh.n_internal} internal faces in {time.perf_counter() - t:.1f}s")
    extra = {}
    if args.max_co is not None:
        extra["max_co"] = args.max_co
    if args.gmres_cycles is not None:
        extra["max_iter"] = args.gmres_cycles
    if args.krylov_dirs is not None:
        extra["n_directions"] = args.krylov_dirs
    if args.rel_tol is not None:
        extra["rel_tol"] = args.rel_tol
    if args.initial_co is not None:
        extra["initial_co"] = args.initial_co
    if args.jacobian != "Roe":
        extra["jacobian"] = args.jacobian
    wd = None
    mesh_dev = None
    if args.turbulence:
        from rhodst_amd._capi import DeviceMesh
        t = time.perf_counter()
        mesh_dev = DeviceMesh(mesh)
        wd = mesh_dev.wall_distance()
        extra.update(turbulence=args.turbulence, wall_dist=wd)
        log(f"mesh upload + device wall distance in {time.perf_counter() - t:.1f}s")
    spec64 = cases.swept_wing_case(mesh=mesh, **extra)
    fast = dict(ilu_order="partitioned", gram_schmidt="classical", **extra)
    spec64p = cases.swept_wing_case(mesh=mesh, **fast)
    spec32p = cases.swept_wing_case(mesh=mesh, precision="mixed", **fast)
    if args.converge:
        return converge(args, mesh, spec64p, spec32p)
    if args.flux_only:
        return flux_only(args, spec64)

    log("fp64 path, natural-order ILU(0) (the plain path)")
    dc64, ms64, res64 = run_path(spec64, args.steps, args.warmup)
    log(f"fp64 {ms64:.2f} ms/step, |R| {res64[0]:.4e} -> {res64[-1]:.4e}")
    info = dc64.mesh.info
    nC, nI, nB = info["n_cells"], info["n_internal"], info["n_bfaces"]
    flux_ms = dc64.time_flux(args.flux_calls)
    mesh_dev = dc64.mesh
    del dc64
    log(f"face-flux kernel {flux_ms:.4f} ms")
    log("fp64 path, partitioned block-ILU(0)")
    dc64p, ms64p, res64p = run_path(spec64p, args.steps, args.warmup, mesh_dev)
    log(f"fp64 partitioned {ms64p:.2f} ms/step, |R| {res64p[0]:.4e} -> {res64p[-1]:.4e}")
    del dc64p
    log("mixed path, partitioned block-ILU(0)")
    dc32, ms32, res32 = run_path(spec32p, args.steps, args.warmup, mesh_dev)
    log(f"mixed {ms32:.2f} ms/step, |R| {res32[0]:.4e} -> {res32[-1]:.4e}")
    del dc32

    gfaces = (nI + nB) / (flux_ms * 1e-3) / 1e9
    bytes_k1 = (k1_bytes_viscous if args.turbulence else k1_bytes)(nC, nI, nB)
    achieved = bytes_k1 / (flux_ms * 1e-3) / 1e9
    kname = "k_flux_residual"
    traffic = measured_traffic(kname) if not (args.small or args.turbulence) else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "%s<AUSM+up%s>" % (kname, ", viscous" if args.turbulence else ""),
                "kernel_ms": round(flux_ms, 5),
                "algorithmic_bytes": int(bytes_k1)}

    ttc = None
    if not (args.skip_converge or args.small or args.turbulence):
        # time to a 1e-8 drop: the exact-Jacobian (Newton) operator with the
        # Roe ILU(0) preconditioner, timeDST's CFL ramp started at 2 (initialCo) up to
        # 1e4, GMRES(30) to relTol 1e-2 in at most 4 cycles, mixed and FP64
        newton = dict(jacobian="exact", initial_co=2.0, max_co=1e4, rel_tol=1e-2, max_iter=4,
                      ilu_order="partitioned", gram_schmidt="classical")
        ttc = {}
        for name, prec in (("mixed", "mixed"), ("fp64", "fp64")):
            log(f"time to 1e-8, {name} path, exact Jacobian")
            ttc[name], _ = converge_path(name, cases.swept_wing_case(mesh=mesh, precision=prec, **newton), mesh,
                                         args.converge_steps, mesh_dev)
        ttc["settings"] = "exact Jacobian + Roe-ILU(0) preconditioner, CFL ramp 2 -> 1e4, GMRES(30) relTol 1e-2, 4 cycles"
        if ttc["mixed"]["reached"] and ttc["fp64"]["reached"]:
            ttc["speedup_mixed_vs_fp64"] = round(ttc["fp64"]["seconds"] / ttc["mixed"]["seconds"], 3)
        ttc["abs_dCd_mixed_vs_fp64"] = abs(ttc["mixed"]["Cd"] - ttc["fp64"]["Cd"])

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 thread)")
        from oracle import binding
        oc = binding.OracleCase(spec64)
        n_calls, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n_calls < 50:
            t0 = time.perf_counter()
            oc.flux_residual()
            t_cpu += time.perf_counter() - t0
            n_calls += 1
        cpu = {"value": round((nI + nB) * n_calls / t_cpu / 1e9, 5), "unit": "Gfaces/s", "cores": 1,
               "kind": "port",
               "sample": f"{n_calls} face-flux + residual calls of the CPU FP64 oracle on the same "
                         f"{nC}-cell mesh and uniform free-stream state ({t_cpu:.1f} s)"}

    out = {
        "metric": "time to 1e-8 residual drop + Gfaces/s flux, 2M-cell 3D SA, mixed vs FP64",
        "value": round(gfaces, 3),
        "unit": "Gfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "ms_per_step_fp64": round(ms64, 3),
        "ms_per_step_fp64_partitioned": round(ms64p, 3),
        "ms_per_step_mixed": round(ms32, 3),
        "mixed_speedup_vs_plain_fp64": round(ms64 / ms32, 3),
        "mixed_speedup_same_preconditioner": round(ms64p / ms32, 3),
        "residual_drop_fp64": float(res64[-1] / res64[0]),
        "residual_drop_fp64_partitioned": float(res64p[-1] / res64p[0]),
        "residual_drop_mixed": float(res32[-1] / res32[0]),
        "time_to_1e-8_s": (ttc["mixed"]["seconds"] if ttc and ttc["mixed"]["reached"] else None),
        "time_to_1e-8": ttc,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"swept wing C-H mesh {nC} cells ({nI} internal + {nB} boundary faces), "
                               "M0.84 AoA3.06, AUSM+up first order, "
                               + (f"{args.turbulence} at Re 1.2e7, no-slip wing" if args.turbulence
                                  else "inviscid slip wall")
                               + ", LTS CFL ramp to 20, GMRES(30)+block-ILU(0), 1 restart cycle per step; "
                                 "ms_per_step is the mixed-precision path with the partitioned block-ILU(0) and "
                                 "classical Gram-Schmidt",
                   "cells": nC, "faces": nI + nB, "flux_scheme": "AUSM+up"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
