"""Benchmark of the block-coupled implicit step on one MI355X.

Workload (default): the transonic swept wing (semi-span 1, 30 deg sweep,
taper 0.56, 10 % analytic section) on the 2 027 520-cell C-H hex mesh with
32 wall layers from a first cell of 2e-6 chord (y+ ~ 1), Mach 0.84,
AoA 3.06 deg, Re 1.2e7, no-slip wing with the Spalart-Allmaras model with
rotation-curvature correction (SARC2), viscous terms on, AUSM+up first
order, the Roe-Jacobian implicit step with local time stepping (CFL ramp to
20), GMRES(30) + block-ILU(0), characteristic far field.

A "step" is one Single::iterate (face flux + residual, Jacobian assembly,
linear solve, primitive/BC update, nuTilda equation, CFL/LTS) with every
array resident in HBM.  The line reports:
  value           Gfaces/s of the face-flux + residual kernel of this
                  workload (viscous, muEff with the eddy viscosity), HIP
                  events over back-to-back launches;
  ms_per_step_*   the plain FP64 path (natural-order ILU, modified
                  Gram-Schmidt), FP64 with the partitioned ILU, and mixed
                  precision (FP32 Jacobian/ILU/Krylov in FP64 defect
                  correction, FP64 residual);
  roofline        of the face-flux kernel: algorithmic bytes / kernel time
                  against 8 TB/s, the measured HBM traffic and the FP64
                  vector-ALU rate against 78.6 TFLOP/s from the committed
                  rocprofv3 counter summaries;
  time_to_1e-8    wall-clock to |R|_2/|R|_2(0) <= 1e-8 (timeDST.C:328-331)
                  with the reference's Roe operator at CFL 20, mixed and
                  FP64, over a bounded number of steps (--roe-steps): on this
                  mesh the drop falls below 1e-5 in a few thousand steps and
                  then stalls, so 1e-8 is not reached and the line reports the
                  time each path takes to the levels both reach (1e-3, 1e-4,
                  ...), their ratio, the drop reached and Cd of both; with
                  --newton also time to 1e-8 of the exact-Jacobian
                  Newton-Krylov operator (an extension: exact first-order
                  linearisation, Roe ILU(0) preconditioner, CFL ramp 2 ->
                  1e4, each cell's rho/rhoE update limited to 20 %) on the
                  inviscid form of the same wing, mixed and FP64, which it
                  converges (the SA form stalls, as above);
  cpu_baseline    the CPU FP64 oracle's face flux + residual on the same
                  mesh and state, all host threads (OpenMP) and one.

Other modes, one JSON line each: --workloads (the forward step to 1e-8 and the
hybrid airfoil per step, GPU and CPU oracle), --adjoint (the discrete adjoint
on the viscous hybrid airfoil: exact viscous J, the adjoint solve, J v against
central differences, the 400-wall-node gradient), --converge-only (the Newton
time to 1e-8), --flux-only (the face-flux kernel and a copy of known size for
the counter runs), --tetprism (8.32 M tetrahedra and prisms, Kurganov + SAneg,
mixed and FP64 per step).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--small]
       [--turbulence SARC2|SAneg|none] [--skip-converge] [--no-cpu]
       [--workloads | --adjoint | --tetprism | --converge-only | --flux-only]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6
PROFILES = os.path.join(ROOT, "profiles")


def k1_bytes(n_cells: int, n_int: int, n_bnd: int) -> float:
    """Algorithmic bytes of one face-flux + residual launch, first-order
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


def profile_value(name: str, key: str):
    """A number a committed rocprofv3 summary under profiles/ states as
    'key: value' (the commands that produce it are in the file)."""
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def device_copy(nbytes: int) -> int:
    """One device-to-device copy of nbytes through the HIP runtime (a blit
    kernel of known traffic: nbytes read, nbytes written), the calibration
    of the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return nbytes


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_path(spec, steps: int, warmup: int, mesh=None):
    from rhodst_amd import DeviceCase
    dc = DeviceCase(spec, mesh)
    if warmup:
        dc.iterate(warmup)
    t0 = time.perf_counter()
    res = dc.iterate(steps)
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


def converge_path(name, spec, mesh, max_steps, dmesh=None, chunk=20):
    """Iterate one path until |R|_2 / |R|_2(0) <= 1e-8 (timeDST.C:328-331 norm)
    or max_steps; the time is taken at the end of the chunk that reaches it."""
    from rhodst_amd import DeviceCase
    dc = DeviceCase(spec, dmesh)
    r0 = None
    hist = []
    t0 = time.perf_counter()
    steps = 0
    reached = False
    next_log = 100
    levels = (1e-3, 1e-4, 1e-5, 1e-6)
    at = {}  # level -> (first step at or below it, seconds at the end of that chunk)
    while steps < max_steps:
        res = dc.iterate(min(chunk, max_steps - steps))
        if r0 is None:
            r0 = res[0]
        steps += len(res)
        hist.extend(res.tolist())
        drop = res[-1] / r0
        if steps >= next_log or not np.isfinite(drop):
            log(f"{name}: step {steps} drop {drop:.3e} t {time.perf_counter() - t0:.1f}s")
            next_log += 100
        if not np.isfinite(drop):
            break
        if np.min(np.array(hist) / r0) <= 1e-8:
            reached = True
            break
        for lvl in levels:
            if lvl not in at and np.min(np.array(hist) / r0) <= lvl:
                at[lvl] = (int(np.argmax(np.array(hist) / r0 <= lvl)) + 1, round(time.perf_counter() - t0, 2))
    t1 = time.perf_counter()
    h = np.array(hist) / r0
    k = int(np.argmax(h <= 1e-8)) + 1 if reached else None
    cd, cl = wing_coefficients(dc, mesh)
    out = {"reached": reached, "steps": k if reached else steps, "seconds": round(t1 - t0, 2),
           "final_drop": float(h[-1]), "min_drop": float(h.min()), "Cd": cd, "Cl": cl,
           "first_reached": {f"{lvl:g}": {"step": a[0], "seconds": a[1]} for lvl, a in sorted(at.items(), reverse=True)}}
    log(f"{name}: {'reached 1e-8 at step ' + str(k) if reached else 'not reached'} in {t1 - t0:.1f}s, Cd {cd:.12f}")
    return out


def workloads_mode(args) -> int:
    """The forward step (120 x 40, AUSM+, CFL 3, GMRES(20) + ILU(0)) run to a
    1e-8 drop, and the hybrid airfoil (HLLC + SAneg) per step, on the GPU
    (plain FP64 path and mixed precision) and on the CPU oracle."""
    from rhodst_amd import DeviceCase, cases, meshgen
    from rhodst_amd._capi import DeviceMesh
    out = {"forward_step": {}, "hybrid_airfoil": {}}
    for name, extra in (("fp64_plain", {}), ("mixed", dict(precision="mixed", ilu_order="partitioned",
                                                             gram_schmidt="classical"))):
        spec = cases.forward_step_case(120, 40, initial_co=3.0, min_co=3.0, max_co=3.0, **extra)
        dc = DeviceCase(spec)
        hist, t0, at = [], time.perf_counter(), None
        while len(hist) < 20000:
            hist.extend(dc.iterate(100).tolist())
            h = np.array(hist) / hist[0]
            if h.min() <= 1e-8 or not np.isfinite(h[-1]):
                at = int(np.argmax(h <= 1e-8)) + 1 if h.min() <= 1e-8 else None
                break
        sec = time.perf_counter() - t0
        out["forward_step"][name] = {"steps": len(hist), "seconds": round(sec, 3), "ms_per_step": round(sec / len(hist) * 1e3, 4),
                                     "step_1e-8": at, "final_drop": float(hist[-1] / hist[0])}
        log(f"forward step {name}: {out['forward_step'][name]}")
    mesh = meshgen.hybrid_airfoil()
    dm = DeviceMesh(mesh)
    wd = dm.wall_distance()
    for name, extra in (("fp64_plain", {}), ("mixed", dict(precision="mixed", ilu_order="partitioned",
                                                             gram_schmidt="classical"))):
        dc = DeviceCase(cases.airfoil_case(mesh=mesh, wall_dist=wd, **extra), dm)
        dc.iterate(2)
        t0 = time.perf_counter()
        r = dc.iterate(20)
        sec = time.perf_counter() - t0
        out["hybrid_airfoil"][name] = {"cells": mesh.n_cells, "ms_per_step": round(sec / 20 * 1e3, 3),
                                       "drop_after_22_steps": float(r[-1] / r[0])}
        log(f"hybrid airfoil {name}: {out['hybrid_airfoil'][name]}")
    cpu = {}
    all_threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    for wl, budget in (("forward_step", 200.0), ("hybrid_airfoil", 25.0)):
        for threads in (all_threads, 1):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), ORACLE_OPENMP="1")
            log(f"cpu oracle {wl}, {threads} thread(s)")
            r = subprocess.run([sys.executable, "-m", "oracle.cpu_workloads", "--workload", wl, "--seconds",
                                str(budget)], cwd=ROOT, env=env, stdout=subprocess.PIPE, text=True, timeout=900)
            if r.returncode != 0:
                log(f"cpu {wl} failed (exit {r.returncode})")
                continue
            cpu[f"{wl}_{threads}t"] = json.loads(r.stdout.strip().splitlines()[-1])
    out["cpu_oracle"] = cpu
    print(json.dumps(out), flush=True)
    return 0


def adjoint_mode(args) -> int:
    """The discrete adjoint on the viscous hybrid airfoil (205 k cells, HLLC +
    SAneg, M 0.73, Re 6.5e6) at a shape seeded with Hicks-Henne bumps of 1e-4
    chord: the flow converged by the mixed path, the exact viscous J with the
    eddy viscosity frozen, J^T lambda = dCd/dW solved with the FP32
    transposed ILU inside the FP64 defect correction (and in FP64 for the
    comparison), J v against central differences of the residual (h 1e-7),
    and dCd/d(displacement) of the 400 wall nodes (batched central
    differences of the frozen-state partials, 16 mesh builds), a few nodes
    also displaced alone."""
    from rhodst_amd.adjoint import AirfoilProblem
    out = {"workload": "hybrid airfoil 205191 cells, HLLC + SAneg (frozen eddy viscosity in the adjoint), "
                       "400 wall nodes, seeded Hicks-Henne bumps 1e-4 chord"}
    heat = 0 if args.adjoint_no_heat else 1
    prob = AirfoilProblem(precision="mixed", ilu_order="partitioned", gram_schmidt="classical",
                          heat_conduction=heat)
    out["heat_conduction"] = heat
    t0 = time.perf_counter()
    cd, steps, hist = prob.solve(max_steps=args.adjoint_steps, drop=1e-8, log=log)
    out["primal"] = {"steps": steps, "seconds": round(time.perf_counter() - t0, 2), "drop": hist[-1][1], "Cd": cd}
    dc = prob.dc
    T = dc.get(4)
    out["primal"]["T_max"] = float(T.max())
    out["primal"]["cells_at_T_bound"] = int(np.sum(T >= 2000.0 * (1 - 1e-12)))
    log(f"primal: {out['primal']}")
    patch = dc.spec.mesh.patch_index("airfoil")
    t = time.perf_counter()
    dc.adjoint_linearise()
    m = dc.mesh
    out["linearise_s"] = round(time.perf_counter() - t, 3)
    g = dc.adjoint_objective_gradient(patch, prob.drag, prob.q)
    sol = {}
    # the defect correction (FP64 residual on the exact viscous J^T, FP32
    # inner cycles on the shifted face-neighbour part, FP32 ILU), and for the
    # record plain FP64 GMRES on the exact J^T with the same ILU
    runs = (("defect_correction_mixed", {}),
            ("gmres_fp64", dict(precision="fp64", inner_operator=0, max_cycles=400)))
    for name, kw in runs:
        t = time.perf_counter()
        lam, perf = prob.adjoint(**kw)
        sec = time.perf_counter() - t
        r = np.linalg.norm(dc.adjoint_apply(lam, transpose=True) - g) / np.linalg.norm(g)
        sol[name] = lam
        out[f"adjoint_{name}"] = {"seconds": round(sec, 3), "cycles": int(perf.n_iterations),
                                  "residual_rel": float(r)}
        log(f"adjoint {name}: {out[f'adjoint_{name}']}")
    out["adjoint_precond_shift"] = prob.precond_shift
    out["adjoint_converged"] = bool(out["adjoint_defect_correction_mixed"]["residual_rel"] < 1e-6)
    # J v against central differences of the residual; rows whose error
    # exceeds 1e-5 of the component's scale are counted (faces on a kink of
    # the flux -- HLLC's wave-speed switches at the near-wall contact speed
    # ~0 -- inside the step)
    W = dc.get(9).reshape(-1, 5)
    rng = np.random.default_rng(5)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    out["Jv_vs_fd"] = {}
    for h in (1e-7, 1e-8):
        dc.put(9, W + h * v)
        Rp = dc.adjoint_residual()
        dc.put(9, W - h * v)
        Rm = dc.adjoint_residual()
        dc.put(9, W)
        fd = (Rp - Rm) / (2 * h)
        e = np.abs(Jv - fd) / np.abs(fd).max(axis=0)
        out["Jv_vs_fd"][f"{h:g}"] = {"max_rel": e.max(axis=0).tolist(), "median_rel": np.median(e, axis=0).tolist(),
                                     "rows_above_1e-5": int(np.sum(e.max(axis=1) > 1e-5))}
    log(f"J v vs central differences: {out['Jv_vs_fd']}")
    lam = sol["defect_correction_mixed"]
    t = time.perf_counter()
    gn = prob.node_gradient(lam)
    out["node_gradient_s"] = round(time.perf_counter() - t, 2)
    picks = [0, 97, 200, 301, 399]
    gs = prob.node_gradient(lam, nodes=picks)
    out["node_gradient"] = {"n_nodes": len(gn), "max_abs": float(np.abs(gn).max()),
                            "batched_vs_alone_rel": float(np.abs(gn[picks] - gs).max() / np.abs(gn).max()),
                            "sample": {int(k): float(gn[k]) for k in picks}}
    log(f"node gradient: {out['node_gradient']} in {out['node_gradient_s']} s")
    print(json.dumps(out), flush=True)
    return 0


def tetprism_mode(args) -> int:
    """Tetrahedra over prism layers at 8.32 M cells (meshgen.tet_prism_box
    160 x 100 x 100 hexes, 20 prism layers from a first height of 2e-5 over a
    1 x 0.6 x 0.4 box): a no-slip flat wall under the characteristic far
    field, M 0.3, Sutherland air, Kurganov first order, SAneg, viscous terms,
    the Roe operator, GMRES(30) + partitioned block-ILU(0), classical
    Gram-Schmidt: ms per step in mixed precision and in FP64, and the face-flux
    kernel's rate."""
    from rhodst_amd import DeviceCase, cases, meshgen
    from rhodst_amd._capi import DeviceMesh
    t0 = time.perf_counter()
    mesh = meshgen.tet_prism_box(160, 100, 100, prism_layers=20, lx=1.0, ly=0.6, lz=0.4, first=2e-5, jitter=0.1)
    log(f"mesh: {mesh.n_cells} cells, {mesh.n_faces} faces in {time.perf_counter() - t0:.1f} s")
    dm = DeviceMesh(mesh)
    wd = dm.wall_distance()
    nF = mesh.n_internal + mesh.n_boundary_nonempty()
    out = {"workload": f"tetrahedra over prism layers, {mesh.n_cells} cells ({nF} faces), flat no-slip wall, "
                       f"M0.3, Kurganov first order, SAneg, viscous, Roe operator, GMRES(30) + partitioned ILU(0)",
           "cells": mesh.n_cells, "faces": nF}
    for name, prec in (("mixed", "mixed"), ("fp64_partitioned", "fp64")):
        dc = DeviceCase(cases.tet_prism_case(mesh=mesh, wall_dist=wd, precision=prec, ilu_order="partitioned",
                                             gram_schmidt="classical", n_directions=30, max_iter=1, rel_tol=1e-2),
                        dm)
        dc.iterate(args.warmup)
        t = time.perf_counter()
        r = dc.iterate(args.steps)
        sec = time.perf_counter() - t
        out[f"ms_per_step_{name}"] = round(sec / args.steps * 1e3, 3)
        out[f"drop_{name}"] = float(r[-1] / r[0])
        if name == "mixed":
            ms = dc.time_flux(args.flux_calls)
            out["flux_kernel_ms"] = round(ms, 5)
            out["flux_gfaces_per_s"] = round(nF / (ms * 1e-3) / 1e9, 3)
            if args.tetprism_adjoint:
                # the viscous adjoint at this size (eddy viscosity frozen): the
                # exact J on the second-neighbour pattern and the defect correction
                dc.update_fields()
                dc.update_fields()
                patch = mesh.patch_index("bottom")
                t = time.perf_counter()
                dc.adjoint_linearise()
                out["adjoint_linearise_s"] = round(time.perf_counter() - t, 2)
                g = dc.adjoint_objective_gradient(patch, (1.0, 0.0, 0.0), 1.0)
                t = time.perf_counter()
                lam, perf = dc.adjoint_solve(patch, (1.0, 0.0, 0.0), 1.0, precision="mixed", n_directions=30,
                                             max_cycles=args.adjoint_steps, tolerance=0.0, rel_tol=1e-8,
                                             precond_shift=0.3, inner_operator=1)
                res = np.linalg.norm(dc.adjoint_apply(lam, transpose=True) - g) / np.linalg.norm(g)
                out["adjoint_solve"] = {"seconds": round(time.perf_counter() - t, 2), "cycles": int(perf.n_iterations),
                                        "residual_rel": float(res)}
                log(f"adjoint: {out['adjoint_solve']} after a {out['adjoint_linearise_s']} s linearisation")
        log(f"{name}: {out[f'ms_per_step_{name}']} ms/step")
        del dc
    out["mixed_speedup_same_preconditioner"] = round(out["ms_per_step_fp64_partitioned"] / out["ms_per_step_mixed"], 3)
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(args, mesh_kw, case_kw, wall_dist, faces):
    """The oracle's face flux + residual on the host: all threads (OpenMP
    build, bit-identical to the serial one) and one thread, each in a child
    process started before any OpenMP runtime."""
    out = {}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "spec.npz")
        extra = {} if wall_dist is None else {"wall_dist": wall_dist}
        np.savez(path, mesh_kw=np.array(json.dumps(mesh_kw)), case_kw=np.array(json.dumps(case_kw)), **extra)
        all_threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        for threads in (all_threads, 1):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), ORACLE_OPENMP="1")
            log(f"cpu baseline: oracle on {threads} thread(s)")
            r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--spec", path, "--seconds",
                                str(args.cpu_seconds)], cwd=ROOT, env=env, capture_output=True, text=True,
                               timeout=600)
            if r.returncode != 0:
                log(f"cpu baseline failed: {r.stderr[-500:]}")
                return None
            out[threads] = json.loads(r.stdout.strip().splitlines()[-1])
    a, one = out[all_threads], out[1]
    return {"value": round(a["gfaces_per_s"], 5), "unit": "Gfaces/s", "cores": all_threads, "kind": "port",
            "value_1_core": round(one["gfaces_per_s"], 5), "cpu_model": a["cpu_model"], "nproc": a["nproc"],
            "omp_num_threads": all_threads,
            "sample": f"{a['calls']} (all threads) and {one['calls']} (1 thread) face-flux + residual calls of the "
                      f"CPU FP64 oracle (OpenMP build, results bit-identical to the serial checker) on the same "
                      f"{a['cells']}-cell mesh and free-stream state ({a['seconds']:.1f} s / {one['seconds']:.1f} s)"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--small", action="store_true", help="a 0.36 M-cell mesh instead of 2.03 M")
    ap.add_argument("--mesh-order", default="ijk", choices=("ijk", "tiles"),
                    help="cell numbering of the wing mesh (meshgen.swept_wing order)")
    ap.add_argument("--turbulence", default="SARC2", help="SARC2, SARC1, SAneg, or none (inviscid slip wing)")
    ap.add_argument("--flux-calls", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-converge", action="store_true", help="leave out the time-to-1e-8 runs")
    ap.add_argument("--roe-steps", type=int, default=500, help="steps of the Roe operator at CFL 20 per path")
    ap.add_argument("--newton", action="store_true",
                    help="also run the Newton operator (extension) toward 1e-8 on the workload, mixed and FP64")
    ap.add_argument("--converge-steps", type=int, default=2500, help="step limit of the Newton time-to-1e-8 runs")
    ap.add_argument("--paths", default="mixed,fp64", help="time-to-1e-8 paths, in order")
    ap.add_argument("--converge-only", action="store_true", help="only the Newton time-to-1e-8 runs (one JSON line)")
    ap.add_argument("--workloads", action="store_true",
                    help="only the two smaller workloads (forward step to 1e-8, hybrid airfoil per step) on the GPU "
                         "and the CPU oracle (all threads and one): one JSON line")
    ap.add_argument("--adjoint", action="store_true",
                    help="only the discrete adjoint on the viscous hybrid airfoil (400 wall nodes): one JSON line")
    ap.add_argument("--tetprism", action="store_true",
                    help="only the 8.32 M-cell tetrahedron/prism workload (Kurganov + SAneg), mixed and FP64: one JSON line")
    ap.add_argument("--tetprism-adjoint", action="store_true",
                    help="with --tetprism: also the viscous adjoint (J and the defect-correction solve) at that size")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="step limit of the airfoil flow before the adjoint (--adjoint); cycle limit of the "
                         "tetrahedron/prism adjoint solve (--tetprism-adjoint)")
    ap.add_argument("--adjoint-no-heat", action="store_true",
                    help="the airfoil flow of --adjoint without the heat-conduction extension")
    ap.add_argument("--flux-only", action="store_true",
                    help="only the face-flux kernel (n calls) and a 256 MiB device copy of known size: the "
                         "short program the FETCH_SIZE / WRITE_SIZE counter runs profile")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print("bench.py measures one GPU: every path here runs on a single MI355X; "
              "the sharded multi-GPU path is not implemented", file=sys.stderr)
        return 2

    if args.workloads:
        return workloads_mode(args)
    if args.adjoint:
        return adjoint_mode(args)
    if args.tetprism:
        return tetprism_mode(args)
    from rhodst_amd import cases, meshgen
    from rhodst_amd._capi import DeviceMesh
    mesh_kw = dict(n_wake=24, n_surf=64, n_normal=64, n_span_wing=24, n_span_out=8) if args.small else {}
    mesh_kw["order"] = args.mesh_order
    turb = None if args.turbulence.lower() == "none" else args.turbulence
    t = time.perf_counter()
    mesh = meshgen.swept_wing(**mesh_kw)
    log(f"mesh {mesh.n_cells} cells, {mesh.n_internal} internal faces in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    dm = DeviceMesh(mesh)
    wd = dm.wall_distance() if turb else None
    log(f"mesh upload{' + device wall distance' if turb else ''} in {time.perf_counter() - t:.1f}s")
    base = dict(turbulence=turb, wall_dist=wd) if turb else {}
    fast = dict(ilu_order="partitioned", gram_schmidt="classical")
    spec64 = cases.swept_wing_case(mesh=mesh, **base)
    spec64p = cases.swept_wing_case(mesh=mesh, **fast, **base)
    spec32p = cases.swept_wing_case(mesh=mesh, precision="mixed", **fast, **base)
    # the Newton extension runs on the inviscid form of the wing, which it converges
    newton = dict(jacobian="exact", initial_co=2.0, max_co=1e4, rel_tol=1e-2, max_iter=4, update_limit=0.2, **fast)

    if args.flux_only:
        from rhodst_amd import DeviceCase
        dc = DeviceCase(spec64, dm)
        nbytes = device_copy(256 * 1024 * 1024)
        ms = dc.time_flux(args.flux_calls)
        print(json.dumps({"flux_ms": ms, "copy_bytes": nbytes, "calls": args.flux_calls}), flush=True)
        return 0
    if args.converge_only:
        out = {}
        for name in args.paths.split(","):
            prec = "mixed" if name == "mixed" else "fp64"
            out[name] = converge_path(f"newton {name}", cases.swept_wing_case(mesh=mesh, precision=prec, **newton),
                                      mesh, args.converge_steps, dm)
        print(json.dumps({"converge": out, "cells": mesh.n_cells}), flush=True)
        return 0

    log("fp64 path, natural-order ILU(0) (the plain path)")
    dc64, ms64, res64 = run_path(spec64, args.steps, args.warmup, dm)
    log(f"fp64 {ms64:.2f} ms/step, |R| {res64[0]:.4e} -> {res64[-1]:.4e}")
    info = dm.info
    nC, nI, nB = info["n_cells"], info["n_internal"], info["n_bfaces"]
    flux_ms = dc64.time_flux(args.flux_calls)
    del dc64
    log(f"face-flux kernel {flux_ms:.4f} ms")
    log("fp64 path, partitioned block-ILU(0)")
    dc64p, ms64p, res64p = run_path(spec64p, args.steps, args.warmup, dm)
    log(f"fp64 partitioned {ms64p:.2f} ms/step, |R| {res64p[0]:.4e} -> {res64p[-1]:.4e}")
    del dc64p
    log("mixed path, partitioned block-ILU(0)")
    dc32, ms32, res32 = run_path(spec32p, args.steps, args.warmup, dm)
    log(f"mixed {ms32:.2f} ms/step, |R| {res32[0]:.4e} -> {res32[-1]:.4e}")
    del dc32

    gfaces = (nI + nB) / (flux_ms * 1e-3) / 1e9
    viscous = turb is not None
    bytes_k1 = (k1_bytes_viscous if viscous else k1_bytes)(nC, nI, nB)
    achieved = bytes_k1 / (flux_ms * 1e-3) / 1e9
    kname = "k_flux_residual"
    tag = "swept_wing_2M_" + ("sa" if viscous else "inviscid")
    prof = f"pmc_{kname}_{tag}.md"
    traffic = None if args.small else profile_value(prof, "traffic_bytes_per_launch")
    flops = None if args.small else profile_value(prof, "fp64_flop_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic else None,
                "kernel": "%s<AUSM+up%s>" % (kname, ", viscous" if viscous else ""),
                "kernel_ms": round(flux_ms, 5), "algorithmic_bytes": int(bytes_k1)}
    if flops:
        tf = flops / (flux_ms * 1e-3) / 1e12
        roofline["fp64_valu"] = {"achieved": round(tf, 2), "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": round(tf / FP64_VECTOR_PEAK_TFLOPS, 4), "flop_per_launch": int(flops)}

    ttc = None
    if not (args.skip_converge or args.small):
        ttc = {"roe_cfl20": {}, "newton": {}}
        for name, spec in (("mixed", spec32p), ("fp64", spec64p)):
            log(f"Roe operator at CFL 20, {name}, {args.roe_steps} steps")
            ttc["roe_cfl20"][name] = converge_path(f"roe {name}", spec, mesh, args.roe_steps, dm)
        rc = ttc["roe_cfl20"]
        rc["abs_dCd_mixed_vs_fp64"] = abs(rc["mixed"]["Cd"] - rc["fp64"]["Cd"])
        if rc["mixed"]["reached"] and rc["fp64"]["reached"]:
            rc["speedup_mixed_vs_fp64"] = round(rc["fp64"]["seconds"] / rc["mixed"]["seconds"], 3)
        # the time to each residual level both precisions reached, and its ratio
        rc["speedup_mixed_vs_fp64_to_level"] = {
            lvl: round(rc["fp64"]["first_reached"][lvl]["seconds"] / rc["mixed"]["first_reached"][lvl]["seconds"], 3)
            for lvl in rc["mixed"]["first_reached"] if lvl in rc["fp64"]["first_reached"]}
        for name in (args.paths.split(",") if args.newton else []):
            prec = "mixed" if name == "mixed" else "fp64"
            log(f"time to 1e-8, {name} path, Newton operator")
            ttc["newton"][name] = converge_path(f"newton {name}",
                                                cases.swept_wing_case(mesh=mesh, precision=prec, **newton),
                                                mesh, args.converge_steps, dm)
        nw = ttc["newton"]
        if not args.newton:
            nw["not_run"] = "pass --newton"
        if all(nw.get(p, {}).get("reached") for p in ("mixed", "fp64")):
            nw["speedup_mixed_vs_fp64"] = round(nw["fp64"]["seconds"] / nw["mixed"]["seconds"], 3)
        if "mixed" in nw and "fp64" in nw:
            nw["abs_dCd_mixed_vs_fp64"] = abs(nw["mixed"]["Cd"] - nw["fp64"]["Cd"])
        nw["settings"] = ("extension, on the inviscid form of the wing: exact first-order Jacobian, Roe-ILU(0) "
                          f"preconditioner, CFL ramp 2 -> {newton['max_co']:g}, GMRES(30) relTol 1e-2, 4 cycles, "
                          "rho/rhoE update limited to 20 % per cell and step")
        ttc["roe_cfl20"]["settings"] = "the reference's Roe operator, LTS CFL ramp to 20, GMRES(30) 1 cycle relTol 1e-2"

    cpu = None
    if not args.no_cpu:
        case_kw = {"turbulence": turb} if turb else {}
        cpu = cpu_baseline(args, mesh_kw, case_kw, wd, nI + nB)

    roe_mixed = ttc["roe_cfl20"].get("mixed") if ttc else None
    out = {
        "metric": "time to 1e-8 residual drop + Gfaces/s flux, 2M-cell 3D SA, mixed vs FP64",
        "value": round(gfaces, 3),
        "unit": "Gfaces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "ms_per_step_path": "mixed precision (FP32 Jacobian, ILU and Krylov in FP64 defect correction; FP64 residual)",
        "ms_per_step_fp64": round(ms64, 3),
        "ms_per_step_fp64_partitioned": round(ms64p, 3),
        "ms_per_step_mixed": round(ms32, 3),
        "mixed_speedup_vs_plain_fp64": round(ms64 / ms32, 3),
        "mixed_speedup_same_preconditioner": round(ms64p / ms32, 3),
        "residual_drop_fp64": float(res64[-1] / res64[0]),
        "residual_drop_fp64_partitioned": float(res64p[-1] / res64p[0]),
        "residual_drop_mixed": float(res32[-1] / res32[0]),
        "time_to_1e-8_s": roe_mixed["seconds"] if roe_mixed and roe_mixed["reached"] else None,
        "time_to_1e-8_path": "mixed precision, the reference's Roe operator at CFL 20 (null: not reached within "
                             "the step budget; time_to_1e-8.roe_cfl20 gives the drop reached)",
        "time_to_1e-8": ttc,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "dtype_note": "value: the FP64 face-flux kernel; ms_per_step: the mixed path",
        "data": "synthetic",
        "config": {"workload": f"swept wing C-H mesh {nC} cells ({nI} internal + {nB} boundary faces), "
                               "M0.84 AoA3.06 Re1.2e7, AUSM+up first order, "
                               + (f"{turb} Spalart-Allmaras, no-slip wing, viscous terms" if turb
                                  else "inviscid slip wall")
                               + ", Roe-Jacobian implicit step, LTS CFL ramp to 20, GMRES(30)+block-ILU(0)",
                   "cells": nC, "faces": nI + nB, "flux_scheme": "AUSM+up", "turbulence": turb or "none"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
