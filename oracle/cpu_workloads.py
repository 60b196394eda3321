"""CPU baseline of the two smaller workloads: the oracle's Single::iterate on
the Mach-3 forward step (120 x 40, AUSM+, CFL 3, GMRES(20) + ILU(0)) run to a
1e-8 drop of |R|_2 (timeDST.C:328-331) or a time budget, and on the hybrid
airfoil (about 205 k cells, HLLC + SAneg) over a few steps.

TEST INFRASTRUCTURE ONLY (the oracle is the checker; here it is the CPU
reference the GPU numbers are reported beside).  Run as a child process of
bench.py so that the thread count is fixed before the OpenMP runtime starts:

    OMP_NUM_THREADS=16 ORACLE_OPENMP=1 python -m oracle.cpu_workloads --workload forward_step --seconds 240

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("forward_step", "hybrid_airfoil"), required=True)
    ap.add_argument("--seconds", type=float, default=120.0, help="time budget of the iteration")
    ap.add_argument("--max-steps", type=int, default=20000)
    args = ap.parse_args(argv)
    from rhodst_amd import cases, meshgen

    from oracle import binding
    if args.workload == "forward_step":
        spec = cases.forward_step_case(120, 40, initial_co=3.0, min_co=3.0, max_co=3.0)
        chunk = 100
    else:
        mesh = meshgen.hybrid_airfoil()
        wd = meshgen.wall_distance(mesh, ["airfoil"], centres=binding.geometry(mesh)["C"])
        spec = cases.airfoil_case(mesh=mesh, wall_dist=wd)
        chunk = 1
    oc = binding.OracleCase(spec)
    hist, spent, reached, at = [], 0.0, False, None
    next_log = 20.0
    while spent < args.seconds and len(hist) < args.max_steps:
        t0 = time.perf_counter()
        r = oc.iterate(chunk)
        spent += time.perf_counter() - t0
        hist.extend(r.tolist())
        h = np.array(hist) / hist[0]
        if spent >= next_log:
            print(f"[cpu {args.workload}] step {len(hist)} drop {h[-1]:.3e} {spent:.0f}s", file=sys.stderr, flush=True)
            next_log += 20.0
        if not np.isfinite(h[-1]):
            break
        if h.min() <= 1e-8:
            reached, at = True, int(np.argmax(h <= 1e-8)) + 1
            break
    h = np.array(hist) / hist[0]
    print(json.dumps({"workload": args.workload, "cells": spec.mesh.n_cells, "steps": len(hist), "seconds": spent,
                      "ms_per_step": spent / len(hist) * 1e3, "reached_1e-8": reached, "step_1e-8": at,
                      "final_drop": float(h[-1]), "threads": int(os.environ.get("OMP_NUM_THREADS", "1")),
                      "openmp": binding._OMP}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
