"""Benchmark of the WENOFV time-stepping path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one SSP-RK3 time step (3 fused stage launches) over the workload; inputs are
resident in HBM before the timed region, which is bracketed by device synchronisation and
timed with HIP events recorded on the solver's stream.  Prints one JSON line.

Workloads: liquid_column (1024x512, walls, WENO-Z + HLLC + limiter, FP64 or mixed).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "two-phase-3equation-2d-pal-column_amd")
sys.path.insert(0, PKG)

BYTES_PER_CELL_STAGE = 256.0 / 3.0  # FP64 state, 4 fields, SSP-RK3: (64 + 96 + 96) B per cell-step
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="liquid_column")
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    return ap.parse_args()


def cpu_baseline(wl, seconds):
    """The oracle (FP64 CPU restatement, OpenMP over rows) on the same initial state, for as
    many steps as fit in `seconds` (at least one)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    kw = dict(wl.params)
    o = orc.Oracle(wl.nx, wl.ny, domain=wl.domain, bc=wl.bc, nranks=1, scheme=wl.opts["scheme"],
                   recon=wl.opts["recon"], flux=wl.opts["flux"], bp=wl.opts["bp"], rk=wl.opts["rk"],
                   cfl=wl.opts["cfl"], **kw)
    U = wl.U0
    steps = 0
    t0 = time.perf_counter()
    while True:
        U, dts, t, err = o.run(U, fixed_steps=1)
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rate = wl.nx * wl.ny * 3 * steps / el
    return {"value": rate, "unit": "cell-stage-updates/s", "cores": cores, "kind": "port",
            "sample": f"{wl.name} {wl.nx}x{wl.ny}, {steps} RK3 steps from the initial state, FP64 oracle "
                      f"(OpenMP, {cores} threads), {el:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); multi-GPU sharding of this path "
              "is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)
    import torch
    import cases
    import wenofv

    if not torch.cuda.is_available():
        sys.exit("no GPU visible")
    if a.workload == "liquid_column":
        wl = cases.liquid_column(precision=a.precision)
    elif a.workload == "perturbed_column":
        wl = cases.perturbed_column(precision=a.precision)
    else:
        sys.exit(f"unknown workload {a.workload}")
    sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs())
    sol.set_state(wl.U0)
    stream = torch.cuda.Stream()
    sol.set_stream(stream.cuda_stream)
    if a.warmup:
        sol.advance(fixed_steps=max(a.warmup, 32))  # also captures the step graph
    sol.set_state(wl.U0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    n, t, _ = sol.advance(fixed_steps=a.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    flags = sol.flags()
    assert n == a.steps, (n, a.steps)
    cells = wl.nx * wl.ny
    stages = 3 * a.steps
    rate = cells * stages / (ms * 1e-3)
    stage_ms = ms / stages
    achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9
    U = sol.get_state()
    out = {
        "metric": "cell-stage-updates/s",
        "value": rate,
        "unit": "cell-stage-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if wl.opts.get("precision", 0) == 0 else "f32+f64",
        "data": "synthetic (seeded initial condition, no dataset)",
        "config": {"workload": f"{wl.name} {wl.nx}x{wl.ny} walls WENO-Z HLLC BP RK3 CFL0.4",
                   "precision": "fp64" if wl.opts.get("precision", 0) == 0 else "mixed",
                   "cells": cells, "final_time": t, "flags": flags,
                   "finite": bool(np.isfinite(U).all())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "VALU-bound stencil; achieved = 85.33 B/cell-stage x cells / mean stage time"},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
