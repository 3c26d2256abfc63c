"""Benchmark of the WENOFV time-stepping path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one SSP-RK3 time step (3 fused stage launches) over the whole workload; inputs
are resident in HBM before the timed region, which is bracketed by device synchronisation
and timed with HIP events recorded on the solver's stream.  Prints one JSON line.

Workloads
  batch256 (default)  256 independent liquid-column instances, grids drawn from
                      {64,128,256,512}^2 (seed 4), WENO-Z + HLLC + limiter, RK3, CFL 0.4,
                      one launch per stage over all instances, --precision (default 1,
                      mixed).  The same batch in all-FP64 gives the baseline throughput and
                      the reference state for the error check (relative L1 per field, always
                      after the workload's 500 steps, outside the timed region); a mixed
                      headline also times and checks the other mixed precision.
  settings_cfg        the reference's shipped config/settings.cfg (SMOOTH 100x100, WENO-Z,
                      characteristic, LLF, CFL 0.5) to the case's final time, --precision
  smooth_1024         the same options and case at 1024x1024, 400 steps, --precision [--compare]
  tanh_interface      128x128 periodic smooth interface, WENO-JS + LLF, to t = 1, --precision
  liquid_column       1024x512 single case (walls, WENO-Z + HLLC + limiter), --precision
  perturbed_column    4096x2048 single case, --precision [--compare]
  multi_column        8192x8192, 64 columns and a ratio-3 pressure step, --precision [--compare]
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

METRIC = "cell-stage-updates/s mixed+batched vs FP64 GPU at ≤1e-5 L1; achieved HBM GB/s"
BYTES_PER_CELL_STAGE = 256.0 / 3.0  # FP64 state, 4 fields, SSP-RK3: (64 + 96 + 96) B per cell-step
HBM_PEAK_GBS = 8000.0
# the algorithm's floating-point operations per cell-stage (reconstruction, limiter, Riemann
# fluxes, RK combination and the CFL step; each value counted once; divide = square root =
# compare = 1), keyed by (scheme, reconstruction, flux, limiter): the operation count of the
# oracle, which performs the FP64 path's arithmetic (oracle/flop_count.cpp; the figures are
# checked against it by tests/test_flop_count.py)
ALGO_FLOPS = {(0, 0, 0, 0): 2781.0, (1, 0, 1, 1): 3868.3, (1, 0, 0, 0): 3259.6, (1, 1, 0, 0): 5313.7,
              (1, 1, 1, 1): 5902.0, (1, 2, 1, 1): 3965.5, (1, 3, 1, 1): 3981.5, (1, 4, 1, 1): 3629.9,
              (2, 0, 1, 1): 4377.9}
VECTOR_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}  # MI355X, FMA = 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0: the workload's own count)")
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--workload", default="batch256")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--ty", type=int, default=0, help="rows per wavefront strip (0: default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="batch256: skip the FP64 baseline run")
    ap.add_argument("--compare", action="store_true", help="single workloads: also run all-FP64 and report errors")
    ap.add_argument("--scheme", type=int, default=None, help="single workloads: override SCHEME (0 JS, 1 Z, 2 EXP)")
    ap.add_argument("--recon", type=int, default=None, help="single workloads: override RECONSTRUCTIONTYPE (0-4)")
    ap.add_argument("--flux", type=int, default=None, help="single workloads: override RIEMANNFLUX (0 LLF, 1 HLLC)")
    ap.add_argument("--tune-cache", action="store_true",
                    help="reuse earlier autotuner choices (WENOFV_TUNING_CACHE) instead of timing the candidates again")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--spread-directions", type=int, default=3,
                    help="batch256: seeded random directions of the 6e-8 perturbation of the FP64 spread")
    ap.add_argument("--plain-steps", type=int, default=32,
                    help="batch256: timed steps per instance of the plain FP64 per-instance baseline (0: skip it)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a device buffer of this size once (HBM counter calibration)")
    return ap.parse_args()


def derived(U, rho10, rho20, sound1, sound2, **_):
    c1, c2 = sound1 ** 2, sound2 ** 2
    q = rho20 * c2 - rho10 * c1
    qs = U[1] * c2 - U[0] * c1
    r = ((q - qs) + np.sqrt((q - qs) ** 2 + 4.0 * U[0] * c1 * U[1] * c2)) / (2.0 * U[1] * c2)
    z1 = r / (1 + r)
    rho = U[0] + U[1]
    return z1, rho, U[2] / rho, U[3] / rho


def field_errors(Um, Ud, params):
    """Relative L1 of the mixed state against the FP64 state: conserved fields, z1, rho, u,
    v.  Momentum and velocity components are normalised by their vector's magnitude."""
    z1m, rm, um, vm = derived(Um, **params)
    z1d, rd, ud, vd = derived(Ud, **params)
    mom = np.abs(np.hypot(Ud[2], Ud[3])).sum()
    vel = np.abs(np.hypot(ud, vd)).sum()

    def l1(a, b, s=None):
        return float(np.abs(a - b).sum() / max(np.abs(b).sum() if s is None else s, 1e-300))

    errs = {"z1rho1": l1(Um[0], Ud[0]), "z2rho2": l1(Um[1], Ud[1]), "rhou": l1(Um[2], Ud[2], mom),
            "rhov": l1(Um[3], Ud[3], mom), "z1": l1(z1m, z1d), "rho": l1(rm, rd), "u": l1(um, ud, vel),
            "v": l1(vm, vd, vel)}
    linf = {"z1": float(np.abs(z1m - z1d).max()), "rho": float(np.abs(rm - rd).max() / np.abs(rd).max()),
            "u": float(np.abs(um - ud).max() / np.hypot(ud, vd).max()),
            "v": float(np.abs(vm - vd).max() / np.hypot(ud, vd).max())}
    return errs, linf, (float(z1m.min()), float(z1m.max()))


def cpu_baseline_oracle(nx, ny, domain, bc, params, opts, U0, seconds, label):
    """The oracle (FP64 CPU restatement, OpenMP over rows) on the same initial state, for as
    many RK3 steps as fit in `seconds` (at least one)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    o = orc.Oracle(nx, ny, domain=domain, bc=bc, nranks=1, scheme=opts["scheme"], recon=opts["recon"],
                   flux=opts["flux"], bp=opts["bp"], rk=opts["rk"], cfl=opts["cfl"], **params)
    U, steps = U0, 0
    t0 = time.perf_counter()
    while True:
        U, dts, t, err = o.run(U, fixed_steps=1)
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nx * ny * 3 * steps / el, "unit": "cell-stage-updates/s", "cores": cores, "kind": "port",
            "sample": f"{label}: {nx}x{ny}, {steps} RK3 steps from the initial state, all-FP64 oracle "
                      f"(OpenMP {cores} threads), {el:.1f} s"}


def timed(torch, obj, stream, steps):
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    res = obj.advance(fixed_steps=steps)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1), res


def profiled(name, key):
    """A machine-readable figure of a committed rocprofv3 summary under profiles/, or None."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def roofline(cells, stage_ms, opts, traffic=None):
    """The bound that applies to the stage kernel: vector ALU throughput.  achieved = the
    algorithm's operations per cell-stage x cells / mean stage time, against MI355X's vector
    peak without packed FP32 (78.6 TFLOP/s, FP32 and FP64 alike; packed FP32 instructions
    are not used).  The HBM roofline of the same launches sits beside it under "hbm"."""
    hbm_achieved = BYTES_PER_CELL_STAGE * cells / (stage_ms * 1e-3) / 1e9
    hbm = {"bound": "hbm", "achieved": hbm_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": hbm_achieved / HBM_PEAK_GBS, "traffic": traffic,
           "note": "achieved = 85.33 algorithmic B per cell-stage (FP64 state in and out) x cells / mean stage "
                   "time; traffic = measured HBM bytes per stage launch (rocprofv3 summary under profiles/)"}
    f = ALGO_FLOPS.get((opts["scheme"], opts["recon"], opts["flux"], opts["bp"]))
    if f is None:
        return hbm
    tf = f * cells / (stage_ms * 1e-3) / 1e12
    peak = VECTOR_PEAK_TFLOPS["fp64"]
    return {"bound": "valu", "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
            "traffic": traffic, "per_cell_stage_flops": f, "hbm": hbm,
            "note": f"achieved = {f} algorithmic operations per cell-stage (oracle/flop_count.cpp; divide = "
                    "square root = compare = 1) x cells / mean stage time (HIP events over the timed region); "
                    "peak = vector FP64 = non-packed FP32 peak"}


def algo_flops(opts, rate):
    """Algorithmic FLOP rate of a run at `rate` cell-stage updates/s, against the vector peaks."""
    f = ALGO_FLOPS.get((opts["scheme"], opts["recon"], opts["flux"], opts["bp"]))
    if f is None:
        return None
    tf = f * rate / 1e12
    return {"per_cell_stage": f, "achieved_tflops": tf,
            "frac_fp64_vector_peak": tf / VECTOR_PEAK_TFLOPS["fp64"],
            "frac_fp32_vector_peak": tf / VECTOR_PEAK_TFLOPS["fp32"],
            "note": "operations of the algorithm (each value once, divide = sqrt = compare = 1; "
                    "oracle/flop_count.cpp) x cell-stage updates/s; the hardware executes more: an FP64 "
                    "division or square root is a sequence of ~10 instructions, and packed FP32 is not used"}


def plain_fp64_per_instance(torch, wenofv, wls, opts, stream, steps):
    """The plain all-FP64 single-domain solver run on one instance after another (the path a
    user without batching takes): each instance warms up (graph capture), then `steps`
    steps are timed on the solver's stream.  Returns cell-stage-updates/s over the batch."""
    total_ms, work = 0.0, 0
    for w in wls:
        s = wenofv.Solver(w.nx, w.ny, **w.solver_kwargs(**dict(opts, precision=0)))
        s.set_stream(stream.cuda_stream)
        s.set_state(w.U0)
        s.autotune(trials=2)  # rows per strip for this grid (cached per grid size)
        s.advance(fixed_steps=32)
        ms, (n, _, _) = timed(torch, s, stream, steps)
        assert n == steps, (n, steps)
        total_ms += ms
        work += w.nx * w.ny * 3 * steps
        s.close()
    return work / (total_ms * 1e-3), total_ms


GATE = ("z1", "rho", "u", "v")
FIELDS = ("z1rho1", "z2rho2", "rhou", "rhov", "z1", "rho", "u", "v")


def field_maps(U, params):
    z1, r, u, v = derived(U, **params)
    return {"z1rho1": U[0], "z2rho2": U[1], "rhou": U[2], "rhov": U[3], "z1": z1, "rho": r, "u": u, "v": v}


def where_worst(fm, fd, w):
    """Where the largest point error of the worst velocity component sits: the cell, z1 there,
    the largest z1 and relative density jumps to a neighbour (interface and shock indicators)
    and the distance to the nearest wall; class = wall (<= 3 cells), interface (z1 jump >
    1e-3), shock (density jump > 1e-2) or smooth."""
    du = np.hypot(fm["u"] - fd["u"], fm["v"] - fd["v"])
    j, i = (int(q) for q in np.unravel_index(np.argmax(du), du.shape))
    z1, rho = fd["z1"], fd["rho"]
    jn = [(j, min(i + 1, w.nx - 1)), (j, max(i - 1, 0)), (min(j + 1, w.ny - 1), i), (max(j - 1, 0), i)]
    zj = max(abs(float(z1[q]) - float(z1[j, i])) for q in jn)
    rj = max(abs(float(rho[q]) - float(rho[j, i])) for q in jn) / float(rho[j, i])
    wall = min(i, j, w.nx - 1 - i, w.ny - 1 - j)
    cls = "wall" if wall <= 3 else ("interface" if zj > 1e-3 else ("shock" if rj > 1e-2 else "smooth"))
    return {"i": i, "j": j, "z1": float(z1[j, i]), "z1_jump": zj, "rho_jump_rel": rj, "wall_dist": int(wall),
            "class": cls}


def accuracy_block(wls, S, Sd, S13, S8s, steps, Sfma=None):
    """Relative L1 per field of the states S against the all-FP64 states Sd over the batch
    (aggregate) and per instance, beside the FP64 solver's own spread from initial data
    perturbed by 1e-13 (S13) and by 6e-8 (FP32's unit roundoff) in several seeded random
    directions (S8s: one list of states per direction).  An instance's 6e-8 spread is the
    largest over the directions; the first direction's alone is reported beside it.  Sfma:
    the states of the all-FP64 solver compiled with fused multiply-add (precision 3), which
    differs from the bit-faithful one by rounding alone, at every step; an instance's FP64
    floor is the largest of its 1e-13 and 6e-8 spreads and that solver's deviation."""
    num = {k: 0.0 for k in FIELDS}
    den = dict(num)
    num13, num8 = dict(num), dict(num)
    inst, sp13, sp8, sp8a, fma, zb, mass, beyond = [], [], [], [], [], [1.0, 0.0], 0.0, []

    def gmax(e):
        return max(e[0][k] for k in GATE)

    for k, w in enumerate(wls):
        fm, fd, f13, f8 = (field_maps(U[k], w.params) for U in (S, Sd, S13, S8s[0]))
        mom, vel = np.hypot(Sd[k][2], Sd[k][3]).sum(), np.hypot(fd["u"], fd["v"]).sum()
        scale = {"rhou": mom, "rhov": mom, "u": vel, "v": vel}
        for key in FIELDS:
            num[key] += np.abs(fm[key] - fd[key]).sum()
            num13[key] += np.abs(f13[key] - fd[key]).sum()
            num8[key] += np.abs(f8[key] - fd[key]).sum()
            den[key] += scale[key] if key in scale else np.abs(fd[key]).sum()
        e = field_errors(S[k], Sd[k], w.params)
        inst.append(gmax(e))
        sp13.append(gmax(field_errors(S13[k], Sd[k], w.params)))
        per_dir = [gmax(field_errors(S8[k], Sd[k], w.params)) for S8 in S8s]
        sp8a.append(per_dir[0])
        sp8.append(max(per_dir))
        fma.append(gmax(field_errors(Sfma[k], Sd[k], w.params)) if Sfma is not None else 0.0)
        if inst[-1] > max(1e-5, sp8[-1]):
            worst = max(GATE, key=lambda q: e[0][q])
            beyond.append({"instance": k, "grid": [w.nx, w.ny], "field": worst, "rel_l1": inst[-1],
                           "fp64_6e-8_spread": sp8[-1], "fp64_6e-8_spread_per_direction": per_dir,
                           "fp64_1e-13_spread": sp13[-1], "fp64_fma_deviation": fma[-1],
                           "where": where_worst(fm, fd, w)})
        zb = [min(zb[0], float(fm["z1"].min())), max(zb[1], float(fm["z1"].max()))]
        m0 = w.U0[:2].sum(axis=(1, 2))
        mass = max(mass, float(np.abs(S[k][:2].sum(axis=(1, 2)) / m0 - 1).max()))
    agg = {k: float(num[k] / den[k]) for k in FIELDS}
    inst, sp13, sp8, sp8a, fma = np.array(inst), np.array(sp13), np.array(sp8), np.array(sp8a), np.array(fma)
    floor = np.maximum.reduce([np.full_like(inst, 1e-5), sp8, sp13, fma])

    def dist(x):
        return {"median": float(np.median(x)), "p90": float(np.percentile(x, 90)), "max": float(x.max())}

    return {
        "steps": steps, "rel_l1_per_field_batch": agg, "gate_fields": list(GATE),
        "max_rel_l1_batch": max(agg[k] for k in GATE),
        "max_rel_l1_batch_conserved": max(agg[k] for k in FIELDS[:4]),
        "max_rel_l1_batch_all_fields": max(agg.values()),
        "tolerance": 1e-5, "batch_within_tolerance": max(agg.values()) <= 1e-5,
        "per_instance_max_field_rel_l1": dict(dist(inst), fraction_within_tolerance=float((inst <= 1e-5).mean())),
        "fp64_self_spread_1e-13_perturbation": dict(dist(sp13), fraction_beyond_tolerance=float((sp13 > 1e-5).mean())),
        "fp64_self_spread_6e-8_perturbation": dict(dist(sp8), directions=len(S8s)),
        "fp64_self_spread_per_field_batch": {"1e-13": {k: float(num13[k] / den[k]) for k in FIELDS},
                                             "6e-8": {k: float(num8[k] / den[k]) for k in FIELDS}},
        "fraction_within_tolerance_or_fp64_6e-8_spread": float((inst <= np.maximum(1e-5, sp8)).mean()),
        "fraction_within_tolerance_or_fp64_6e-8_spread_first_direction":
            float((inst <= np.maximum(1e-5, sp8a)).mean()),
        "beyond_tolerance_and_spread": beyond,
        "fraction_within_tolerance_or_fp64_floor": float((inst <= floor).mean()),
        "beyond_tolerance_and_fp64_floor": [b["instance"] for b in beyond if b["rel_l1"] > floor[b["instance"]]],
        "fp64_fma_deviation": (dict(dist(fma), fraction_beyond_tolerance=float((fma > 1e-5).mean()))
                               if Sfma is not None else None),
        "z1_range": zb, "z1_in_unit_interval": zb[0] >= 0.0 and zb[1] <= 1.0, "max_phase_mass_drift": mass}


def run_batch(a, torch, wenofv, cases):
    """The batch workload.  Timed runs at --steps (default: the workload's own 500): the
    requested precision (the headline), all-FP64 and, for a mixed headline, the other mixed
    precision.  The accuracy blocks always compare states after the workload's own 500 steps
    (the timed runs' states when --steps is 500, else untimed runs after the timed ones)."""
    specs = cases.batch_specs(256, seed=4)
    wls = [cases.batch_instance(sp) for sp in specs]
    opts = dict(wls[0].opts)
    acc_steps = wls[0].steps
    steps = a.steps or acc_steps
    insts = [dict(nx=w.nx, ny=w.ny, bc=w.bc, domain=w.domain, **w.params) for w in wls]
    cells = sum(w.nx * w.ny for w in wls)
    stream = torch.cuda.Stream()
    tiles = {}

    def go(prec, nsteps, warmup=0, perturb=0.0, tune=False, seed=17):
        b = wenofv.Batch(insts, rows_per_strip=a.ty, **dict(opts, precision=prec))
        b.set_stream(stream.cuda_stream)
        for k, w in enumerate(wls):
            b.set_state(k, w.U0)
        if tune and not a.ty:  # rows per strip per grid class and the schedule from the autotuner
            tiles[prec] = b.autotune(trials=2, cache=a.tune_cache)
        if warmup:
            b.advance(fixed_steps=warmup)  # also captures the step graph
        rng = np.random.default_rng(13 if perturb < 1e-10 else seed)
        for k, w in enumerate(wls):
            b.set_state(k, w.U0 * (1 + perturb * rng.standard_normal(w.U0.shape)) if perturb else w.U0)
        ms, (n, t, f) = timed(torch, b, stream, nsteps)
        assert (n == nsteps).all(), n
        res = dict(ms=ms, flags=int(np.bitwise_or.reduce(f)), states=[b.get_state(k) for k in range(len(wls))])
        b.close()
        return res

    head = a.precision
    other = {1: 2, 2: 1}.get(head) if not a.no_fp64 else None
    timed_runs = {head: go(head, steps, a.warmup, tune=True)}
    if not a.no_fp64 and head != 0:
        timed_runs[0] = go(0, steps, a.warmup, tune=True)
    if other is not None:
        timed_runs[other] = go(other, steps, a.warmup, tune=True)
    rate = cells * 3 * steps / (timed_runs[head]["ms"] * 1e-3)
    pname = {0: "fp64", 1: "mixed", 2: "mixed_w64"}
    out = {
        "metric": METRIC, "value": rate, "unit": "cell-stage-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": timed_runs[head]["ms"] / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if head == 0 else "f32+f64",
        "data": "synthetic (seeded initial conditions, no dataset)",
        "config": {"workload": f"batch256: 256 liquid-column instances (seed 4), grids {{64,128,256,512}}^2, walls, "
                               f"WENO-Z HLLC BP RK3 CFL0.4, {steps} steps each, one launch per stage",
                   "precision": pname[head], "instances": len(wls), "cells": cells,
                   "flags": timed_runs[head]["flags"], "tiles": tiles.get(head, {"rows_per_strip": a.ty})},
    }
    prof = f"batch_stage_{pname[head]}_batch256.md"
    stage_ms = timed_runs[head]["ms"] / (3 * steps)
    out["roofline"] = roofline(cells, stage_ms, opts, profiled(prof, "traffic_bytes_per_launch"))
    out["compute"] = {"busy_frac": profiled(prof, "valu_busy_frac"), "issue_frac": profiled(prof, "valu_issue_frac"),
                      "lane_instructions_per_cell_stage": profiled(prof, "lane_instructions_per_cell_stage"),
                      "profile_stage_ms": (profiled(prof, "stage_kernel_avg_ns") or 0) / 1e6 or None,
                      "bench_stage_ms": stage_ms, "source": f"profiles/{prof}",
                      "note": "issue_frac = SQ_INSTS_VALU per SIMD cycle / 0.25 (a wave64 VALU instruction "
                              "occupies a SIMD for 4 cycles); busy_frac = SQ_ACTIVE_INST_VALU quad-cycles / "
                              "(SIMDs x duration x clock / 4), multi-cycle transcendental and FP64 instructions "
                              "included (a sum over waves: it can pass 1 when two waves' instructions overlap)"}
    if not a.no_fp64 and a.plain_steps > 0:
        plain_rate, plain_ms = plain_fp64_per_instance(torch, wenofv, wls, opts, stream, a.plain_steps)
        out["fp64_plain_gpu"] = {
            "value": plain_rate, "steps_per_instance": a.plain_steps, "ms_total": plain_ms,
            "note": "the plain all-FP64 single-domain solver, one instance after another (graph of steps, "
                    "device-side dt), timed per instance after warm-up"}
        out["speedup_vs_fp64_plain_gpu"] = rate / plain_rate
    if 0 in timed_runs and head != 0:
        fp64_rate = cells * 3 * steps / (timed_runs[0]["ms"] * 1e-3)
        out["fp64_batched_gpu"] = {"value": fp64_rate, "ms_per_step": timed_runs[0]["ms"] / steps,
                                   "flags": timed_runs[0]["flags"], "algorithmic_flops": algo_flops(opts, fp64_rate),
                                   "note": "the same batch and launches in all-FP64 (precision is the only change)"}
        out["speedup_vs_fp64_batched_gpu"] = rate / fp64_rate
        # accuracy at the workload's own step count, outside the timed region
        states = {p: (r["states"] if steps == acc_steps else go(p, acc_steps)["states"]) for p, r in timed_runs.items()}
        S13 = go(0, acc_steps, perturb=1e-13)["states"]
        S8s = [go(0, acc_steps, perturb=6e-8, seed=17 + d)["states"] for d in range(a.spread_directions)]
        Sfma = go(3, acc_steps)["states"]
        out["accuracy"] = accuracy_block(wls, states[head], states[0], S13, S8s, acc_steps, Sfma)
        if other is not None:
            orate = cells * 3 * steps / (timed_runs[other]["ms"] * 1e-3)
            out[f"precision{other}"] = {
                "precision": pname[other], "value": orate, "ms_per_step": timed_runs[other]["ms"] / steps,
                "flags": timed_runs[other]["flags"], "speedup_vs_fp64_batched_gpu": orate / fp64_rate,
                "accuracy": accuracy_block(wls, states[other], states[0], S13, S8s, acc_steps, Sfma),
                "note": {1: "FP32 reconstruction and fluxes, FP64 state, EOS, RK update and CFL reduction",
                         2: "FP64 reconstruction, FP32 fluxes; FP64 state, EOS, RK update and CFL "
                            "reduction"}[other]}
    if not a.no_cpu:
        w = next(w for w in wls if (w.nx, w.ny) == (256, 256))
        out["cpu_baseline"] = cpu_baseline_oracle(w.nx, w.ny, w.domain, w.bc, w.params, opts, w.U0, a.cpu_seconds,
                                                  "one 256x256 instance of the batch")
    return out


def run_single(a, torch, wenofv, cases):
    """One case on the single-domain solver.  Fixed-step workloads are timed over their
    steps; tanh_interface runs the reference's time loop to its final time.  --compare also
    runs the case in all-FP64 and reports the field errors of the timed run against it."""
    def settings_cfg():
        # the reference's shipped config/settings.cfg: TESTCASE 0 (SMOOTH) 100x100, RK3, CFL 0.5,
        # WENO-Z, characteristic, LLF, no limiter, to the case's final time; 4-rank cell widths
        kw, t_end = wenofv.reference_case(0)
        U0 = wenofv.reference_initial_state(0, 100, 100, geometry_ranks=4)
        prm = {k: kw[k] for k in ("rho10", "rho20", "sound1", "sound2", "preb")}
        w = cases.Workload("settings_cfg", 100, 100, kw["domain"], prm, U0, kw["bc"],
                           dict(scheme=1, recon=1, flux=0, bp=0, rk=2, cfl=0.5, precision=a.precision), t_end=t_end)
        w.geometry_ranks = 4
        return w

    def smooth_1024():
        # the shipped config/settings.cfg's scheme options (WENO-Z, characteristic, LLF, no
        # limiter, RK3, CFL 0.5) on its case (TESTCASE 0, SMOOTH) at 1024x1024, 400 steps
        kw, _ = wenofv.reference_case(0)
        U0 = wenofv.reference_initial_state(0, 1024, 1024, geometry_ranks=1)
        prm = {k: kw[k] for k in ("rho10", "rho20", "sound1", "sound2", "preb")}
        return cases.Workload("smooth_1024", 1024, 1024, kw["domain"], prm, U0, kw["bc"],
                              dict(scheme=1, recon=1, flux=0, bp=0, rk=2, cfl=0.5, precision=a.precision), steps=400)

    makers = {"settings_cfg": settings_cfg,
              "smooth_1024": smooth_1024,
              "tanh_interface": lambda: cases.tanh_interface(),
              "liquid_column": lambda: cases.liquid_column(precision=a.precision),
              "perturbed_column": lambda: cases.perturbed_column(precision=a.precision),
              "multi_column": lambda: cases.multi_column(precision=a.precision)}
    if a.workload not in makers:
        sys.exit(f"unknown workload {a.workload}")
    wl = makers[a.workload]()
    if a.workload == "tanh_interface":
        wl.opts["precision"] = a.precision
    for key, val in (("scheme", a.scheme), ("recon", a.recon), ("flux", a.flux)):
        if val is not None:
            wl.opts[key] = val
    prec = wl.opts.get("precision", 0)
    steps = a.steps or wl.steps or 0
    stream = torch.cuda.Stream()

    def run(precision, U0=None):
        U0 = wl.U0 if U0 is None else U0
        sol = wenofv.Solver(wl.nx, wl.ny, **wl.solver_kwargs(rows_per_strip=a.ty, precision=precision,
                                                             geometry_ranks=getattr(wl, "geometry_ranks", 1)))
        sol.set_stream(stream.cuda_stream)
        sol.set_state(U0)
        tiles = sol.autotune(trials=2, cache=a.tune_cache) if not a.ty else {"rows_per_strip": a.ty}
        if a.warmup:
            sol.advance(fixed_steps=a.warmup)  # also captures the step graph
        sol.set_state(U0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        n, t, _ = sol.advance(fixed_steps=steps) if steps else sol.advance(t_end=wl.t_end)
        e1.record(stream)
        torch.cuda.synchronize()
        if steps:
            assert n == steps, (n, steps)
        return e0.elapsed_time(e1), n, t, sol.get_state(), sol.flags(), tiles

    ms, n, t, U, flags, tiles = run(prec)
    cells = wl.nx * wl.ny
    bcname = "walls" if wl.bc == 1 else "periodic"
    o = wl.opts
    recon_names = ("conservative", "characteristic", "primitive", "primitive2", "primitive3")
    desc = (f"{wl.name} {wl.nx}x{wl.ny} {bcname} "
            f"{'primitive3/THINC' if o['recon'] == 4 else ('WENO-JS', 'WENO-Z', 'WENO-EXP')[o['scheme']]} "
            f"{recon_names[o['recon']] if o['recon'] != 4 else ''} "
            f"{'HLLC' if o['flux'] else 'LLF'}{' BP' if o['bp'] else ''} RK3 CFL{o['cfl']}, "
            + (f"{steps} steps" if steps else f"to t={wl.t_end}"))
    out = {
        "metric": METRIC, "value": cells * 3 * n / (ms * 1e-3), "unit": "cell-stage-updates/s", "n_gpus": 1,
        "steps": int(n), "warmup": a.warmup, "ms_per_step": ms / n, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32+f64",
        "data": "synthetic (seeded initial condition, no dataset)",
        "config": {"workload": desc, "precision": ("fp64", "mixed", "mixed_w64")[prec], "cells": cells,
                   "final_time": t, "flags": flags, "finite": bool(np.isfinite(U).all()), "tiles": tiles},
        "roofline": roofline(cells, ms / (3 * n), wl.opts),
        "algorithmic_flops": algo_flops(wl.opts, cells * 3 * n / (ms * 1e-3)),
    }
    z1 = derived(U, **wl.params)[0]
    out["config"]["z1_range"] = [float(z1.min()), float(z1.max())]
    m0 = wl.U0[:2].sum(axis=(1, 2))
    out["config"]["max_phase_mass_drift"] = float(np.abs(U[:2].sum(axis=(1, 2)) / m0 - 1).max())
    if a.workload == "tanh_interface" and abs(t - 1.0) < 1e-9:
        # after t = 1 the exact solution is the initial state shifted by (1, 0.5): by half a
        # period in y
        z1x = derived(np.roll(wl.U0, wl.ny // 2, axis=1), **wl.params)[0]
        out["config"]["z1_l1_vs_exact"] = float(np.abs(z1 - z1x).mean())
    if a.compare and prec != 0:
        ms64, n64, t64, U64, f64, _ = run(0)
        errs, linf, _ = field_errors(U, U64, wl.params)
        gate = ("z1", "rho", "u", "v")
        out["fp64_gpu"] = {"value": cells * 3 * n64 / (ms64 * 1e-3), "ms_per_step": ms64 / n64, "flags": f64}
        out["speedup_vs_fp64_gpu"] = out["value"] / out["fp64_gpu"]["value"]
        # the FP64 solver's own sensitivity: the same run from the initial state perturbed by
        # 1e-13 relative (seeded); where it exceeds the tolerance the case is ill-conditioned
        # in FP64 itself and the mixed error is compared with 10x that spread
        rng = np.random.default_rng(13)
        *_, Up, _, _ = run(0, wl.U0 * (1 + 1e-13 * rng.standard_normal(wl.U0.shape)))
        spread, _, _ = field_errors(Up, U64, wl.params)
        # and by 6e-8 (FP32's unit roundoff): the FP64 solution's response to one float
        # rounding of its input, the floor for any arithmetic that rounds to FP32
        *_, Up8, _, _ = run(0, wl.U0 * (1 + 6e-8 * rng.standard_normal(wl.U0.shape)))
        spread8, _, _ = field_errors(Up8, U64, wl.params)
        worst = max(errs[k] for k in gate)
        spread_max = max(spread[k] for k in gate)
        spread8_max = max(spread8[k] for k in gate)
        out["accuracy"] = {"rel_l1_per_field": errs, "rel_linf": linf, "max_rel_l1": worst,
                           "gate_fields": list(gate), "tolerance": 1e-5, "within_tolerance": worst <= 1e-5,
                           "fp64_self_spread_1e-13": spread_max, "fp64_self_spread_6e-8": spread8_max,
                           "within_tolerance_or_10x_fp64_spread": worst <= max(1e-5, 10 * spread_max),
                           "within_tolerance_or_fp64_6e-8_spread": worst <= max(1e-5, spread8_max)}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_oracle(wl.nx, wl.ny, wl.domain, wl.bc, wl.params, wl.opts, wl.U0,
                                                  a.cpu_seconds, wl.name)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): this build has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import cases
    import wenofv

    if not torch.cuda.is_available():
        sys.exit("no GPU visible")
    out = run_batch(a, torch, wenofv, cases) if a.workload == "batch256" else run_single(a, torch, wenofv, cases)
    if a.copy_calibration_mib:
        n = a.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = {"read": n * 8, "write": n * 8}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
