#!/usr/bin/env python3
"""Grain-steps/s of the per-step granular pipeline on one MI355X.

Workload (one "step" = one take_step of the whole pipeline over the grain
set, with the workload's diagnostics): the 32,768-grain (32^3 jittered
lattice) column collapse, mu = 0.5, semi-implicit Euler at dt = 1e-3, with the
total energy recorded every step (fused into the force pass) and a 64^3
cloud-in-cell deposit every 40 steps (kept on the device).  Grains are
synthetic (a seeded lattice), state resident in HBM before timing starts.

Also measured, as secondary lines inside the same JSON object: the
1,048,576-grain column (128 x 128 x 64, 4 x 4 m box, CIC 128^3 every 100
steps) and the CPU oracle on the host's cores (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
# Algorithmic bytes per grain-step of the semi-implicit Euler step: cell key
# (read position 12 + write key 4), stable reorder (read key 4 + state 24 +
# material 4, write state 24 + material 4 + permutation 4), fused
# force + integrate (read state 24 + material 4, write state 24).
B_STEP = 132.0
# The force kernel performs the force + integrate pass (52 B) and writes the
# next cell key (4 B).
B_FORCE = 56.0


def cic_bytes_per_grain_step(n: int, nodes: int, every: int) -> float:
    # read position 12 + velocity 12 + material 4 per grain, write 16 B per
    # node, once every `every` steps
    return (28.0 + 16.0 * nodes / n) / every


def run_gpu(cfg, steps: int, warmup: int, cic_every: int, energy: bool, profile_steps: int):
    from world_amd.sim import ParticleDynamics

    sim = ParticleDynamics(cfg)
    n = sim.n
    if energy:
        sim.enable_energy_log(steps + warmup + profile_steps + 8)
    ncic = (steps + warmup) // cic_every + 2
    sim.enable_density_log(ncic)
    # warmup (also captures and instantiates the step graphs)
    done = 0
    while done < warmup:
        k = min(cic_every, warmup - done)
        sim.take_step(k)
        done += k
    sim.sync()
    t0 = time.perf_counter()
    left = steps
    while left > 0:
        k = min(cic_every, left)
        sim.take_step(k)
        left -= k
        if k == cic_every:
            sim.record_density_grid()
    sim.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    grids = sim.density_log()
    elog = sim.energy_log() if energy else None
    # Per-kernel durations: HIP events around every launch on the engine's
    # stream, for profile_steps further steps of the same run.
    kernels = sim.profile_step(profile_steps)
    stable = sim.is_stable()
    overflow = sim.grid_overflow_count()
    sim.close()
    return {
        "n": n, "elapsed": elapsed, "kernels_ms": kernels, "stable": stable,
        "overflow": overflow, "grids": grids, "energy": elog,
    }


def roofline_of(kernels_ms: dict, n: int) -> dict:
    ms = kernels_ms.get("k_force", 0.0)
    achieved = B_FORCE * n / (ms * 1e-3) if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": "k_force", "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": None, "bytes_per_grain": B_FORCE, "avg_launch_ms": ms}


def cpu_baseline(cfg, target_s: float = 12.0) -> dict:
    from oracle import cpu
    from world_amd import _capi

    threads = min(16, os.cpu_count() or 1)
    st, mat = _capi.initial_state(cfg)
    n = st.size // 6
    # calibrate, then time a bounded sample of steps
    t = time.perf_counter()
    st1, _, _ = cpu.steps(cfg, st, mat, 2, threads=threads)
    per = (time.perf_counter() - t) / 2
    k = int(max(4, min(400, target_s / max(per, 1e-6))))
    t = time.perf_counter()
    cpu.steps(cfg, st1, mat, k, threads=threads)
    dt = time.perf_counter() - t
    return {"value": n * k / dt, "unit": "grain-steps/s", "cores": threads, "kind": "port",
            "sample": f"{k} steps of the {n}-grain workload from its initial lattice, "
                      f"FP32 oracle (oracle/oracle.cpp), OpenMP over grains"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--large-steps", type=int, default=500,
                    help="steps of the 1,048,576-grain column (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-steps", type=int, default=50)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus must be 1 "
              "(multi-GPU slab decomposition is a design, not part of this build)",
              file=sys.stderr)
        return 2

    from world_amd import workloads

    cfg = workloads.column_collapse_32k(friction=0.5)
    main_run = run_gpu(cfg, args.steps, args.warmup, 40, True, args.profile_steps)
    n = main_run["n"]
    value = n * args.steps / main_run["elapsed"]
    nodes = 64 ** 3
    b_total = B_STEP + cic_bytes_per_grain_step(n, nodes, 40)
    out = {
        "metric": "grain-steps/s (32,768-grain column collapse, SIE, mu 0.5, energy every step, "
                  "64^3 CIC every 40 steps)",
        "value": value,
        "unit": "grain-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * main_run["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": "column_collapse_32k", "grains": n, "integrator": "semi_implicit_euler",
                   "friction": 0.5, "dt": 1e-3, "cic_nodes": "64^3", "cic_every": 40,
                   "energy_every": 1, "parallelism": "single_gpu"},
        "roofline": roofline_of(main_run["kernels_ms"], n),
        "step_roofline": {"bytes_per_grain_step": b_total,
                          "achieved_GBps": value * b_total / 1e9,
                          "frac": value * b_total / HBM_PEAK},
        "kernels_ms": main_run["kernels_ms"],
        "stable": main_run["stable"],
        "overflow_cells": main_run["overflow"],
    }
    if args.large_steps > 0:
        big = workloads.column_collapse_1m()
        lr = run_gpu(big, args.large_steps, 10, 100, False, 20)
        lv = lr["n"] * args.large_steps / lr["elapsed"]
        lb = B_STEP + cic_bytes_per_grain_step(lr["n"], 128 ** 3, 100)
        out["large"] = {
            "workload": "column_collapse_1m", "grains": lr["n"], "steps": args.large_steps,
            "value": lv, "unit": "grain-steps/s", "ms_per_step": 1e3 * lr["elapsed"] / args.large_steps,
            "step_roofline": {"bytes_per_grain_step": lb, "achieved_GBps": lv * lb / 1e9,
                              "frac": lv * lb / HBM_PEAK},
            "roofline": roofline_of(lr["kernels_ms"], lr["n"]),
            "kernels_ms": lr["kernels_ms"], "stable": lr["stable"], "overflow_cells": lr["overflow"],
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
