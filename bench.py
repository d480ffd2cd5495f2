#!/usr/bin/env python3
"""Grain-steps/s of the per-step granular pipeline on one MI355X.

Workload (one "step" = one take_step of the whole pipeline over the grain
set, with the workload's diagnostics): the 32,768-grain (32^3 jittered
lattice) column collapse, mu = 0.5, semi-implicit Euler at dt = 1e-3, with the
total energy recorded every step (fused into the force pass) and a 64^3
cloud-in-cell deposit every 40 steps (kept on the device).  Grains are
synthetic (a seeded lattice), state resident in HBM before timing starts.

Also measured, as secondary entries of the same JSON object:
  variants  the same column at mu = 0 and with backward-Euler-Picard (3
            iterations), mu = 0.5 and 0;
  large     the 1,048,576-grain column (128 x 128 x 64, 4 x 4 m box, CIC 128^3
            every 100 steps);
  ensemble  64 uneven members (2,048..32,768 grains, aspect 1..4, mu in
            [0, 0.6], 50 um Gaussian perturbation) advanced by one batched
            handle, against looping the same members one at a time;
  adjoint   the 8,192-grain window (mu 0.3, checkpoints every 20 steps): cost
            of a backward step against a forward step;
  batched_adjoint  16 uneven members (4,096..16,384 grains) of that adjoint
            in one batch handle, 100-step windows, checkpoints every 10
            steps, against the same windows member by member;
  sliding   one grain launched at 1 m/s on the floor at mu 0.5: fitted
            deceleration against -mu g = -4.905 m/s^2;
  cpu_baseline  the CPU oracle on the host's cores.

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
# With the spring_history tangential model the force pass also reads and
# writes each grain's stored springs: a 4 B count each way and 16 B per
# contact each way, so 8 + 32 Z bytes per grain for Z contacts per grain.


def history_bytes(z: float) -> float:
    return 8.0 + 32.0 * z


def cic_bytes_per_grain_step(n: int, nodes: int, every: int) -> float:
    # read position 12 + velocity 12 + material 4 per grain, write 16 B per
    # node, once every `every` steps
    return (28.0 + 16.0 * nodes / n) / every


def run_gpu(cfg, steps: int, warmup: int, cic_every: int, energy: bool, profile_steps: int, pin=None):
    from world_amd.sim import ParticleDynamics

    sim = ParticleDynamics(cfg)
    if pin:
        sim.set_force_variant(*pin)
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
    stable = sim.is_stable()
    overflow = sim.grid_overflow_count()
    variant = sim.force_variant()
    passes, contacts = sim.contact_stats()
    z = contacts / max(1, passes * n)
    sim.close()
    # Per-kernel durations over the same work: the run is deterministic, so a
    # second handle, with the force-pass variant the first one tuned, replays
    # the warmup + timed steps with kernel-stamped HIP events on every launch
    # on the engine's stream (profile_steps of them).
    kernels = {}
    if profile_steps > 0:
        prof = ParticleDynamics(cfg)
        prof.set_force_variant(*variant)
        if energy:
            prof.enable_energy_log(profile_steps + 8)
        kernels = prof.profile_step(profile_steps)
        prof.close()
    return {
        "n": n, "elapsed": elapsed, "kernels_ms": kernels, "stable": stable,
        "overflow": overflow, "grids": grids, "energy": elog,
        "force_variant": {"block": variant[0], "lanes_per_grain": variant[1]},
        "contacts_per_grain": z, "history": cfg.tangential_model == 1,
    }


def roofline_of(run: dict) -> dict:
    n = run["n"]
    ms = run["kernels_ms"].get("k_force", 0.0)
    bpg = B_FORCE + (history_bytes(run["contacts_per_grain"]) if run["history"] else 0.0)
    achieved = bpg * n / (ms * 1e-3) if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": "k_force", "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic_of("k_force", n, run["force_variant"]), "bytes_per_grain": bpg,
            "contacts_per_grain": run["contacts_per_grain"], "avg_launch_ms": ms}


def step_bytes(run: dict, nodes: int, every: int) -> float:
    b = B_STEP + cic_bytes_per_grain_step(run["n"], nodes, every)
    return b + (history_bytes(run["contacts_per_grain"]) if run["history"] else 0.0)


def ensemble_leg(steps: int) -> dict:
    """64-member uneven ensemble: one batched handle vs the members looped."""
    from world_amd import _capi, workloads
    from world_amd.sim import Ensemble, ParticleDynamics

    cs = workloads.uneven_ensemble(64)
    n = sum(_capi.grain_count(c) for c in cs)
    ens = Ensemble(cs)
    ens.take_step(20)
    ens.sync()
    t = time.perf_counter()
    ens.take_step(steps)
    ens.sync()
    tb = time.perf_counter() - t
    stable = ens.is_stable()
    ens.close()
    tl = 0.0
    for c in cs:
        sim = ParticleDynamics(c)
        sim.take_step(20)
        sim.sync()
        t = time.perf_counter()
        sim.take_step(steps)
        sim.sync()
        tl += time.perf_counter() - t
        sim.close()
    return {"workload": "uneven_ensemble_64", "members": len(cs), "grains": n, "steps": steps,
            "batched_ms_per_step": 1e3 * tb / steps, "batched_grain_steps_per_s": n * steps / tb,
            "looped_ms_per_step": 1e3 * tl / steps, "looped_grain_steps_per_s": n * steps / tl,
            "speedup": tl / tb, "stable": stable}


def adjoint_window(make, cot_of, window: int, every: int):
    """Forward, record and backward cost of one window, all over the same
    steps: a handle advanced 20 + window steps times a forward step over the
    next window; a second handle records and differentiates the window
    before it (its allocations and graph captures) and then the same window.
    Returns seconds (forward, record, backward) for the window."""
    with make() as a:
        a.take_step(20 + window)
        a.sync()
        t = time.perf_counter()
        a.take_step(window)
        a.sync()
        fwd = time.perf_counter() - t
    with make() as h:
        cot = cot_of(h)
        h.take_step(20)
        h.sync()
        rec = bwd = 0.0
        for _ in range(2):
            t = time.perf_counter()
            h.adjoint_record(window, every)
            h.sync()
            rec = time.perf_counter() - t
            t = time.perf_counter()
            h.adjoint_backward(cot)
            bwd = time.perf_counter() - t
    return fwd, rec, bwd


def adjoint_leg(window: int = 200, every: int = 20) -> dict:
    """Backward-step cost of the 8,192-grain density-loss window against a
    forward step over the same window."""
    import numpy as np

    from world_amd import workloads
    from world_amd.sim import ParticleDynamics

    c = workloads.column_8k(friction=0.3)
    fwd, rec, bwd = adjoint_window(
        lambda: ParticleDynamics(c),
        lambda h: np.random.default_rng(0).normal(size=4 * h.density_grid_nodes()).astype(np.float32),
        window, every)
    return {"workload": "column_8k_window", "grains": 8192, "window_steps": window,
            "checkpoint_every": every, "forward_ms_per_step": 1e3 * fwd / window,
            "record_ms_per_step": 1e3 * rec / window, "backward_ms_per_step": 1e3 * bwd / window,
            "backward_over_forward": bwd / fwd, "record_plus_backward_over_forward": (rec + bwd) / fwd}


def batched_adjoint_leg(members: int = 16, window: int = 100, every: int = 10) -> dict:
    """The density-loss adjoint over 16 uneven columns (4,096 to 16,384 grains,
    mu 0.3, aspect 2 as the 8,192-grain window) in one batch handle: 100-step
    windows with a checkpoint every sqrt(100) steps, against the same windows
    recorded and differentiated member by member.  Forward and backward are
    timed over the same steps."""
    import numpy as np

    from world_amd import workloads
    from world_amd.sim import Ensemble, ParticleDynamics

    cs = workloads.uneven_ensemble(members, seed=77, n_min=4096, n_max=16384, aspect=(2.0, 2.0),
                                   mu=(0.3, 0.3))
    rng = np.random.default_rng(0)
    with Ensemble(cs) as probe:
        n = probe.n
    fwd, rec, bwd = adjoint_window(
        lambda: Ensemble(cs),
        lambda h: rng.normal(size=members * 4 * h.density_grid_nodes()).astype(np.float32),
        window, every)
    lrec = lbwd = 0.0
    for c in cs:
        _, r, b = adjoint_window(
            lambda: ParticleDynamics(c),
            lambda h: rng.normal(size=4 * h.density_grid_nodes()).astype(np.float32),
            window, every)
        lrec += r
        lbwd += b
    return {"workload": "uneven_adjoint_16", "members": members, "grains": n, "window_steps": window,
            "checkpoint_every": every, "forward_ms_per_step": 1e3 * fwd / window,
            "record_ms_per_step": 1e3 * rec / window, "backward_ms_per_step": 1e3 * bwd / window,
            "backward_over_forward": bwd / fwd,
            "looped_record_plus_backward_ms_per_step": 1e3 * (lrec + lbwd) / window,
            "speedup_over_looped": (lrec + lbwd) / (rec + bwd)}


def sliding_leg() -> dict:
    import numpy as np

    from world_amd import workloads
    from world_amd.sim import ParticleDynamics

    c = workloads.sliding_grain(friction=0.5)
    sim = ParticleDynamics(c)
    t, vx = [], []
    for k in range(180):
        sim.take_step(100)
        t.append((k + 1) * 100 * c.dt)
        vx.append(float(sim.state()[0, 3]))
    sim.close()
    t, vx = np.array(t), np.array(vx)
    sel = vx > 0.1  # fit where the grain still slides
    decel = float(np.polyfit(t[sel], vx[sel], 1)[0])
    return {"deceleration": decel, "analytic": -0.5 * 9.81, "rel_error": abs(decel / (-0.5 * 9.81) - 1)}


def copy_probe(mib: int = 256) -> None:
    """One device copy of known size, the calibration dispatch that HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) are scaled by."""
    import torch

    a = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)  # one elementwise kernel: reads and writes `mib` MiB
    torch.cuda.synchronize()


def traffic_of(kernel: str, n: int, variant: dict):
    """HBM bytes per launch of `kernel` at n grains with the force-pass
    variant the run used, from the committed per-grain counter summary
    (profiles/hbm_traffic.json), or None when it has no such entry."""
    p = ROOT / "profiles" / "hbm_traffic.json"
    if not p.exists():
        return None
    try:
        per = json.loads(p.read_text())["bytes_per_grain"][kernel]
    except (KeyError, ValueError):
        return None
    tag = f"{variant['block']}/{variant['lanes_per_grain']}"
    for e in per:
        if e.get("variant") == tag and abs(e["grains"] - n) <= 0.05 * n:
            return e["bytes_per_grain"] * n
    return None


def cpu_baseline(cfg, target_s: float = 10.0) -> dict:
    from oracle import cpu
    from world_amd import _capi

    threads = min(16, os.cpu_count() or 1)
    st, mat = _capi.initial_state(cfg)
    n = st.size // 6
    # calibrate on the released lattice, then time a bounded sample of steps;
    # steps get dearer as the column collapses and contacts build up, so the
    # sample runs somewhat longer than target_s
    t = time.perf_counter()
    st1, _, _ = cpu.steps(cfg, st, mat, 2, threads=threads)
    per = (time.perf_counter() - t) / 2
    k = int(max(4, min(3000, target_s / max(per, 1e-6))))
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
    ap.add_argument("--profile-steps", type=int, default=-1,
                    help="steps replayed with HIP events around every kernel for the "
                         "roofline (default: the warmup + timed steps of each workload)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the variants, ensemble, adjoint, batched adjoint and sliding legs")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 256 MiB device copy (counter calibration)")
    ap.add_argument("--pin-32k", default="", metavar="BLOCK/LANES",
                    help="fix the 32,768-grain force-pass variant instead of tuning it online "
                         "(counter runs: the counters perturb the timings the tuner reads)")
    ap.add_argument("--pin-1m", default="", metavar="BLOCK/LANES",
                    help="the same for the 1,048,576-grain column")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: measures one GPU only; --gpus must be 1 "
              "(the multi-GPU slab decomposition is designed, not implemented)",
              file=sys.stderr)
        return 2

    from world_amd import workloads

    if args.copy_probe:
        copy_probe()
    cfg = workloads.column_collapse_32k(friction=0.5)
    prof_main = args.warmup + args.steps if args.profile_steps < 0 else args.profile_steps
    def pin_of(text):
        return tuple(int(v) for v in text.split("/")) if text else None

    main_run = run_gpu(cfg, args.steps, args.warmup, 40, True, prof_main, pin_of(args.pin_32k))
    n = main_run["n"]
    value = n * args.steps / main_run["elapsed"]
    nodes = 64 ** 3
    b_total = step_bytes(main_run, nodes, 40)
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
        "roofline": roofline_of(main_run),
        "step_roofline": {"bytes_per_grain_step": b_total,
                          "achieved_GBps": value * b_total / 1e9,
                          "frac": value * b_total / HBM_PEAK},
        "kernels_ms": main_run["kernels_ms"],
        "force_variant": main_run["force_variant"],
        "stable": main_run["stable"],
        "overflow_cells": main_run["overflow"],
    }
    if args.large_steps > 0:
        big = workloads.column_collapse_1m()
        prof_big = 10 + args.large_steps if args.profile_steps < 0 else args.profile_steps
        lr = run_gpu(big, args.large_steps, 10, 100, False, prof_big, pin_of(args.pin_1m))
        lv = lr["n"] * args.large_steps / lr["elapsed"]
        lb = step_bytes(lr, 128 ** 3, 100)
        out["large"] = {
            "workload": "column_collapse_1m", "grains": lr["n"], "steps": args.large_steps,
            "value": lv, "unit": "grain-steps/s", "ms_per_step": 1e3 * lr["elapsed"] / args.large_steps,
            "step_roofline": {"bytes_per_grain_step": lb, "achieved_GBps": lv * lb / 1e9,
                              "frac": lv * lb / HBM_PEAK},
            "roofline": roofline_of(lr),
            "kernels_ms": lr["kernels_ms"], "force_variant": lr["force_variant"], "stable": lr["stable"],
            "overflow_cells": lr["overflow"],
        }
    if not args.no_extras:
        variants = {}
        for name, mu, picard in (("sie_mu0", 0.0, False), ("picard3_mu0.5", 0.5, True),
                                 ("picard3_mu0", 0.0, True)):
            vr = run_gpu(workloads.column_collapse_32k(friction=mu, picard=picard), args.steps,
                         args.warmup, 40, True, 10)
            variants[name] = {"value": vr["n"] * args.steps / vr["elapsed"], "unit": "grain-steps/s",
                              "ms_per_step": 1e3 * vr["elapsed"] / args.steps, "stable": vr["stable"]}
        out["variants"] = variants
        out["ensemble"] = ensemble_leg(args.ensemble_steps)
        out["adjoint"] = adjoint_leg()
        out["batched_adjoint"] = batched_adjoint_leg()
        out["sliding"] = sliding_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
