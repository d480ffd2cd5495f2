#!/usr/bin/env python3
"""Grain-steps/s of the per-step granular pipeline on one MI355X.

Headline (`value`): the 32,768-grain (32^3 jittered lattice) column collapse,
mu = 0.5, tangential spring with history, semi-implicit Euler at dt = 1e-3,
the total energy recorded every step (fused into the force pass) and a 64^3
cloud-in-cell deposit after every step whose absolute number is a multiple of
40 (kept on the device).  One "step" is one take_step of the whole pipeline:
a fused step (the force pass, which also counts its grains into the next
step's cells; the cell scan, whose last blocks reduce the pass's partials;
the reorder) or, for ensembles, the classic six-launch sequence.
Exactly --steps K steps are timed, in a window in the middle of the
configured 5,000-step run: the handle is first advanced untimed to step
S0 = 2480 - ceil(min(K, 40) / 2) (2480 is the multiple of 40 nearest step
2,500, so the window holds at least one deposit; later if --warmup W asks for
more untimed steps), the force pass's variant is tuned online before S0 and
fixed for the window, and the step graphs are captured before the clock
starts.
Grains are synthetic (a seeded lattice) and resident in HBM.

Secondary entries of the same JSON object:
  column_32k_runs  the configured 5,000-step runs of that column: mu 0.5 and
            0, semi-implicit Euler and backward-Euler-Picard (3 iterations),
            from the released lattice, online tuning included;
  column_1m  the 1,048,576-grain column (128 x 128 x 64, 4 x 4 m box, 128^3
            CIC every 100 steps): a window of whole deposit periods around
            step 5,000, and the configured 10,000-step run;
  uneven_ensemble_64  64 uneven members (2,048..32,768 grains, aspect 1..4,
            mu in [0, 0.6], 50 um Gaussian perturbation), 2,000 steps, one
            batched handle against looping the members one at a time;
  adjoint   the 8,192-grain window (mu 0.3, 200 steps, checkpoints every 20):
            a backward step against a forward step;
  batched_adjoint  16 uneven members (4,096..16,384 grains) of that adjoint
            in one handle, 100-step windows, checkpoints every 10 steps,
            against the same windows member by member;
  sliding   one grain launched at 1 m/s on the floor at mu 0.5: fitted
            deceleration against -mu g = -4.905 m/s^2;
  cpu_baseline  the CPU oracle (the reference has no CPU path) on the
            host's cores from the GPU's state and springs at S0: the same
            window, extended to about ten seconds of work, plus one thread;
            with column_1m, also the 1M window (column_1m.cpu_baseline).

`roofline` is the force pass of the headline window: algorithmic bytes per
grain (56 B of state and cell key, 8 B of the fused count (the cell
counter's increment and the grain's bucket entry or slot), plus 8 B of
spring counts and 32 B per contact for the stored springs; Z, the mean
contact count, is read from the device counter over the window) times grains
over its mean launch time,
measured by replaying the same window on a second handle with HIP events
stamped by the kernels on the engine's stream.  `traffic` is the force
pass's HBM bytes per launch over the same window, from the committed
rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this command
(profiles/hbm_traffic.json), or null when it holds no entry for this window.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs LIST]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
# Algorithmic bytes per grain of the force pass: read state 24 + material 4,
# write state 24 (the force + integrate pass) and the next cell key 4.
B_FORCE = 56.0
# A fused step's force pass also counts its grains for the next bin: the cell
# counter's increment (4 B) and the grain's bucket entry or slot (4 B).
B_COUNT = 8.0
# Algorithmic bytes per grain-step of the whole semi-implicit Euler step: cell
# key (read position 12 + write key 4), stable reorder (read key 4 + state 24
# + material 4, write state 24 + material 4 + permutation 4), fused force +
# integrate (read state 24 + material 4, write state 24).
B_STEP = 132.0
LEGS = ("main", "runs32k", "large", "ensemble", "adjoint", "batched_adjoint", "sliding", "cpu")
# One online tuning: three rounds over the six force-pass variants, then (fused
# steps) the six (scan tile, binning) pairs twice with the chosen variant.
TUNE_STEPS = 30
CAPTURE_STEPS = 31  # 16 + 8 + 4 + 2 + 1: every step graph the engine replays


def progress(msg: str) -> None:
    """One line on stderr per finished leg (stdout carries only the JSON)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def history_bytes(z: float) -> float:
    """Spring counts (4 B read + 4 B written) and 16 B per contact each way."""
    return 8.0 + 32.0 * z


def cic_bytes_per_grain_step(n: int, nodes: int, every: int) -> float:
    # read position 12 + velocity 12 + material 4 per grain, write 16 B per
    # node, once every `every` steps
    return (28.0 + 16.0 * nodes / n) / every


def window_start(center: int, k: int, every: int, warmup: int) -> int:
    """Start s0 of a k-step window around `center` that holds a deposit:
    steps s0+1 .. s0+k straddle the multiple of `every` nearest the centre
    (the first one at least warmup + ceil(min(k, every) / 2) steps in)."""
    half = -(-min(k, every) // 2)
    c = every * int(round(center / every))
    while c - half < warmup:
        c += every
    return c - half


def deposit_steps(s0: int, k: int, every: int) -> list[int]:
    """Absolute step numbers in (s0, s0 + k] after which a deposit is due."""
    first = (s0 // every + 1) * every
    return list(range(first, s0 + k + 1, every))


def run_segment(sim, s0: int, k: int, every: int) -> int:
    """Advances k steps from absolute step s0, recording a CIC deposit after
    every step whose number is a multiple of `every`.  Returns the number
    of deposits."""
    s = s0
    deps = deposit_steps(s0, k, every)
    for d in deps:
        if d > s:
            sim.take_step(d - s)
        sim.record_density_grid()
        s = d
    if s0 + k > s:
        sim.take_step(s0 + k - s)
    return len(deps)


def advance(sim, steps: int, chunk: int = 64) -> None:
    """Untimed steps, waiting for the device every `chunk` steps so that few
    launches are ever queued at once (a profiler attached to the process
    records every queued dispatch)."""
    while steps > 0:
        k = min(chunk, steps)
        sim.take_step(k)
        sim.sync()
        steps -= k


def prepare(sim, s0: int, pin=None) -> tuple[int, int]:
    """Advances untimed to step s0 and leaves the force pass's variant fixed
    and the step and deposit graphs captured: one online tuning round ends
    CAPTURE_STEPS before s0, its choice is pinned, the last CAPTURE_STEPS
    steps capture the step graphs the window replays, and one deposit into
    the density log's spare slot captures the deposit's graph."""
    if pin:
        sim.set_force_variant(*pin)
        advance(sim, s0)
        sim.record_density_grid()  # captures the deposit's graph (the log's spare slot)
        sim.sync()
        return tuple(pin)
    pre = s0 - TUNE_STEPS - CAPTURE_STEPS
    if pre < 0:
        raise ValueError(f"window start {s0} leaves no room for tuning before it")
    advance(sim, pre)
    sim.set_force_variant(0, 0)  # tune on the next call
    sim.take_step(TUNE_STEPS)
    v = sim.force_variant()
    sim.set_force_variant(*v)
    sim.take_step(CAPTURE_STEPS)
    sim.record_density_grid()  # captures the deposit's graph (the log's spare slot)
    sim.sync()
    return v


def window(cfg, k: int, s0: int, every: int, energy: bool, pin=None, profile: bool = True,
           keep_state: bool = False) -> dict:
    """Times steps s0+1 .. s0+k of cfg's run (deposits at their absolute
    cadence) after prepare(); then replays the same window on a second handle
    with kernel-stamped events for the per-kernel means."""
    from world_amd.sim import ParticleDynamics

    sim = ParticleDynamics(cfg)
    n = sim.n
    if energy:
        sim.enable_energy_log(s0 + k + 64)
    sim.enable_density_log(len(deposit_steps(s0, k, every)) + 1)
    variant = prepare(sim, s0, pin)
    tile = sim.scan_tile()
    binning = sim.binning()
    passes0, contacts0 = sim.contact_stats()
    hist = cfg.tangential_model == 1
    start = None
    if keep_state:
        start = {"state": sim.state(), "history": sim.history() if hist else None}
    sim.sync()
    t0 = time.perf_counter()
    ndep = run_segment(sim, s0, k, every)
    sim.sync()
    elapsed = time.perf_counter() - t0
    passes1, contacts1 = sim.contact_stats()
    z = (contacts1 - contacts0) / max(1, (passes1 - passes0) * n)
    out = {
        "n": n, "elapsed": elapsed, "s0": s0, "k": k, "deposits": ndep,
        "stable": sim.is_stable(), "overflow": sim.grid_overflow_count(),
        "variant": variant, "scan_tile": tile, "binning": binning, "contacts_per_grain": z, "history": hist,
        "eager": cfg.kernel_timing != 0,
        "fused": sim.fused_steps(),
        "tuned": sim.force_variant() == tuple(variant),
    }
    if keep_state:
        out["start"] = start
        out["end_state"] = sim.state()
    sim.close()
    out["kernels_ms"] = {}
    if profile:
        prof = ParticleDynamics(cfg)
        prof.set_force_variant(*variant)
        prof.set_scan_tile(tile)
        prof.set_binning(binning)
        if energy:
            prof.enable_energy_log(s0 + k + 64)
        advance(prof, s0)
        out["kernels_ms"] = prof.profile_step(k)
        prof.close()
    return out


def full_run(cfg, steps: int, every: int, energy: bool) -> dict:
    """The configured run from the released lattice, as a user runs it:
    online tuning and graph captures included."""
    from world_amd.sim import ParticleDynamics

    sim = ParticleDynamics(cfg)
    if energy:
        sim.enable_energy_log(steps + 64)
    sim.enable_density_log(steps // every + 1)
    sim.sync()
    t0 = time.perf_counter()
    ndep = run_segment(sim, 0, steps, every)
    sim.sync()
    elapsed = time.perf_counter() - t0
    passes, contacts = sim.contact_stats()
    out = {"steps": steps, "deposits": ndep, "value": sim.n * steps / elapsed,
           "unit": "grain-steps/s", "ms_per_step": 1e3 * elapsed / steps, "stable": sim.is_stable(),
           "overflow_cells": sim.grid_overflow_count(),
           "contacts_per_grain": contacts / max(1, passes * sim.n),
           "force_variant_at_end": list(sim.force_variant()), "includes_online_tuning": True}
    sim.close()
    return out


def traffic_of(workload: str, s0: int, k: int, variant) -> float | None:
    """HBM bytes per launch of the force pass over steps s0+1 .. s0+k of
    `workload` with `variant`, from the committed counter summary, or None
    when it holds no entry for exactly that window."""
    p = ROOT / "profiles" / "hbm_traffic.json"
    if not p.exists():
        return None
    try:
        entries = json.loads(p.read_text())["force_pass_windows"]
    except (KeyError, ValueError):
        return None
    tag = f"{variant[0]}/{variant[1]}"
    for e in entries:
        if (e.get("workload") == workload and e.get("first_step") == s0 + 1 and
                e.get("last_step") == s0 + k and e.get("variant") == tag):
            return float(e["bytes_per_launch"])
    return None


def roofline_of(run: dict, workload: str) -> dict:
    n = run["n"]
    ms = run["kernels_ms"].get("k_force", 0.0)
    bpg = B_FORCE + (B_COUNT if run["fused"] else 0.0)
    bpg += history_bytes(run["contacts_per_grain"]) if run["history"] else 0.0
    achieved = bpg * n / (ms * 1e-3) if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": "k_force", "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic_of(workload, run["s0"], run["k"], run["variant"]),
            "bytes_per_grain": bpg, "contacts_per_grain": run["contacts_per_grain"],
            "avg_launch_ms": ms}


def step_roofline(run: dict, nodes: int, every: int, value: float) -> dict:
    b = B_STEP + cic_bytes_per_grain_step(run["n"], nodes, every)
    if run["history"]:
        b += history_bytes(run["contacts_per_grain"])
    return {"bytes_per_grain_step": b, "achieved_GBps": value * b / 1e9, "frac": value * b / HBM_PEAK}


def window_config(run: dict, every: int) -> dict:
    return {"first_step": run["s0"] + 1, "last_step": run["s0"] + run["k"],
            "untimed_steps_before": run["s0"], "deposits_in_window": run["deposits"],
            "deposit_every": every, "force_variant": {"block": run["variant"][0],
                                                      "lanes_per_grain": run["variant"][1]},
            "scan_tile_cells": run["scan_tile"],
            "binning": "buckets" if run["binning"] == 2 else "scatter",
            "tuned_before_window": True, "launch": "eager" if run["eager"] else "graphs captured before the window",
            "step_sequence": "fused: force + count, scan + reduction, reorder" if run["fused"] else "classic: six launches"}


def runs32k_leg(steps: int = 5000) -> dict:
    from world_amd import workloads

    out = {}
    for name, mu, picard in (("sie_mu0.5", 0.5, False), ("sie_mu0", 0.0, False),
                             ("picard3_mu0.5", 0.5, True), ("picard3_mu0", 0.0, True)):
        out[name] = full_run(workloads.column_collapse_32k(friction=mu, picard=picard), steps, 40, True)
    return out


def large_leg(k: int, warmup: int, pin, full_steps: int, eager: bool = False,
              cpu_seconds: float | None = None) -> dict:
    """The 1,048,576-grain column: a window of whole deposit periods (at
    least k steps) around step 5,000, and the configured run; with
    cpu_seconds, the CPU oracle over the same window (from the GPU's state and
    springs at its start, its end state checked against the GPU's)."""
    from world_amd import workloads

    cfg = workloads.column_collapse_1m()
    if eager:
        cfg.kernel_timing = 1
    kl = 100 * max(1, -(-k // 100))
    s0 = window_start(5000, kl, 100, warmup)
    r = window(cfg, kl, s0, 100, False, pin, keep_state=cpu_seconds is not None)
    v = r["n"] * kl / r["elapsed"]
    out = {"workload": "column_collapse_1m", "grains": r["n"], "steps": kl, "value": v,
           "unit": "grain-steps/s", "ms_per_step": 1e3 * r["elapsed"] / kl,
           "window": window_config(r, 100), "roofline": roofline_of(r, "column_collapse_1m"),
           "step_roofline": step_roofline(r, 128 ** 3, 100, v), "kernels_ms": r["kernels_ms"],
           "stable": r["stable"], "overflow_cells": r["overflow"]}
    out["cell_scale_tuned"] = cell_scale_window(cfg, kl, s0, 100, r["variant"])
    if full_steps > 0:
        out["full_run"] = full_run(cfg, full_steps, 100, False)
    if cpu_seconds is not None:
        out["cpu_baseline"] = cpu_baseline(cfg, r, cpu_seconds)
    return out


def cell_scale_window(cfg, k: int, s0: int, every: int, variant) -> dict:
    """The same window with the opt-in coarser collision cells: the engine
    times the scales 1, 1.25 and 1.5 on real steps just before the window
    (6 steps each) and keeps the fastest; results then match the FP32
    arithmetic on that grid, not on the configured one."""
    from world_amd.sim import ParticleDynamics

    sim = ParticleDynamics(cfg)
    sim.set_force_variant(*variant)
    advance(sim, s0 - 18 - CAPTURE_STEPS)
    chosen = sim.tune_cell_scale(6)
    sim.take_step(CAPTURE_STEPS)
    sim.enable_density_log(len(deposit_steps(s0, k, every)) + 1)
    sim.record_density_grid()
    sim.sync()
    t0 = time.perf_counter()
    run_segment(sim, s0, k, every)
    sim.sync()
    elapsed = time.perf_counter() - t0
    scale, grid = sim.cell_scale()
    out = {"first_step": s0 + 1, "last_step": s0 + k, "cell_scale": scale, "collision_grid": list(grid),
           "ms_per_step": 1e3 * elapsed / k, "value": sim.n * k / elapsed, "stable": sim.is_stable(),
           "overflow_cells": sim.grid_overflow_count()}
    sim.close()
    return out


def ensemble_leg(steps: int) -> dict:
    """64-member uneven ensemble over its configured run: one batched handle
    vs the members looped, both from the released columns."""
    from world_amd import _capi, workloads
    from world_amd.sim import Ensemble, ParticleDynamics

    cs = workloads.uneven_ensemble(64)
    n = sum(_capi.grain_count(c) for c in cs)
    ens = Ensemble(cs)
    ens.sync()
    t = time.perf_counter()
    ens.take_step(steps)
    ens.sync()
    tb = time.perf_counter() - t
    stable = bool(ens.member_stable().all())
    ens.close()
    tl = 0.0
    for c in cs:
        sim = ParticleDynamics(c)
        sim.sync()
        t = time.perf_counter()
        sim.take_step(steps)
        sim.sync()
        tl += time.perf_counter() - t
        sim.close()
    return {"workload": "uneven_ensemble_64", "members": len(cs), "grains": n, "steps": steps,
            "batched_ms_per_step": 1e3 * tb / steps, "batched_grain_steps_per_s": n * steps / tb,
            "looped_ms_per_step": 1e3 * tl / steps, "looped_grain_steps_per_s": n * steps / tl,
            "speedup": tl / tb, "stable": stable, "includes_online_tuning": True}


def adjoint_window(make, cot_of, window_steps: int, every: int, pin=None):
    """Forward, record and backward cost of one window, all over the same
    steps, with the same tuned choices and the graphs captured beforehand.  A
    handle tunes (TUNE_STEPS steps), pins its force variant, scan tile and
    binning, captures the step graphs (31 steps) and runs one window untimed,
    then times a forward step over the next window; a second handle takes the
    same choices (so the record and the recomputed steps run the forward's
    variant), runs the same steps, records and differentiates its first window
    untimed (allocations, adjoint graph captures) and then the same second
    window.  Returns seconds (forward, record, backward) for the window."""
    with make() as a:
        if pin:
            a.set_force_variant(*pin)  # no tuning: the default scan tile and binning
        a.take_step(TUNE_STEPS)
        choice = (a.force_variant(), a.scan_tile(), a.binning())
        a.set_force_variant(*choice[0])
        a.take_step(CAPTURE_STEPS)
        a.take_step(window_steps)
        a.sync()
        t = time.perf_counter()
        a.take_step(window_steps)
        a.sync()
        fwd = time.perf_counter() - t
    with make() as h:
        cot = cot_of(h)
        h.set_force_variant(*choice[0])
        h.set_scan_tile(choice[1])
        h.set_binning(choice[2])
        h.take_step(TUNE_STEPS + CAPTURE_STEPS)
        h.sync()
        rec = bwd = 0.0
        for _ in range(2):
            t = time.perf_counter()
            h.adjoint_record(window_steps, every)
            h.sync()
            rec = time.perf_counter() - t
            t = time.perf_counter()
            h.adjoint_backward(cot)
            bwd = time.perf_counter() - t
    return fwd, rec, bwd


def adjoint_leg(window_steps: int = 200, every: int = 20, pin=None) -> dict:
    """Backward-step cost of the 8,192-grain density-loss window against a
    forward step over the same window."""
    import numpy as np

    from world_amd import workloads
    from world_amd.sim import ParticleDynamics

    c = workloads.column_8k(friction=0.3)
    fwd, rec, bwd = adjoint_window(
        lambda: ParticleDynamics(c),
        lambda h: np.random.default_rng(0).normal(size=4 * h.density_grid_nodes()).astype(np.float32),
        window_steps, every, pin)
    return {"workload": "column_8k_window", "grains": 8192, "window_steps": window_steps,
            "checkpoint_every": every, "forward_ms_per_step": 1e3 * fwd / window_steps,
            "record_ms_per_step": 1e3 * rec / window_steps,
            "backward_ms_per_step": 1e3 * bwd / window_steps,
            "backward_over_forward": bwd / fwd, "record_plus_backward_over_forward": (rec + bwd) / fwd}


def batched_adjoint_leg(members: int = 16, window_steps: int = 100, every: int = 10) -> dict:
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
        window_steps, every)
    lrec = lbwd = 0.0
    for c in cs:
        _, r, b = adjoint_window(
            lambda: ParticleDynamics(c),
            lambda h: rng.normal(size=4 * h.density_grid_nodes()).astype(np.float32),
            window_steps, every)
        lrec += r
        lbwd += b
    return {"workload": "uneven_adjoint_16", "members": members, "grains": n,
            "window_steps": window_steps, "checkpoint_every": every,
            "forward_ms_per_step": 1e3 * fwd / window_steps,
            "record_ms_per_step": 1e3 * rec / window_steps,
            "backward_ms_per_step": 1e3 * bwd / window_steps,
            "backward_over_forward": bwd / fwd,
            "looped_record_plus_backward_ms_per_step": 1e3 * (lrec + lbwd) / window_steps,
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


# The saved 1M window state lives in the temporary directory (24 MiB, not part of the tree).
WINDOW_STATE = Path(os.environ.get("TMPDIR") or "/tmp") / "world_amd_column_1m_step4950.npy"


def save_window_state(steps: int = 4950) -> None:
    """Advances the 1,048,576-grain column `steps` steps and saves its grain
    state (original order), so that a short process under a counter
    profiler can start from the 1M window (rocprofv3 on this image fails
    in long or graph-replaying runs)."""
    import numpy as np
    from world_amd import workloads
    from world_amd.sim import ParticleDynamics

    with ParticleDynamics(workloads.column_collapse_1m()) as sim:
        advance(sim, steps, 256)
        WINDOW_STATE.parent.mkdir(parents=True, exist_ok=True)
        np.save(WINDOW_STATE, sim.state())


def restored_steps(k: int = 20) -> None:
    """k steps of the 1M column from the saved window state, launched
    eagerly (no graph replay), with the force variant, scan tile and binning
    the tuner keeps in the window pinned (one lane per grain, 4,096-cell
    tiles, scatter binning).  The tangential springs are not part of the
    saved state: they restart at zero and are rebuilt in the first step, so
    the later of the k launches are the ones to read."""
    import numpy as np
    from world_amd import _capi, workloads
    from world_amd.sim import ParticleDynamics

    cfg = workloads.column_collapse_1m()
    with ParticleDynamics(cfg) as sim:
        sim.set_graph_replay(False)
        sim.set_force_variant(64, 1)
        sim.set_scan_tile(4096)
        sim.set_binning(_capi.WD_BINNING_SCATTER)
        sim.set_state(np.load(WINDOW_STATE))
        for _ in range(k):
            sim.take_step(1)
        sim.sync()


def host_cpu() -> dict:
    """The host's processor and the threads this process may use: the CPU
    set it is bound to, capped by OMP_NUM_THREADS where that is set (the
    host's share for one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    model = ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"threads": threads, "affinity_cpus": avail, "cpu_count": os.cpu_count(),
            "omp_num_threads": omp or None, "model": model}


def cpu_baseline(cfg, run: dict, target_s: float = 10.0) -> dict:
    """The CPU oracle from the GPU's state and springs at the window's start:
    the window itself on every thread this process may use (its end state
    checked bit for bit against the GPU's), extended to about target_s of
    work, and a sample of about target_s on one thread."""
    import numpy as np

    from oracle import cpu
    from world_amd import _capi

    host = host_cpu()
    threads = host["threads"]
    _, mat = _capi.initial_state(cfg)
    st0, hist0 = run["start"]["state"], run["start"]["history"]
    n, k, s0 = run["n"], run["k"], run["s0"]

    def make():
        s = cpu.Sim(cfg, st0, mat)
        if hist0 is not None:
            s.set_history(*hist0)
        return s

    sim = make()
    t = time.perf_counter()
    sim.steps(k, threads)
    tk = time.perf_counter() - t
    same = bool(np.array_equal(sim.state(), run["end_state"]))
    extra = max(0, int(target_s / max(tk / k, 1e-9)) - k)
    extra = min(extra, 4000)
    t = time.perf_counter()
    if extra:
        sim.steps(extra, threads)
    te = time.perf_counter() - t
    sim.close()
    steps_all = k + extra
    rate_all = n * steps_all / (tk + te)
    one = make()
    one.steps(1, 1)  # calibration
    t = time.perf_counter()
    one.steps(1, 1)
    per1 = time.perf_counter() - t
    m = max(1, min(k, int(target_s / max(per1, 1e-9))))
    t = time.perf_counter()
    one.steps(m, 1)
    t1 = time.perf_counter() - t
    one.close()
    return {"value": rate_all, "unit": "grain-steps/s", "cores": threads, "kind": "port",
            "sample": f"steps {s0 + 1}..{s0 + steps_all} of the {n}-grain workload on {threads} "
                      f"threads (OpenMP over grains), from the GPU's state and springs at step {s0}: "
                      f"the timed GPU window and {extra} steps beyond it; FP32 oracle "
                      f"(oracle/oracle.cpp, g++ -O2)",
            "window_bit_exact_vs_gpu": same,
            "one_thread": {"value": n * m / t1, "unit": "grain-steps/s",
                           "sample": f"steps {s0 + 3}..{s0 + 2 + m} on one thread"},
            "host": host}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of the headline window")
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed steps at least before the headline window")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated subset of " + ",".join(LEGS))
    ap.add_argument("--large-full-steps", type=int, default=10000,
                    help="steps of the configured 1,048,576-grain run (0 skips it)")
    ap.add_argument("--runs32k-steps", type=int, default=5000)
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the event-timed replay of each window (no per-kernel means)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel of the headline and 1M legs eagerly (the engine's "
                         "kernel-timing mode) instead of replaying step graphs: for profiler runs, "
                         "since rocprofv3 has crashed following long runs of graph launches")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 256 MiB device copy (counter calibration)")
    ap.add_argument("--pin-32k", default="", metavar="BLOCK/LANES",
                    help="fix the 32,768-grain force-pass variant instead of tuning it before the window")
    ap.add_argument("--pin-8k", default="", metavar="BLOCK/LANES",
                    help="the same for the 8,192-grain adjoint window (for profiler runs, whose "
                         "serialized dispatches mislead the tuner)")
    ap.add_argument("--pin-1m", default="", metavar="BLOCK/LANES",
                    help="the same for the 1,048,576-grain column")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: measures one GPU only; --gpus must be 1 "
              "(the multi-GPU slab decomposition is designed, not implemented)",
              file=sys.stderr)
        return 2
    legs = set(v for v in args.legs.split(",") if v)
    unknown = legs - set(LEGS)
    if unknown or args.steps < 1 or args.warmup < 0:
        print(f"bench.py: bad --legs {sorted(unknown)} or --steps / --warmup", file=sys.stderr)
        return 2

    from world_amd import workloads

    def pin_of(text):
        return tuple(int(v) for v in text.split("/")) if text else None

    if args.copy_probe:
        copy_probe()
    k = args.steps
    cfg = workloads.column_collapse_32k(friction=0.5)
    if args.eager:
        cfg.kernel_timing = 1
    s0 = window_start(2500, k, 40, args.warmup)
    out = {"steps": k, "warmup": args.warmup}
    main_run = None
    if "main" in legs:
        main_run = window(cfg, k, s0, 40, True, pin_of(args.pin_32k), profile=not args.no_profile,
                          keep_state="cpu" in legs)
        n = main_run["n"]
        value = n * k / main_run["elapsed"]
        progress(f"headline window steps {s0 + 1}..{s0 + k}: {1e3 * main_run['elapsed'] / k:.4f} ms/step")
        cfg_block = {"workload": "column_collapse_32k", "grains": n, "integrator": "semi_implicit_euler",
                     "friction": 0.5, "tangential_model": "spring_history", "dt": 1e-3,
                     "cic_nodes": "64^3", "cic_every": 40, "energy_every": 1, "parallelism": "single_gpu"}
        cfg_block.update(window_config(main_run, 40))
        out = {
            "metric": "grain-steps/s (32,768-grain column collapse, SIE, mu 0.5, energy every step, "
                      "64^3 CIC every 40 steps)",
            "value": value,
            "unit": "grain-steps/s",
            "n_gpus": 1,
            "steps": k,
            "warmup": args.warmup,
            "ms_per_step": 1e3 * main_run["elapsed"] / k,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "fp32",
            "data": "synthetic (seeded jittered lattice)",
            "config": cfg_block,
            "launch": "eager" if args.eager else "graphs",
            "roofline": roofline_of(main_run, "column_collapse_32k"),
            "step_roofline": step_roofline(main_run, 64 ** 3, 40, value),
            "kernels_ms": main_run["kernels_ms"],
            "stable": main_run["stable"],
            "overflow_cells": main_run["overflow"],
        }
    if "runs32k" in legs:
        out["column_32k_runs"] = runs32k_leg(args.runs32k_steps)
        progress("column_32k_runs")
    if "large" in legs:
        out["column_1m"] = large_leg(k, args.warmup, pin_of(args.pin_1m), args.large_full_steps,
                                     args.eager, args.cpu_seconds if "cpu" in legs else None)
        progress("column_1m")
    if "ensemble" in legs:
        out["uneven_ensemble_64"] = ensemble_leg(args.ensemble_steps)
        progress("uneven_ensemble_64")
    if "adjoint" in legs:
        out["adjoint"] = adjoint_leg(pin=pin_of(args.pin_8k))
        progress("adjoint")
    if "batched_adjoint" in legs:
        out["batched_adjoint"] = batched_adjoint_leg()
        progress("batched_adjoint")
    if "sliding" in legs:
        out["sliding"] = sliding_leg()
        progress("sliding")
    if "cpu" in legs and main_run is not None:
        out["cpu_baseline"] = cpu_baseline(cfg, main_run, args.cpu_seconds)
        progress("cpu_baseline")
    progress("done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
