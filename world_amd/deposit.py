"""Deposit left by a granular column collapse: runout radius and height.

Restates the reduction of the reference's deposit script
(scripts/measure_deposit.cjs:42-62): for every grain the radial extent
hypot(x, y) from the z axis and the height z - z_min above the domain's
lower bound, each sorted, and the quantile q taken as the element at index
round(q (n - 1)) clamped to [0, n - 1], with JavaScript's round (halves up).
The script measures once the broadcast server's simulated time, advanced
steps_per_frame steps per "run" request and accumulated in FP32 one dt per
step (src/particle_dynamics.cu:280), first reaches the target time
(scripts/measure_deposit.cjs:38-40): frames_until() gives that step count.

The reference publishes these quantities for its demo column
(demos/sand.yaml, mu 0.5, and demos/fluid.yaml, mu 0) at t = 2.4 s, measured
more than once: MEASUREMENTS lists every published value with its source.
The frictional deposit was measured at dt = 4e-4 with 100 steps per frame
(the demo files) and again at dt = 1/2520 with 42 steps per frame (the video
configuration, which the reference describes as the same physics,
CLAUDE.md:1264-1270).
"""
from __future__ import annotations

import math

import numpy as np

# Every published median at t = 2.4 s: (friction, dt, steps per frame, key,
# value in metres, source).
MEASUREMENTS = [
    (0.5, 4e-4, 100, "radius_p50", 0.550, "CLAUDE.md:1326-1328"),
    (0.5, 4e-4, 100, "height_p50", 0.042, "CLAUDE.md:1326-1328"),
    (0.5, 4e-4, 100, "height_p50", 0.043, "README.md:27-28"),
    (0.0, 4e-4, 100, "radius_p50", 0.872, "CLAUDE.md:1326-1328"),
    (0.0, 4e-4, 100, "height_p50", 0.026, "CLAUDE.md:1326-1328"),
    (0.5, 1.0 / 2520.0, 42, "radius_p50", 0.540, "CLAUDE.md:1268-1270"),
    (0.5, 1.0 / 2520.0, 42, "height_p50", 0.043, "CLAUDE.md:1268-1270"),
    (0.0, 1.0 / 2520.0, 42, "radius_p50", 0.872, "CLAUDE.md:1268-1270"),
    (0.0, 1.0 / 2520.0, 42, "height_p50", 0.026, "CLAUDE.md:1268-1270"),
]
PUBLISHED_DECIMALS = 3
# The two time-step configurations the measurements were taken with.
CONFIGS = [(4e-4, 100), (1.0 / 2520.0, 42)]


def published(friction: float, key: str) -> list[float]:
    return [m[4] for m in MEASUREMENTS if m[0] == friction and m[3] == key]


def quantile(sorted_values: np.ndarray, q: float) -> float:
    """The script's quantile of an ascending array."""
    n = len(sorted_values)
    idx = min(n - 1, max(0, int(math.floor(q * (n - 1) + 0.5))))
    return float(sorted_values[idx])


def extent(state: np.ndarray, z_min: float) -> dict[str, float]:
    """p50, p99 and max of the radial extent and of the height above z_min,
    from positions (the first three columns of an [n, 6] state or an [n, 3]
    array), reduced in double as the script does."""
    p = np.asarray(state, np.float32).reshape(len(state), -1)[:, :3].astype(np.float64)
    radius = np.sort(np.hypot(p[:, 0], p[:, 1]))
    height = np.sort(p[:, 2] - float(np.float32(z_min)))
    out = {}
    for name, v in (("radius", radius), ("height", height)):
        out[f"{name}_p50"] = quantile(v, 0.50)
        out[f"{name}_p99"] = quantile(v, 0.99)
        out[f"{name}_max"] = float(v[-1])
    return out


def frames_until(dt: float, target: float, steps_per_frame: int) -> int:
    """Steps the script lets run: whole frames of steps_per_frame steps until
    the FP32 simulated time is at least `target`."""
    t = np.float32(0.0)
    d = np.float32(dt)
    steps = 0
    while True:
        for _ in range(steps_per_frame):
            t = np.float32(t + d)
        steps += steps_per_frame
        if t >= np.float32(target):
            return steps
