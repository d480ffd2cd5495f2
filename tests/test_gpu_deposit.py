"""Bulk observables of the column collapse against the reference's published
deposit: the median radial extent and the median height above the floor at
t = 2.4 s of its demo column (demos/sand.yaml, mu 0.5; demos/fluid.yaml,
mu 0), reduced as scripts/measure_deposit.cjs:42-62 does.

The reference measured the frictional deposit twice, with the demo's dt 4e-4
and with its video configuration's dt 1/2520, "the same physics"
(CLAUDE.md:1264-1270): radius 0.550 and 0.540 m, height 0.042 and 0.043 m (and
0.043 m in README.md:27-28); the frictionless one gave 0.872 / 0.026 m both
times.  The collapse is chaotic, so each of those is one realization.  Here
both configurations are run as ensembles of distinct realizations: every
grain offset uniformly by up to a quarter radius (the jitter's magnitude;
the reference's chaos configuration uses it to measure "the scatter between
distinct realizations of the same macroscopic state", chaos/ensemble.yaml:
92-97), perturbation seeds 1000 + e as the chaos driver seeds its members,
member 0 unperturbed.

Agreement: the mean of the published values and the mean of the ensemble
members differ by at most three standard deviations of one realization, the
two sides' sample scatters combined in quadrature (the published values' own
spread is the reference's run-to-run scatter), plus half a unit in the
published last digit.  Each configuration's ensemble against its own published values is
printed beside it."""
import math

import numpy as np
import pytest

from world_amd import deposit, workloads
from world_amd.sim import Ensemble

pytestmark = pytest.mark.gpu

MEMBERS = 12


def run_ensemble(mu, dt, per_frame):
    cs = workloads.demo_ensemble(mu, MEMBERS, dt=dt)
    steps = deposit.frames_until(dt, 2.4, per_frame)
    with Ensemble(cs) as ens:
        ens.take_step(steps)
        assert ens.time >= np.float32(2.4)
        states = ens.member_states()
        assert ens.member_stable().all()
        assert (ens.member_overflow_counts() == 0).all()
    return [deposit.extent(s, cs[0].z_min) for s in states]


@pytest.mark.parametrize("mu", [0.5, 0.0], ids=["sand", "fluid"])
def test_demo_deposit_within_measured_scatter(mu):
    runs = {cfg: run_ensemble(mu, *cfg) for cfg in deposit.CONFIGS}
    half_unit = 0.5 * 10.0 ** -deposit.PUBLISHED_DECIMALS
    for key in ("radius_p50", "height_p50"):
        ours = np.array([e[key] for ex in runs.values() for e in ex])
        ref = np.array(deposit.published(mu, key))
        sd_ours = float(ours.std(ddof=1))
        sd_ref = float(ref.std(ddof=1)) if len(ref) > 1 else 0.0
        sd_diff = math.sqrt(sd_ours ** 2 + sd_ref ** 2)
        gap = abs(float(ours.mean()) - float(ref.mean()))
        tol = 3.0 * sd_diff + half_unit
        for (dt, per_frame), ex in runs.items():
            v = np.array([e[key] for e in ex])
            pub = [m[4] for m in deposit.MEASUREMENTS
                   if m[0] == mu and m[1] == dt and m[3] == key]
            print(f"mu {mu} dt {dt:.3e} {key}: published {pub}, ensemble mean {v.mean():.4f} "
                  f"sd {v.std(ddof=1):.4f} [{v.min():.4f}, {v.max():.4f}]", flush=True)
        print(f"mu {mu} {key}: published mean {ref.mean():.4f} (sd {sd_ref:.4f}), ours "
              f"{ours.mean():.4f} (sd {sd_ours:.4f}), gap {gap:.4f}, tolerance {tol:.4f}", flush=True)
        assert sd_ours > 0.0  # the members did decorrelate
        assert gap <= tol, key
    for key in ("radius_p99", "height_p99"):
        v = np.array([e[key] for ex in runs.values() for e in ex])
        print(f"mu {mu} {key}: mean {v.mean():.4f} sd {v.std(ddof=1):.4f}", flush=True)
