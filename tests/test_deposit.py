"""The deposit observables' definitions and the demo column's parameters.

The reduction restates scripts/measure_deposit.cjs:42-62; the demo workload
must be the configuration the product loader reads from demos/sand.yaml and
demos/fluid.yaml (committed in tests/golden/demo_configs.json by
tests/golden/make_demo_configs.py)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from world_amd import _capi, deposit, workloads

GOLDEN = Path(__file__).parent / "golden" / "demo_configs.json"
REF = Path("/root/reference/demos")


def fields_of(cfg):
    out = {}
    for name, _ in _capi.WdConfig._fields_:
        v = getattr(cfg, name)
        if name == "masses":
            v = [float(x) for x in list(v)[: cfg.num_masses]]
        elif isinstance(v, bytes):
            v = v.decode()
        out[name] = v
    return out


@pytest.mark.parametrize("name,mu", [("sand", 0.5), ("fluid", 0.0)])
def test_demo_workload_is_the_demo_config(name, mu):
    want = json.loads(GOLDEN.read_text())[name]
    got = fields_of(workloads.demo_column(mu))
    for k, v in want.items():
        if k == "kernel_timing":  # the demos time every kernel; no physics
            continue
        assert got[k] == v, k
    assert _capi.grain_count(workloads.demo_column(mu)) == 25 * 25 * 52


@pytest.mark.skipif(not REF.exists(), reason="reference demos absent")
@pytest.mark.parametrize("name", ["sand", "fluid"])
def test_golden_demo_configs_are_current(name):
    want = json.loads(GOLDEN.read_text())[name]
    assert fields_of(_capi.load_config(REF / f"{name}.yaml")) == want


def test_quantile_rounds_halves_up_and_clamps():
    v = np.arange(4, dtype=np.float64)  # q (n - 1) = 1.5 at q = 0.5: JavaScript rounds up
    assert deposit.quantile(v, 0.5) == 2.0
    assert deposit.quantile(v, 0.0) == 0.0
    assert deposit.quantile(v, 1.0) == 3.0
    assert deposit.quantile(v, 0.99) == 3.0  # round(2.97) = 3
    assert deposit.quantile(np.arange(101.0), 0.99) == 99.0
    assert deposit.quantile(np.array([7.0]), 0.5) == 7.0


def test_extent_of_known_rings():
    # 200 grains on rings of radius 0.1..0.2 (100 each) at heights 1 and 3
    # above z_min = -1: medians and 99th percentiles are known exactly
    ang = np.linspace(0, 2 * math.pi, 100, endpoint=False)
    pts = []
    for r, z in ((0.1, 0.0), (0.2, 2.0)):
        pts.append(np.stack([r * np.cos(ang), r * np.sin(ang), np.full_like(ang, z)], axis=1))
    st = np.concatenate([np.concatenate(pts), np.zeros((200, 3))], axis=1).astype(np.float32)
    e = deposit.extent(st, -1.0)
    assert e["radius_p50"] == pytest.approx(0.2, abs=1e-6)   # index round(99.5) = 100
    assert e["radius_p99"] == pytest.approx(0.2, abs=1e-6)
    assert e["height_p50"] == pytest.approx(3.0)
    e2 = deposit.extent(st[:100], -1.0)
    assert e2["radius_p50"] == pytest.approx(0.1, abs=1e-6) and e2["height_max"] == pytest.approx(1.0)


def test_published_measurements_are_listed_by_configuration():
    want = json.loads(GOLDEN.read_text())["published_medians_t2.4"]
    assert [list(m) for m in deposit.MEASUREMENTS] == want
    assert deposit.published(0.5, "radius_p50") == [0.550, 0.540]
    assert sorted(deposit.published(0.0, "height_p50")) == [0.026, 0.026]
    for dt, per_frame in deposit.CONFIGS:  # both frames are 1/25 and 1/60 s of simulated time
        assert abs(dt * per_frame - (0.04 if per_frame == 100 else 1.0 / 60.0)) < 1e-12


def test_measurement_time_is_whole_frames_past_the_target():
    # FP32 dt = 4e-4 rounds below 4e-4, so even the exact sum of 6,000 steps
    # stays below 2.4: the script measures after the 61st frame
    dt32 = float(np.float32(4e-4))
    assert math.fsum([dt32] * 6000) < float(np.float32(2.4))
    assert deposit.frames_until(4e-4, 2.4, 100) == 6100
    # dt = 0.5 is exact in FP32 and so is every partial sum: 4 frames of 3 reach 6
    assert deposit.frames_until(0.5, 6.0, 3) == 12
