"""Writes tests/golden/demo_configs.json: the configurations the product's
YAML loader (wd_config_load) reads from the reference's demo files
demos/sand.yaml and demos/fluid.yaml (the column collapse whose deposit the
reference publishes, README.md:24-30), field by field, plus the published
deposit medians.  tests/test_deposit.py checks workloads.demo_column against
it, so the demo parameters are pinned where the reference tree is absent.

Run here, where the reference sources exist:
    python tests/golden/make_demo_configs.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from world_amd import _capi, deposit  # noqa: E402

REF = Path("/root/reference")


def fields(cfg) -> dict:
    out = {}
    for name, _ in _capi.WdConfig._fields_:
        v = getattr(cfg, name)
        if name == "masses":
            v = [float(x) for x in list(v)[: cfg.num_masses]]
        elif isinstance(v, bytes):
            v = v.decode()
        out[name] = v
    return out


def main() -> int:
    data = {"published_medians_t2.4": [list(m) for m in deposit.MEASUREMENTS]}
    for name in ("sand", "fluid"):
        data[name] = fields(_capi.load_config(REF / "demos" / f"{name}.yaml"))
    (Path(__file__).parent / "demo_configs.json").write_text(json.dumps(data, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
