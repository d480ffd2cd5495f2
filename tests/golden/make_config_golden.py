"""Writes tests/golden/config_cases.json: synthetic YAML inputs (written for
this test, exercising every syntax rule and key of the schema) together with
the values the REFERENCE's own loader (src/sim_config.cpp, built from the
reference sources into oracle/_ref/libref_config.so by oracle/Makefile)
parses out of them.

Run here, where the reference sources exist:
    make -C oracle && python tests/golden/make_config_golden.py
"""
from __future__ import annotations

import json
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import cpu  # noqa: E402

CASES = {
    "sections_lists_comments": """
# leading comment
simulation:
  dt: 0.0005   # trailing comment
  collision_grid_size_x: 180
  collision_grid_size_y: 170
  collision_grid_size_z: 50
  particles_per_cell: 8
  threads_per_block: 128
  density_grid_size_x: 64
  density_grid_size_y: 48
  density_grid_size_z: 16
domain:
  x_min: -0.9
  x_max:  0.9
  y_min: -0.8
  y_max:  0.8
  z_min: 0.0
  z_max: 0.5
physics:
  gravity: 9.81
  particle_radius: 0.005
  max_force: 0.75
  floor_z: 0.0
  ceiling_z: 0.5
  wall_x_min: -0.9
  wall_x_max: 0.9
  wall_y_min: -0.8
  wall_y_max: 0.8
  restitution: 0.5
  friction: 0.25
materials:
  masses: [0.0, 0.0013090, 2.5e-3]
initialization:
  type: cube
  x0: -0.0765
  y0: -0.0765
  z0: 0.005
  vx0: 1.5
  vy0: -0.5
  vz0: 0.25
  jitter: 0.04
  seed: 777
  perturbation: 0.01
  perturbation_seed: 1003
  cube_length_x: 0.165
  cube_length_y: 0.165
  cube_length_z: 0.325
  two_particle_separation: 0.2
  particle_material: 2
time_integration:
  solver: backward_euler_picard
  picard_iterations: 4
drivers:
  kernel_timing: false
  profiling_outer_iters: 3
  profiling_steps_per_iter: 7
dataset:
  sim_time: 1.5
  checkpoints: 5
model_problem:
  sim_time: 0.35
""",
    "quotes_and_unknown_keys": """
initialization:
  type: "single_particle"
  state_path: 'grains.bin'
  x0: 0.25
physics:
  not_a_key: 3
  friction: 0.0
time_integration:
  solver: "semi_implicit_euler"
drivers:
  kernel_timing: TRUE
bogus_section:
  anything: 1
""",
    "malformed_lines_and_empty_list": """
simulation
  dt: 0.002
materials:
  masses: []
initialization:
  type: two_particles
  two_particle_separation: 0.05
time_integration:
  solver: semi_implicit_euler
""",
    "defaults_only": """
# nothing but a comment
""",
}


def view_to_dict(v) -> dict:
    return {
        "dt": v.dt,
        "grid": list(v.grid),
        "particles_per_cell": v.particles_per_cell,
        "threads_per_block": v.threads_per_block,
        "density": list(v.density),
        "domain": list(v.domain),
        "physics": list(v.physics),
        "masses": [v.masses[i] for i in range(v.num_masses)],
        "init_type": v.init_type.decode(),
        "init_state_path": v.init_state_path.decode(),
        "init": list(v.init)[:8],
        "seeds": list(v.seeds),
        "cube": list(v.cube),
        "two_particle_separation": v.two_particle_separation,
        "particle_material": v.particle_material,
        "time_integrator": v.time_integrator.decode(),
        "picard_iterations": v.picard_iterations,
        "kernel_timing": v.kernel_timing,
    }


def main() -> None:
    cpu.build()
    if not cpu.ref_loader_available():
        raise SystemExit("reference loader not built (reference sources absent?)")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, text in CASES.items():
            p = Path(d) / f"{name}.yaml"
            p.write_text(text)
            out[name] = {"yaml": text, "expected": view_to_dict(cpu.ref_load_config(str(p)))}
    dest = Path(__file__).with_name("config_cases.json")
    dest.write_text(json.dumps(out, indent=1, sort_keys=True))
    print(f"wrote {dest}")


if __name__ == "__main__":
    main()
