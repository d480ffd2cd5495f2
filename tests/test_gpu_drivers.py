"""The driver counterparts (world_amd/bin/dataset, chaos, model_problem,
stability_and_accuracy) write the reference drivers' file formats, and the
batched runs write exactly the bytes of the sequential ones (one simulation
at a time, as the reference drivers loop).

Layouts checked here are the reference's own: grids.bin (magic RTPGRD01,
int32[7] header, checkpoint times, [R][K][4][nx][ny][nz] float32; documented
in surrogate/dataset_io.py:8-21 and written by src/dataset.cpp:161-214),
manifest.csv (src/dataset.cpp:176-226), divergence.csv and final_fields.bin
(src/chaos.cpp:134-218), history.dat (src/model_problem.cpp:108-142),
stability_results.dat (src/stability_and_accuracy.cpp:251-267)."""
import csv
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "world_amd" / "bin"

BLOB = """
simulation:
  dt: 0.0002
  collision_grid_size_x: 100
  collision_grid_size_y: 100
  collision_grid_size_z: 70
  particles_per_cell: 8
  threads_per_block: 256
  density_grid_size_x: 16
  density_grid_size_y: 16
  density_grid_size_z: 8
domain:
  x_min: -1.0
  x_max: 1.0
  y_min: -1.0
  y_max: 1.0
  z_min: -1.0
  z_max: 0.4
physics:
  gravity: 9.81
  particle_radius: 0.01
  max_force: 2000.0
  floor_z: -1.0
  ceiling_z: 0.4
  wall_x_min: -1.0
  wall_x_max: 1.0
  wall_y_min: -1.0
  wall_y_max: 1.0
  restitution: 0.3
  friction: 0.5
materials:
  masses: [0.0, 0.08]
initialization:
  type: cube
  x0: -0.05
  y0: -0.05
  z0: 0.0
  cube_length_x: 0.1
  cube_length_y: 0.1
  cube_length_z: 0.1
  jitter: 0.25
  seed: 12345
  particle_material: 1
dataset:
  sim_time: 0.06
  checkpoints: 4
"""

DESIGN = """throw_vx,throw_vy,throw_vz,release_z,grain_mass,restitution,friction\r
0.5,0.0,0.0,-0.5,0.08,0.3,0.5\r
-1.0,0.3,0.5,-0.3,0.05,0.6,0.2\r
0.0,0.0,-2.0,-0.8,0.12,0.1,0.8\r
600.0,0.0,0.0,-0.5,0.08,0.3,0.5\r
0.2,-0.4,1.0,-0.6,0.02,0.9,0.0\r
"""


def run(args, cwd):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r


def read_grids(path):
    raw = Path(path).read_bytes()
    assert raw[:8] == b"RTPGRD01"
    hdr = np.frombuffer(raw[8:36], np.int32)
    R, K, ch, nx, ny, nz, npar = (int(v) for v in hdr)
    times = np.frombuffer(raw[36:36 + 4 * K], np.float32)
    grids = np.frombuffer(raw[36 + 4 * K:], np.float32)
    assert grids.size == R * K * ch * nx * ny * nz
    return (R, K, ch, nx, ny, nz, npar), times, grids.reshape(R, K, ch, nx, ny, nz)


def test_dataset_batched_equals_sequential_and_layout(tmp_path):
    (tmp_path / "base.yaml").write_text(BLOB)
    (tmp_path / "design.csv").write_text(DESIGN)
    run([str(BIN / "dataset"), "base.yaml", "design.csv", "batched"], tmp_path)
    run([str(BIN / "dataset"), "base.yaml", "design.csv", "sequential", "--sequential"], tmp_path)
    for name in ("grids.bin", "manifest.csv"):
        assert (tmp_path / "batched" / name).read_bytes() == (tmp_path / "sequential" / name).read_bytes(), name

    (R, K, ch, nx, ny, nz, npar), times, grids = read_grids(tmp_path / "batched" / "grids.bin")
    assert (R, K, ch, nx, ny, nz, npar) == (5, 4, 4, 16, 16, 8, 7)
    np.testing.assert_array_equal(times, np.float32(0.06) * np.arange(1, 5, dtype=np.float32) / np.float32(4))
    rows = list(csv.DictReader(open(tmp_path / "batched" / "manifest.csv")))
    assert list(rows[0].keys())[:8] == ["rollout", "throw_vx", "throw_vy", "throw_vz", "release_z",
                                       "grain_mass", "restitution", "friction"]
    assert [r["status"] for r in rows] == ["ok", "ok", "ok", "diverged", "ok"]
    n = 125  # 0.1 m cube of 0.02 m grains
    for r, row in enumerate(rows):
        if row["status"] == "ok":
            mass = n * float(row["grain_mass"])
            # every grain's mass lands on the grid at every checkpoint
            np.testing.assert_allclose(grids[r, :, 0].sum(axis=(1, 2, 3)), mass, rtol=1e-5)
            assert float(row["deposited_mass"]) == pytest.approx(mass, rel=1e-5)
            assert float(row["final_time"]) == pytest.approx(0.06, rel=1e-4)
        else:
            # the thrown row leaves the box before the first checkpoint
            assert not grids[r].any()
            assert float(row["final_time"]) == pytest.approx(0.015, rel=1e-4)


CHAOS = BLOB.replace("  seed: 12345\n", "  seed: 12345\n  perturbation: 0.02\n  perturbation_seed: 0\n")


def test_chaos_batched_equals_sequential_and_layout(tmp_path):
    (tmp_path / "ens.yaml").write_text(CHAOS)
    run([str(BIN / "chaos"), "ens.yaml", "4", "batched"], tmp_path)
    run([str(BIN / "chaos"), "ens.yaml", "4", "sequential", "--sequential"], tmp_path)
    for name in ("divergence.csv", "final_fields.bin"):
        assert (tmp_path / "batched" / name).read_bytes() == (tmp_path / "sequential" / name).read_bytes(), name
    rows = list(csv.DictReader(open(tmp_path / "batched" / "divergence.csv")))
    assert len(rows) == 3 * 4
    assert [(int(r["member"]), int(r["checkpoint"])) for r in rows] == [(e, k) for e in (1, 2, 3) for k in range(4)]
    assert all(float(r["grain_rms_displacement"]) > 0 for r in rows)  # perturbed members differ
    raw = (tmp_path / "batched" / "final_fields.bin").read_bytes()
    E, nodes = np.frombuffer(raw[:8], np.int32)
    assert (E, nodes) == (4, 16 * 16 * 8)
    fields = np.frombuffer(raw[8:], np.float32).reshape(E, nodes)
    np.testing.assert_allclose(fields.sum(axis=1), 125 * 0.08, rtol=1e-5)


MODEL = """
simulation:
  dt: 0.00001
physics:
  gravity: 9.81
  particle_radius: 0.01
  max_force: 2000.0
  restitution: 0.3
  friction: 0.5
materials:
  masses: [0.0, 0.8, 0.8]
initialization:
  type: single_particle
  x0: -0.5
  y0: 0.0
  z0: -0.98999
  vx0: 1.0
  vy0: 0.0
  vz0: 0.0
sweep:
  restitution: [0.3, 0.6]
model_problem:
  sim_time: 0.002
"""


def test_model_problem_history_layout(tmp_path):
    (tmp_path / "mp.yaml").write_text(MODEL)
    run([str(BIN / "model_problem"), "mp.yaml", "out"], tmp_path)
    text = (tmp_path / "out" / "history.dat").read_text()
    assert text.startswith("# step time x0 y0 z0 vx0 vy0 vz0\n")
    blocks = text.split("\n\n\n")
    assert len(blocks) == 2
    for c, block in enumerate(blocks):
        lines = [l for l in block.strip("\n").split("\n") if l and not l.startswith("# step")]
        assert lines[0].startswith(f"# combo {c}: dt=9.99999975e-06 max_force=2000 restitution=")
        data = np.array([[float(v) for v in l.split()] for l in lines[1:]])
        assert data.shape == (201, 8)  # steps 0..200: the initial row plus one per step
        np.testing.assert_array_equal(data[:, 0], np.arange(201))
        np.testing.assert_allclose(data[0, 2:], [-0.5, 0.0, -0.98999, 1.0, 0.0, 0.0], rtol=1e-6)
        assert data[-1, 5] < 1.0  # friction has slowed the grain


def sweep_rows(path):
    rows = []
    for line in Path(path).read_text().splitlines():
        if line.startswith("#"):
            assert line == "# dt max_force status(0=unstable,1=steady,2=timed_out) energy"
            continue
        dt, mf, st, e = line.split()
        rows.append((float(dt), float(mf), int(st), float(e)))
    return rows


def test_stability_sweep_classifies_each_pair_in_its_own_process(tmp_path):
    """The stability sweep (src/stability_and_accuracy.cpp) runs every
    (dt, max_force) pair in a child process and writes
    stability_results.dat.  A 0.2 s step carries the falling cube out of the
    domain (U), a 0.2 ms step keeps it in bounds (S); in steady_state mode a
    run still falling at the end is timed out (T), and a run whose grid
    overflows ends its process and is counted unstable (U)."""
    base = BLOB.split("dataset:")[0]
    fixed = base + """stability:
  mode: fixed_time
  sim_time: 0.4
sweep:
  dt: [0.0002, 0.2]
  max_force: [2000.0, 4000.0]
"""
    (tmp_path / "fixed.yaml").write_text(fixed)
    r = run([str(BIN / "stability_and_accuracy"), "fixed.yaml", "fixed"], tmp_path)
    assert "Legend: U = unstable, S = steady, T = timed out" in r.stdout
    rows = sweep_rows(tmp_path / "fixed" / "stability_results.dat")
    assert [(round(d, 6), m) for d, m, _, _ in rows] == [(0.0002, 2000.0), (0.0002, 4000.0), (0.2, 2000.0), (0.2, 4000.0)]
    assert [s for _, _, s, _ in rows] == [1, 1, 0, 0]
    assert (tmp_path / "fixed" / "stability_plot.gp").read_text().count("set object") == 4

    # dt 0.2 ms: omega dt = 0.32 for k = max_force / r, well inside the damped
    # semi-implicit Euler limit (at 1 ms the jittered cube's contacts blow up)
    steady = base.replace("  particles_per_cell: 8\n", "  particles_per_cell: 1\n") + """stability:
  mode: steady_state
  sim_time: 0.004
  acceleration_cutoff: 0.01
  steps_per_steadiness_check: 10
sweep:
  dt: [0.0002]
  max_force: [2000.0]
"""
    packed = steady.replace("  jitter: 0.25\n", "  jitter: 0.25\n  lattice_spacing: 0.5\n")
    (tmp_path / "falling.yaml").write_text(steady.replace("  particles_per_cell: 1\n", "  particles_per_cell: 8\n"))
    (tmp_path / "packed.yaml").write_text(packed)
    run([str(BIN / "stability_and_accuracy"), "falling.yaml", "falling"], tmp_path)
    assert [s for _, _, s, _ in sweep_rows(tmp_path / "falling" / "stability_results.dat")] == [2]
    run([str(BIN / "stability_and_accuracy"), "packed.yaml", "packed"], tmp_path)
    assert [s for _, _, s, _ in sweep_rows(tmp_path / "packed" / "stability_results.dat")] == [0]
