"""Configuration layer: the product's YAML loader (wd_config_load) against the
reference loader's results (src/sim_config.cpp:54-195), the reference
defaults (src/sim_config.h:64-200) and its documented quirks."""
import json
from pathlib import Path

import numpy as np
import pytest

from world_amd import _capi

GOLDEN = Path(__file__).parent / "golden" / "config_cases.json"
REFERENCE = Path("/root/reference")

PHYS = ["gravity", "particle_radius", "max_force", "floor_z", "ceiling_z", "wall_x_min",
        "wall_x_max", "wall_y_min", "wall_y_max", "restitution", "friction"]
INIT = ["init_x0", "init_y0", "init_z0", "init_vx0", "init_vy0", "init_vz0", "init_jitter",
        "init_perturbation"]
INIT_TYPES = {"file": _capi.WD_INIT_FILE, "two_particles": _capi.WD_INIT_TWO_PARTICLES,
              "single_particle": _capi.WD_INIT_SINGLE_PARTICLE}


def f32(x):
    return np.float32(x)


def assert_matches(cfg: _capi.WdConfig, exp: dict) -> None:
    assert f32(cfg.dt) == f32(exp["dt"])
    assert [cfg.collision_grid_size_x, cfg.collision_grid_size_y, cfg.collision_grid_size_z] == exp["grid"]
    assert cfg.particles_per_cell == exp["particles_per_cell"]
    assert cfg.threads_per_block == exp["threads_per_block"]
    assert [cfg.density_grid_size_x, cfg.density_grid_size_y, cfg.density_grid_size_z] == exp["density"]
    dom = [cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max]
    assert [f32(v) for v in dom] == [f32(v) for v in exp["domain"]]
    assert [f32(getattr(cfg, k)) for k in PHYS] == [f32(v) for v in exp["physics"]]
    assert [f32(v) for v in cfg.get_masses()] == [f32(v) for v in exp["masses"]]
    assert cfg.init_type == INIT_TYPES.get(exp["init_type"], _capi.WD_INIT_CUBE)
    assert cfg.init_state_path.decode() == exp["init_state_path"]
    assert [f32(getattr(cfg, k)) for k in INIT] == [f32(v) for v in exp["init"]]
    assert [cfg.init_seed, cfg.init_perturbation_seed] == exp["seeds"]
    assert [f32(cfg.cube_length_x), f32(cfg.cube_length_y), f32(cfg.cube_length_z)] == \
        [f32(v) for v in exp["cube"]]
    assert f32(cfg.two_particle_separation) == f32(exp["two_particle_separation"])
    assert cfg.particle_material == exp["particle_material"]
    want = (_capi.WD_SEMI_IMPLICIT_EULER if exp["time_integrator"] == "semi_implicit_euler"
            else _capi.WD_BACKWARD_EULER_PICARD)
    assert cfg.time_integrator == want
    assert cfg.picard_iterations == exp["picard_iterations"]
    assert cfg.kernel_timing == exp["kernel_timing"]


@pytest.mark.parametrize("case", sorted(json.loads(GOLDEN.read_text())))
def test_loader_matches_reference_loader_fixture(case, tmp_path):
    data = json.loads(GOLDEN.read_text())[case]
    p = tmp_path / "c.yaml"
    p.write_text(data["yaml"])
    assert_matches(_capi.load_config(p), data["expected"])


def test_defaults_are_reference_defaults():
    c = _capi.default_config()
    # Literal defaults of the reference SimConfig (src/sim_config.h:68-200).
    assert f32(c.dt) == f32(0.0001)
    assert (c.collision_grid_size_x, c.collision_grid_size_y, c.collision_grid_size_z) == (32, 32, 32)
    assert c.particles_per_cell == 64 and c.threads_per_block == 256
    assert (c.x_min, c.x_max, c.y_min, c.y_max, c.z_min, c.z_max) == (-1, 1, -1, 1, -1, 1)
    assert [f32(getattr(c, k)) for k in PHYS] == [f32(v) for v in
                                                 [9.81, 0.01, 100, -1, 1, -1, 1, -1, 1, 0.3, 0.5]]
    assert [f32(v) for v in c.get_masses()] == [f32(0.0), f32(0.04)]
    assert c.init_type == _capi.WD_INIT_CUBE
    assert (f32(c.init_x0), c.init_y0, c.init_z0) == (f32(-0.5), 0.0, 0.0)
    assert (f32(c.cube_length_x), f32(c.cube_length_y), f32(c.cube_length_z)) == (f32(0.2),) * 3
    assert c.init_seed == 12345 and c.init_perturbation_seed == 1
    assert c.time_integrator == _capi.WD_SEMI_IMPLICIT_EULER and c.picard_iterations == 3
    assert c.kernel_timing == 1
    # Extension defaults reproduce the reference.
    assert c.tangential_model == _capi.WD_TANGENTIAL_VISCOUS
    assert f32(c.lattice_spacing) == f32(1.01)
    assert c.normal_stiffness == 0.0
    assert c.perturbation_distribution == _capi.WD_PERTURB_UNIFORM


def test_missing_file_gives_defaults(tmp_path, capfd):
    c = _capi.load_config(tmp_path / "absent.yaml")
    d = _capi.default_config()
    assert bytes(c) == bytes(d)


def test_extension_keys(tmp_path):
    p = tmp_path / "ext.yaml"
    p.write_text("physics:\n  tangential_model: spring_history\n  tangential_stiffness_ratio: 0.5\n"
                 "  normal_stiffness: 150\ninitialization:\n  lattice_spacing: 1.02\n"
                 "  perturbation_distribution: gaussian\n")
    c = _capi.load_config(p)
    assert c.tangential_model == _capi.WD_TANGENTIAL_SPRING_HISTORY
    assert c.tangential_stiffness_ratio == 0.5 and c.normal_stiffness == 150.0
    assert f32(c.lattice_spacing) == f32(1.02)
    assert c.perturbation_distribution == _capi.WD_PERTURB_GAUSSIAN


def test_malformed_number_is_an_error(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("simulation:\n  dt: fast\n")
    with pytest.raises(_capi.WdError):
        _capi.load_config(p)


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference sources not mounted")
def test_every_reference_yaml_parses_identically():
    from oracle import cpu

    cpu.build()
    if not cpu.ref_loader_available():
        pytest.skip("reference loader not built")
    from tests.golden.make_config_golden import view_to_dict

    files = sorted(REFERENCE.rglob("*.yaml"))
    assert len(files) >= 20
    for f in files:
        assert_matches(_capi.load_config(f), view_to_dict(cpu.ref_load_config(str(f))))
