"""The reference's GoogleTest suite (test/particle_dynamics_test.cpp:16-285),
re-expressed against the Python mirror of ParticleDynamics, same defaults and
tolerances.  The C++ façade runs the same cases in tests/cpp."""
import math

import numpy as np
import pytest

from world_amd import _capi
from world_amd.sim import K_DIM, K_STATE_STRIDE, ParticleDynamics

pytestmark = pytest.mark.gpu


def test_runs_stably_for_fixed_iterations():
    with ParticleDynamics() as sim:
        for _ in range(100):
            sim.take_step()
        assert sim.is_stable()


def test_free_fall_matches_kinematics():
    with ParticleDynamics() as sim:
        sim.unpack_state()
        hs = sim.host_state.reshape(-1, K_STATE_STRIDE)
        hs[:, K_DIM:] = 0.0
        sim.device_state.copy_from_host(sim.host_state)
        z0 = sim.host_state[2]
        for _ in range(100):
            sim.take_step()
        sim.unpack_state()
        expected = -9.81 * sim.dt * sim.dt * 100 * 101 / 2.0
        assert abs((sim.positions[2] - z0) - expected) <= 1e-6


def test_particle_does_not_penetrate_floor():
    with ParticleDynamics() as sim:
        sim.host_state[:6] = [0.5, 0.5, -0.5, 0.0, 0.0, 0.0]
        sim.device_state.copy_from_host(sim.host_state)
        sim.take_step(20000)
        sim.unpack_state()
        assert sim.positions[2] >= -1.0 + 0.01 - 1e-3
        assert sim.is_stable()


def test_repeated_runs_are_deterministic():
    def run():
        with ParticleDynamics() as sim:
            for _ in range(50):
                sim.take_step()
            sim.unpack_state()
            return sim.positions.copy()

    a, b = run(), run()
    assert a.tobytes() == b.tobytes()


def test_grid_overflow_is_counted_and_stays_in_bounds():
    c = _capi.default_config()
    c.collision_grid_size_x = c.collision_grid_size_y = c.collision_grid_size_z = 1
    c.particles_per_cell = 2
    with ParticleDynamics(c) as sim:
        assert sim.n > c.particles_per_cell
        assert sim.grid_overflow_count() == 0
        sim.take_step()
        assert sim.grid_overflow_count() > 0
        sim.unpack_state()
        assert np.all(np.isfinite(sim.positions))


def test_total_energy_never_increases():
    with ParticleDynamics() as sim:
        previous = sim.compute_total_energy()
        initial = previous
        assert initial > 0.0
        tol = 1e-3 * initial
        for _ in range(12):
            for _ in range(100):
                sim.take_step()
            current = sim.compute_total_energy()
            assert current <= previous + tol
            previous = current
        assert previous < initial


def test_density_grid_conserves_total_mass():
    c = _capi.default_config()
    c.init_jitter = 0.3
    with ParticleDynamics(c) as sim:
        for _ in range(250):
            sim.take_step()
        sim.compute_density_grid()
        deposited = float(np.sum(sim.host_density_grid[: sim.density_grid_nodes()], dtype=np.float64))
        sim.unpack_state()
        expected = float(np.sum(sim.host_mass[sim.host_material], dtype=np.float64))
        assert expected > 0
        assert abs(deposited - expected) <= 1e-4 * expected


def single_particle(x, y, z, vx=0.0, vy=0.0, vz=0.0, nodes=32):
    c = _capi.default_config()
    c.init_type = _capi.WD_INIT_SINGLE_PARTICLE
    c.density_grid_size_x = c.density_grid_size_y = c.density_grid_size_z = nodes
    c.init_x0, c.init_y0, c.init_z0 = x, y, z
    c.init_vx0, c.init_vy0, c.init_vz0 = vx, vy, vz
    return c


def test_density_grid_puts_grain_on_node_at_single_node():
    with ParticleDynamics(single_particle(-0.5, 0.0, 0.5, nodes=5)) as sim:
        sim.compute_density_grid()
        m = sim.host_mass[1]
        expected_node = (1 * 5 + 2) * 5 + 3
        g = sim.host_density_grid[: sim.density_grid_nodes()]
        for node, v in enumerate(g):
            assert abs(v - (m if node == expected_node else 0.0)) <= 1e-6 * m


def test_density_grid_splits_cell_center_eight_ways():
    with ParticleDynamics(single_particle(-0.75, -0.25, 0.25, nodes=5)) as sim:
        sim.compute_density_grid()
        m = sim.host_mass[1]
        g = sim.host_density_grid[: sim.density_grid_nodes()]
        nz = g[g > 1e-9]
        assert len(nz) == 8
        assert np.all(np.abs(nz - m / 8) <= 1e-5 * m)
        assert abs(g.sum() - m) <= 1e-5 * m


def test_density_grid_momentum_matches_mass_times_velocity():
    with ParticleDynamics(single_particle(-0.5, 0.0, 0.0, 2.0, -3.0, 0.5)) as sim:
        sim.compute_density_grid()
        nodes = sim.density_grid_nodes()
        m = float(sim.host_mass[1])
        for a, v in enumerate((2.0, -3.0, 0.5)):
            mom = float(np.sum(sim.host_density_grid[(a + 1) * nodes:(a + 2) * nodes], dtype=np.float64))
            assert abs(mom - m * v) <= 1e-4 * abs(m * v) + 1e-9


def test_cube_grains_receive_initial_velocity():
    c = _capi.default_config()
    c.init_vx0, c.init_vy0, c.init_vz0 = 1.5, -0.5, 0.25
    c.gravity = 0.0
    with ParticleDynamics(c) as sim:
        assert sim.n > 1
        hs = sim.host_state.reshape(-1, K_STATE_STRIDE)
        assert np.all(hs[:, 3:] == np.array([1.5, -0.5, 0.25], np.float32))
        before = sim.positions.copy()
        for _ in range(50):
            sim.take_step()
        sim.unpack_state()
        elapsed = sim.dt * 50
        for a, v in enumerate((1.5, -0.5, 0.25)):
            moved = sim.positions[a] - before[a]
            assert abs(moved - v * elapsed) <= 1e-4 * abs(v * elapsed) + 1e-6
