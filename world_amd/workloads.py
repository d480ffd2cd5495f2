"""Granular column-collapse workloads, built as engine configurations.

Material and box shared by all of them: monodisperse spheres of diameter
d = 0.01 m and density 2500 kg/m^3 (m = 1.309e-3 kg), g = 9.81, a linear
spring-dashpot normal contact of stiffness k_n = 150 N/m with restitution
0.5, a tangential spring k_t = (2/7) k_n with contact history capped at the
Coulomb limit (physics.tangential_model: spring_history), collision cells one
diameter wide, 8 grains per cell, and a walled box whose lowest lattice layer
rests one radius above the floor.

The spring is set through physics.normal_stiffness with the force clamp
(max_force) far above any load these columns carry.  The reference's own
parameterization k = max_force / radius would put the clamp at k_n * r =
0.75 N, below the 0.82 N weight of a 64-grain stack, so floor contacts under
the large column would saturate and let grains sink through the floor.
Lattice spacing is 1.02 d with a seeded uniform jitter of +-0.02 d.
"""
from __future__ import annotations

import math

from ._capi import (WD_BACKWARD_EULER_PICARD, WD_INIT_CUBE, WD_PERTURB_GAUSSIAN,
                    WD_SEMI_IMPLICIT_EULER, WD_TANGENTIAL_SPRING_HISTORY, WdConfig,
                    default_config)

DIAMETER = 0.01
RADIUS = DIAMETER / 2
DENSITY = 2500.0
GRAIN_MASS = DENSITY * 4.0 / 3.0 * math.pi * RADIUS ** 3
K_N = 150.0
MAX_FORCE = 1000.0  # clamp on the normal spring force, never reached here
SPACING = 1.02


def column(nx: int, ny: int, nz: int, *, half_width: float = 0.9, ceiling: float = 0.5,
           friction: float = 0.5, dt: float = 1e-3, density_nodes: int = 32,
           integrator: int = WD_SEMI_IMPLICIT_EULER, picard_iterations: int = 3,
           seed: int = 12345) -> WdConfig:
    """A jittered nx x ny x nz lattice column centred on the z axis, resting on
    the floor of a (2*half_width)^2 box of height `ceiling`."""
    c = default_config()
    c.dt = dt
    # The cell domain reaches two diameters beyond every wall: the soft
    # contacts let the lowest centres under a tall column sink slightly
    # through the floor, and a grain reaching a side wall at 2 m/s overlaps it
    # by more than a radius, so its centre crosses the wall's plane.  Every
    # centre then stays inside the domain, which is what is_stable checks.
    margin = 2 * DIAMETER
    cells_xy = int(round(2 * (half_width + margin) / DIAMETER))
    cells_z = int(round((ceiling + 2 * margin) / DIAMETER))
    c.collision_grid_size_x = cells_xy
    c.collision_grid_size_y = cells_xy
    c.collision_grid_size_z = cells_z
    c.particles_per_cell = 8
    c.threads_per_block = 256
    c.density_grid_size_x = c.density_grid_size_y = c.density_grid_size_z = density_nodes
    c.x_min, c.x_max = -half_width - margin, half_width + margin
    c.y_min, c.y_max = -half_width - margin, half_width + margin
    c.z_min, c.z_max = -margin, ceiling + margin
    c.gravity = 9.81
    c.particle_radius = RADIUS
    c.normal_stiffness = K_N
    c.max_force = MAX_FORCE
    c.floor_z, c.ceiling_z = 0.0, ceiling
    c.wall_x_min, c.wall_x_max = -half_width, half_width
    c.wall_y_min, c.wall_y_max = -half_width, half_width
    c.restitution = 0.5
    c.friction = friction
    c.tangential_model = WD_TANGENTIAL_SPRING_HISTORY
    c.tangential_stiffness_ratio = 2.0 / 7.0
    c.set_masses([0.0, GRAIN_MASS])
    c.init_type = WD_INIT_CUBE
    pitch = 2 * SPACING * RADIUS
    c.init_x0 = -0.5 * (nx - 1) * pitch
    c.init_y0 = -0.5 * (ny - 1) * pitch
    c.init_z0 = RADIUS
    c.init_vx0 = c.init_vy0 = c.init_vz0 = 0.0
    c.init_jitter = 0.04
    c.init_seed = seed
    c.init_perturbation = 0.0
    # n_a = int(L_a / d): pick L_a in the middle of [n_a d, (n_a + 1) d).
    c.cube_length_x = (nx + 0.5) * DIAMETER
    c.cube_length_y = (ny + 0.5) * DIAMETER
    c.cube_length_z = (nz + 0.5) * DIAMETER
    c.particle_material = 1
    c.lattice_spacing = SPACING
    c.time_integrator = integrator
    c.picard_iterations = picard_iterations
    c.kernel_timing = 0
    return c


def column_collapse_4k(**kw) -> WdConfig:
    """4,096 grains (16^3), 1.8 m box, mu 0.5, SIE dt 1e-3, 32^3 CIC."""
    return column(16, 16, 16, **kw)


def column_collapse_32k(friction: float = 0.5, picard: bool = False, **kw) -> WdConfig:
    """32,768 grains (32^3), 1.8 m box, 64^3 CIC; SIE or Picard (3 iterations)."""
    return column(32, 32, 32, friction=friction, density_nodes=64,
                  integrator=WD_BACKWARD_EULER_PICARD if picard else WD_SEMI_IMPLICIT_EULER,
                  **kw)


def column_8k(friction: float = 0.3, **kw) -> WdConfig:
    """8,192 grains (16 x 16 x 32), the adjoint window's column."""
    return column(16, 16, 32, friction=friction, **kw)


def column_collapse_1m(**kw) -> WdConfig:
    """1,048,576 grains (128 x 128 x 64) in a 4 x 4 m box, 128^3 CIC."""
    return column(128, 128, 64, half_width=2.0, ceiling=0.8, density_nodes=128, **kw)


def sliding_grain(friction: float = 0.5, vx0: float = 1.0, dt: float = 1e-5) -> WdConfig:
    """One grain launched along the floor at vx0 (analytic deceleration -mu g)."""
    from ._capi import WD_INIT_SINGLE_PARTICLE

    c = column(1, 1, 1, friction=friction, dt=dt)
    c.init_type = WD_INIT_SINGLE_PARTICLE
    c.init_x0, c.init_y0 = -0.5, 0.0
    c.init_z0 = RADIUS - 1e-5
    c.init_vx0 = vx0
    return c


def perturbed(c: WdConfig, sigma_m: float, seed: int) -> WdConfig:
    """Adds an isotropic Gaussian offset of standard deviation sigma_m to every grain."""
    out = c.copy()
    out.init_perturbation = sigma_m / RADIUS
    out.perturbation_distribution = WD_PERTURB_GAUSSIAN
    out.init_perturbation_seed = seed
    return out


def uneven_ensemble(members: int = 64, *, seed: int = 2024, n_min: int = 2048,
                    n_max: int = 32768, aspect: tuple[float, float] = (1.0, 4.0),
                    mu: tuple[float, float] = (0.0, 0.6), sigma_m: float = 50e-6,
                    dt: float = 1e-3) -> list[WdConfig]:
    """Seed-drawn uneven column-collapse ensemble in one 1.8 m x 1.8 m x 1 m box.

    Member e draws a grain budget N_e ~ U{n_min..n_max}, an aspect ratio
    h/w ~ U[aspect] and a friction mu_e ~ U[mu]; its column is w x w x h
    lattice sites with w = round((N_e / a)^(1/3)) and h = round(N_e / w^2),
    every grain offset by an isotropic Gaussian of sigma_m, perturbation seed
    1000 + e (the chaos driver's member seeding, src/chaos.cpp:165)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for e in range(members):
        n = int(rng.integers(n_min, n_max + 1))
        a = float(rng.uniform(*aspect))
        f = float(rng.uniform(*mu))
        w = max(1, int(round((n / a) ** (1.0 / 3.0))))
        h = max(1, int(round(n / (w * w))))
        c = column(w, w, h, friction=f, ceiling=1.0, dt=dt)
        out.append(perturbed(c, sigma_m, 1000 + e))
    return out


# The demo column of the reference's interactive path (demos/sand.yaml and
# demos/fluid.yaml, which differ only in physics.friction): 25 x 25 x 52 =
# 32,500 grains of radius 0.01 m and mass 0.8 in a 2 m box, the reference's
# own spring parameterization (k = max_force / radius = 2e5 N/m), restitution
# 0.3, viscous Coulomb friction, SIE at dt 4e-4; measured after whole frames
# of 100 steps.
DEMO_STEPS_PER_FRAME = 100


def demo_column(friction: float = 0.5, dt: float = 4e-4) -> WdConfig:
    """demos/sand.yaml (friction 0.5) or demos/fluid.yaml (friction 0); the
    reference's video configuration of the same scene changes dt to 1/2520
    (and the frame to 42 steps, CLAUDE.md:1264-1267)."""
    c = default_config()
    c.dt = dt
    c.collision_grid_size_x = c.collision_grid_size_y = c.collision_grid_size_z = 100
    c.particles_per_cell = 8
    c.threads_per_block = 256
    c.density_grid_size_x = c.density_grid_size_y = c.density_grid_size_z = 32
    c.x_min, c.x_max = -1.0, 1.0
    c.y_min, c.y_max = -1.0, 1.0
    c.z_min, c.z_max = -1.0, 1.0
    c.gravity = 9.81
    c.particle_radius = 0.01
    c.max_force = 2000.0
    c.floor_z, c.ceiling_z = -1.0, 1.0
    c.wall_x_min, c.wall_x_max = -1.0, 1.0
    c.wall_y_min, c.wall_y_max = -1.0, 1.0
    c.restitution = 0.3
    c.friction = friction
    c.set_masses([0.0, 0.8])
    c.init_type = WD_INIT_CUBE
    c.init_x0, c.init_y0, c.init_z0 = -0.242, -0.242, -0.99
    c.init_vx0 = c.init_vy0 = c.init_vz0 = 0.0
    c.init_jitter = 0.25
    c.init_seed = 12345
    c.cube_length_x, c.cube_length_y, c.cube_length_z = 0.50, 0.50, 1.05
    c.particle_material = 1
    c.two_particle_separation = 0.1
    c.time_integrator = WD_SEMI_IMPLICIT_EULER
    c.picard_iterations = 3
    c.kernel_timing = 0
    return c


def demo_ensemble(friction: float, members: int, perturbation: float = 0.25,
                  dt: float = 4e-4) -> list[WdConfig]:
    """The demo column and members - 1 distinct realizations of it: every
    grain offset uniformly by up to `perturbation` radii on top of the
    jittered lattice, perturbation seed 1000 + e (the chaos driver's member
    seeding, src/chaos.cpp:163-166); member 0 is unperturbed.  0.25, the
    jitter's own magnitude, is the value the reference's chaos configuration
    gives for measuring "the scatter between distinct realizations of the
    same macroscopic state" (chaos/ensemble.yaml:92-97)."""
    base = demo_column(friction, dt)
    out = [base]
    for e in range(1, members):
        c = base.copy()
        c.init_perturbation = perturbation
        c.init_perturbation_seed = 1000 + e
        out.append(c)
    return out
