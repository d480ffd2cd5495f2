// sim_config.h -- configuration schema of the granular solver.
//
// Same names, fields, defaults and YAML keys as the reference
// (src/sim_config.h:14-205), so drivers written against the reference compile
// unchanged.  Extension fields (marked "ext") are opt-in; their defaults
// reproduce the reference's behaviour.
#pragma once

#include <cstring>
#include <string>
#include <vector>

#include "../world_amd.h"

// Per-grain state is kStateStride floats [x, y, z, vx, vy, vz]; z is up.
constexpr int kDim = 3;
constexpr int kStateStride = 2 * kDim;
// 3x3x3 neighbor-search stencil.
constexpr int kStencilCells = 27;

struct DomainParams {
    float x_min;
    float x_max;
    float y_min;
    float y_max;
    float z_min;
    float z_max;
};

struct PhysicsParams {
    float gravity;
    float particle_radius;
    float max_force;
    float floor_z;
    float ceiling_z;
    float wall_x_min;
    float wall_x_max;
    float wall_y_min;
    float wall_y_max;
    float restitution;
    float friction;
};

struct SimConfig {
    float dt = 0.0001f;
    int collision_grid_size_x = 32;
    int collision_grid_size_y = 32;
    int collision_grid_size_z = 32;
    int particles_per_cell = 64;
    int threads_per_block = 256;

    int density_grid_size_x = 32;
    int density_grid_size_y = 32;
    int density_grid_size_z = 32;

    DomainParams domain{-1.0f, 1.0f, -1.0f, 1.0f, -1.0f, 1.0f};
    PhysicsParams physics{9.81f, 0.01f, 100.0f,
                          -1.0f, 1.0f, -1.0f, 1.0f, -1.0f, 1.0f, 0.3f, 0.5f};

    // ext: "viscous" (reference) or "spring_history".
    std::string tangential_model = "viscous";
    float tangential_stiffness_ratio = 2.0f / 7.0f;  // ext
    float normal_stiffness = 0.0f;                   // ext; <= 0 means max_force/radius

    std::vector<float> masses{0.0f, 0.04f};

    std::string init_type = "cube";  // "cube" | "two_particles" | "single_particle" | "file"
    std::string init_state_path = "";
    float init_x0 = -0.5f;
    float init_y0 = 0.0f;
    float init_z0 = 0.0f;
    float init_vx0 = 0.0f;
    float init_vy0 = 0.0f;
    float init_vz0 = 0.0f;
    float init_jitter = 0.0f;
    unsigned int init_seed = 12345u;
    float init_perturbation = 0.0f;
    unsigned int init_perturbation_seed = 1u;
    float cube_length_x = 0.2f;
    float cube_length_y = 0.2f;
    float cube_length_z = 0.2f;
    float two_particle_separation = 0.1f;
    int particle_material = 1;
    float lattice_spacing = 1.01f;                    // ext, in diameters
    std::string perturbation_distribution = "uniform";  // ext: "uniform" | "gaussian"

    int num_triangles = 20;

    int port = 8081;
    int steps_per_frame = 10;
    int profiling_outer_iters = 10;
    int profiling_steps_per_iter = 10;
    bool kernel_timing = true;

    std::vector<float> sweep_dt{};
    std::vector<float> sweep_max_force{};
    std::vector<float> sweep_restitution{};

    std::string stability_mode = "fixed_time";
    float stability_sim_time = 0.1f;
    float stability_acceleration_cutoff = 0.01f;
    int stability_steps_per_steadiness_check = 100;

    float model_problem_sim_time = 1.0f;

    float dataset_sim_time = 2.0f;
    int dataset_checkpoints = 10;

    std::string time_integrator = "semi_implicit_euler";  // anything else selects Picard
    int picard_iterations = 3;
};

// Loads a YAML-subset file over the defaults (implemented in libworld_amd.so).
SimConfig load_config(const std::string& path);

namespace world {

// Flattens a SimConfig into the C ABI's plain-data wd_config.  The integrator
// test mirrors the reference's take_step, which runs semi-implicit Euler only
// for the exact string "semi_implicit_euler" and Picard for anything else
// (src/particle_dynamics.cu:265).
inline wd_config to_wd_config(const SimConfig& c) {
    wd_config w;
    std::memset(&w, 0, sizeof(w));
    w.dt = c.dt;
    w.collision_grid_size_x = c.collision_grid_size_x;
    w.collision_grid_size_y = c.collision_grid_size_y;
    w.collision_grid_size_z = c.collision_grid_size_z;
    w.particles_per_cell = c.particles_per_cell;
    w.threads_per_block = c.threads_per_block;
    w.density_grid_size_x = c.density_grid_size_x;
    w.density_grid_size_y = c.density_grid_size_y;
    w.density_grid_size_z = c.density_grid_size_z;
    w.x_min = c.domain.x_min; w.x_max = c.domain.x_max;
    w.y_min = c.domain.y_min; w.y_max = c.domain.y_max;
    w.z_min = c.domain.z_min; w.z_max = c.domain.z_max;
    w.gravity = c.physics.gravity;
    w.particle_radius = c.physics.particle_radius;
    w.max_force = c.physics.max_force;
    w.floor_z = c.physics.floor_z;
    w.ceiling_z = c.physics.ceiling_z;
    w.wall_x_min = c.physics.wall_x_min; w.wall_x_max = c.physics.wall_x_max;
    w.wall_y_min = c.physics.wall_y_min; w.wall_y_max = c.physics.wall_y_max;
    w.restitution = c.physics.restitution;
    w.friction = c.physics.friction;
    w.tangential_model = (c.tangential_model == "spring_history")
            ? WD_TANGENTIAL_SPRING_HISTORY : WD_TANGENTIAL_VISCOUS;
    w.tangential_stiffness_ratio = c.tangential_stiffness_ratio;
    w.normal_stiffness = c.normal_stiffness;
    const size_t nm = c.masses.size() < WD_MAX_MATERIALS ? c.masses.size() : WD_MAX_MATERIALS;
    w.num_masses = static_cast<int>(nm);
    for (size_t i = 0; i < nm; ++i) w.masses[i] = c.masses[i];
    if (c.init_type == "file") w.init_type = WD_INIT_FILE;
    else if (c.init_type == "two_particles") w.init_type = WD_INIT_TWO_PARTICLES;
    else if (c.init_type == "single_particle") w.init_type = WD_INIT_SINGLE_PARTICLE;
    else w.init_type = WD_INIT_CUBE;
    std::strncpy(w.init_state_path, c.init_state_path.c_str(), WD_PATH_MAX - 1);
    w.init_x0 = c.init_x0; w.init_y0 = c.init_y0; w.init_z0 = c.init_z0;
    w.init_vx0 = c.init_vx0; w.init_vy0 = c.init_vy0; w.init_vz0 = c.init_vz0;
    w.init_jitter = c.init_jitter;
    w.init_seed = c.init_seed;
    w.init_perturbation = c.init_perturbation;
    w.init_perturbation_seed = c.init_perturbation_seed;
    w.cube_length_x = c.cube_length_x;
    w.cube_length_y = c.cube_length_y;
    w.cube_length_z = c.cube_length_z;
    w.two_particle_separation = c.two_particle_separation;
    w.particle_material = c.particle_material;
    w.lattice_spacing = c.lattice_spacing;
    w.perturbation_distribution = (c.perturbation_distribution == "gaussian")
            ? WD_PERTURB_GAUSSIAN : WD_PERTURB_UNIFORM;
    w.time_integrator = (c.time_integrator == "semi_implicit_euler")
            ? WD_SEMI_IMPLICIT_EULER : WD_BACKWARD_EULER_PICARD;
    w.picard_iterations = c.picard_iterations;
    w.kernel_timing = c.kernel_timing ? 1 : 0;
    return w;
}

}  // namespace world
