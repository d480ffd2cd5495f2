// Test infrastructure only: a C entry point over the REFERENCE's own YAML
// loader (src/sim_config.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/).  The parity test loads every checked-in
// YAML through both this and the product's wd_config_load and compares the
// resulting fields.  Extension fields do not exist in the reference and are
// reported at their defaults.
#include <cstring>
#include <string>

#include "sim_config.h"  // the reference header (-I <reference>/src)

extern "C" {

// Flat view of the reference SimConfig fields the solver consumes.
struct ref_config_view {
    float dt;
    int grid[3];
    int particles_per_cell;
    int threads_per_block;
    int density[3];
    float domain[6];
    float physics[11];
    int num_masses;
    float masses[16];
    char init_type[64];
    char init_state_path[512];
    float init[10];  // x0 y0 z0 vx0 vy0 vz0 jitter perturbation
    unsigned int seeds[2];
    float cube[3];
    float two_particle_separation;
    int particle_material;
    char time_integrator[64];
    int picard_iterations;
    int kernel_timing;
    int profiling[2];
    float dataset_sim_time;
    int dataset_checkpoints;
    float model_problem_sim_time;
};

int ref_load_config(const char* path, ref_config_view* out) {
    const SimConfig c = load_config(path);
    std::memset(out, 0, sizeof(*out));
    out->dt = c.dt;
    out->grid[0] = c.collision_grid_size_x;
    out->grid[1] = c.collision_grid_size_y;
    out->grid[2] = c.collision_grid_size_z;
    out->particles_per_cell = c.particles_per_cell;
    out->threads_per_block = c.threads_per_block;
    out->density[0] = c.density_grid_size_x;
    out->density[1] = c.density_grid_size_y;
    out->density[2] = c.density_grid_size_z;
    const float d[6] = {c.domain.x_min, c.domain.x_max, c.domain.y_min,
                        c.domain.y_max, c.domain.z_min, c.domain.z_max};
    std::memcpy(out->domain, d, sizeof(d));
    const PhysicsParams& p = c.physics;
    const float ph[11] = {p.gravity, p.particle_radius, p.max_force, p.floor_z, p.ceiling_z,
                          p.wall_x_min, p.wall_x_max, p.wall_y_min, p.wall_y_max,
                          p.restitution, p.friction};
    std::memcpy(out->physics, ph, sizeof(ph));
    out->num_masses = static_cast<int>(c.masses.size());
    for (size_t i = 0; i < c.masses.size() && i < 16; ++i) out->masses[i] = c.masses[i];
    std::strncpy(out->init_type, c.init_type.c_str(), 63);
    std::strncpy(out->init_state_path, c.init_state_path.c_str(), 511);
    const float in[10] = {c.init_x0, c.init_y0, c.init_z0, c.init_vx0, c.init_vy0,
                          c.init_vz0, c.init_jitter, c.init_perturbation, 0.f, 0.f};
    std::memcpy(out->init, in, sizeof(in));
    out->seeds[0] = c.init_seed;
    out->seeds[1] = c.init_perturbation_seed;
    out->cube[0] = c.cube_length_x;
    out->cube[1] = c.cube_length_y;
    out->cube[2] = c.cube_length_z;
    out->two_particle_separation = c.two_particle_separation;
    out->particle_material = c.particle_material;
    std::strncpy(out->time_integrator, c.time_integrator.c_str(), 63);
    out->picard_iterations = c.picard_iterations;
    out->kernel_timing = c.kernel_timing ? 1 : 0;
    out->profiling[0] = c.profiling_outer_iters;
    out->profiling[1] = c.profiling_steps_per_iter;
    out->dataset_sim_time = c.dataset_sim_time;
    out->dataset_checkpoints = c.dataset_checkpoints;
    out->model_problem_sim_time = c.model_problem_sim_time;
    return 0;
}

}  // extern "C"
