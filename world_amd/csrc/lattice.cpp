// Host-side initial conditions.
//
// Behaviour follows the reference initializers (src/particle_dynamics.cu:128-235):
//   * cube: n_a = int(L_a / (2r)) grains per axis, grain p = (i*ny + j)*nz + k
//     (z fastest) at origin_a + l_a*2*s*r + U(+-jitter*r) + U(+-perturbation*r),
//     with s = lattice_spacing (reference: 1.01).  Two independent mt19937
//     streams (init_seed, init_perturbation_seed); a stream is drawn from only
//     when its magnitude is positive.  Every grain gets v0 and particle_material.
//   * two_particles: stacked along z at init_{x,y,z}0 and +separation, at rest.
//   * single_particle: one grain at init_{x,y,z}0 with init_v{x,y,z}0.
//   * file: int32 n followed by 6n float32 (the reference exits on a short
//     read; this returns WD_ERR_IO).
// The Gaussian perturbation is an extension: std::normal_distribution<float>
// with standard deviation init_perturbation * r.
//
// This file is compiled by g++ so that the <random> distributions run the
// same libstdc++ code as the CPU oracle.

#include <cstdio>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../include/world_amd.h"
#include "capi_util.h"

namespace {

struct Grains {
    std::vector<float> state;
    std::vector<int> material;
    int n = 0;
};

int cube_counts(const wd_config& c, int counts[3]) {
    const float r = c.particle_radius;
    const float len[3] = {c.cube_length_x, c.cube_length_y, c.cube_length_z};
    for (int a = 0; a < 3; ++a) counts[a] = static_cast<int>(len[a] / (2 * r));
    return counts[0] * counts[1] * counts[2];
}

int read_file_header(const char* path, int* n) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return wd::fail(WD_ERR_IO, std::string("initial state: could not open '") + path + "'");
    int32_t count = 0;
    const bool ok = std::fread(&count, sizeof(int32_t), 1, f) == 1 && count > 0;
    std::fclose(f);
    if (!ok) return wd::fail(WD_ERR_IO, std::string("initial state: bad particle count in '") + path + "'");
    *n = count;
    return WD_OK;
}

int build(const wd_config& c, Grains& g) {
    const float r = c.particle_radius;
    switch (c.init_type) {
    case WD_INIT_TWO_PARTICLES: {
        g.n = 2;
        g.state = {c.init_x0, c.init_y0, c.init_z0, 0.f, 0.f, 0.f,
                   c.init_x0, c.init_y0, c.init_z0 + c.two_particle_separation, 0.f, 0.f, 0.f};
        break;
    }
    case WD_INIT_SINGLE_PARTICLE: {
        g.n = 1;
        g.state = {c.init_x0, c.init_y0, c.init_z0, c.init_vx0, c.init_vy0, c.init_vz0};
        break;
    }
    case WD_INIT_FILE: {
        int n = 0;
        const int rc = read_file_header(c.init_state_path, &n);
        if (rc != WD_OK) return rc;
        g.n = n;
        g.state.assign(static_cast<size_t>(n) * 6, 0.f);
        std::FILE* f = std::fopen(c.init_state_path, "rb");
        if (!f) return wd::fail(WD_ERR_IO, "initial state: file vanished");
        std::fseek(f, sizeof(int32_t), SEEK_SET);
        const size_t got = std::fread(g.state.data(), sizeof(float), g.state.size(), f);
        std::fclose(f);
        if (got != g.state.size()) {
            return wd::fail(WD_ERR_IO, std::string("initial state: '") + c.init_state_path +
                                           "' is short of " + std::to_string(g.state.size()) + " floats");
        }
        break;
    }
    default: {
        int side[3];
        g.n = cube_counts(c, side);
        if (g.n < 0) return wd::fail(WD_ERR_ARG, "cube: negative grain count");
        g.state.assign(static_cast<size_t>(g.n) * 6, 0.f);
        const float origin[3] = {c.init_x0, c.init_y0, c.init_z0};
        const float v0[3] = {c.init_vx0, c.init_vy0, c.init_vz0};
        std::mt19937 jitter_rng(c.init_seed);
        std::uniform_real_distribution<float> jitter(-c.init_jitter * r, c.init_jitter * r);
        std::mt19937 pert_rng(c.init_perturbation_seed);
        std::uniform_real_distribution<float> pert_u(-c.init_perturbation * r, c.init_perturbation * r);
        std::normal_distribution<float> pert_g(0.f, c.init_perturbation * r);
        const bool gaussian = c.perturbation_distribution == WD_PERTURB_GAUSSIAN;
        const float spacing = c.lattice_spacing;
        for (int i = 0; i < side[0]; ++i) {
            for (int j = 0; j < side[1]; ++j) {
                for (int k = 0; k < side[2]; ++k) {
                    const int l[3] = {i, j, k};
                    const size_t p = (static_cast<size_t>(i) * side[1] + j) * side[2] + k;
                    for (int a = 0; a < 3; ++a) {
                        // A disabled stream still contributes +0.f, so the sum
                        // is formed exactly as with both terms present.
                        const float dj = c.init_jitter > 0.f ? jitter(jitter_rng) : 0.f;
                        const float dp = c.init_perturbation > 0.f
                                ? (gaussian ? pert_g(pert_rng) : pert_u(pert_rng)) : 0.f;
                        const float x = origin[a] + static_cast<float>(l[a]) * 2.f * spacing * r + dj + dp;
                        g.state[6 * p + a] = x;
                        g.state[6 * p + 3 + a] = v0[a];
                    }
                }
            }
        }
        break;
    }
    }
    g.material.assign(static_cast<size_t>(g.n), c.particle_material);
    return WD_OK;
}

}  // namespace

extern "C" int wd_config_grain_count(const wd_config* cfg, int* n_out) {
    if (!cfg || !n_out) return wd::fail(WD_ERR_ARG, "wd_config_grain_count: null argument");
    switch (cfg->init_type) {
    case WD_INIT_TWO_PARTICLES: *n_out = 2; return WD_OK;
    case WD_INIT_SINGLE_PARTICLE: *n_out = 1; return WD_OK;
    case WD_INIT_FILE: return read_file_header(cfg->init_state_path, n_out);
    default: {
        int side[3];
        *n_out = cube_counts(*cfg, side);
        return WD_OK;
    }
    }
}

extern "C" int wd_initial_state(const wd_config* cfg, float* state_out, int* material_out,
                                int capacity) {
    if (!cfg || !state_out || !material_out) {
        return wd::fail(WD_ERR_ARG, "wd_initial_state: null argument");
    }
    Grains g;
    const int rc = build(*cfg, g);
    if (rc != WD_OK) return rc;
    if (g.n > capacity) return wd::fail(WD_ERR_ARG, "wd_initial_state: buffers too small");
    for (size_t i = 0; i < g.state.size(); ++i) state_out[i] = g.state[i];
    for (int i = 0; i < g.n; ++i) material_out[i] = g.material[i];
    return WD_OK;
}
