// YAML-subset configuration loader.
//
// Accepted syntax and failure behaviour follow the reference loader
// (src/sim_config.cpp:54-195): a bare "name:" line opens a section, indented
// "key: value" lines assign, "[a, b]" is an inline float list, '#' starts a
// comment, unknown keys print a warning and are skipped, and a file that
// cannot be opened yields the defaults with a notice on stderr.  Numbers go
// through std::stof / std::stoi, so malformed numbers throw just as they do in
// the reference; the C entry point converts that into an error status.
//
// The key set is held in one table of (section, key, setter) rows instead of
// nested branches, which keeps the schema in one place next to the C-ABI
// conversion in include/world/sim_config.h.

#include <cstdio>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/world/sim_config.h"
#include "capi_util.h"

namespace {

std::string strip(const std::string& s) {
    const char* ws = " \t\r\n";
    const auto b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    const auto e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

// Removes one pair of matching single or double quotes.
std::string dequote(const std::string& s) {
    if (s.size() >= 2 && (s[0] == '"' || s[0] == '\'') && s.back() == s[0]) {
        return s.substr(1, s.size() - 2);
    }
    return s;
}

std::vector<float> float_list(const std::string& raw) {
    std::string body = strip(raw);
    if (body.size() >= 2 && body.front() == '[' && body.back() == ']') {
        body = body.substr(1, body.size() - 2);
    }
    std::vector<float> out;
    std::stringstream ss(body);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = strip(item);
        if (!item.empty()) out.push_back(std::stof(item));
    }
    return out;
}

using Setter = std::function<void(SimConfig&, const std::string&)>;

Setter F(float SimConfig::*m) {
    return [m](SimConfig& c, const std::string& v) { c.*m = std::stof(v); };
}
Setter I(int SimConfig::*m) {
    return [m](SimConfig& c, const std::string& v) { c.*m = std::stoi(v); };
}
Setter U(unsigned int SimConfig::*m) {
    return [m](SimConfig& c, const std::string& v) { c.*m = static_cast<unsigned int>(std::stoi(v)); };
}
Setter B(bool SimConfig::*m) {
    return [m](SimConfig& c, const std::string& v) { c.*m = (v == "true"); };
}
Setter S(std::string SimConfig::*m, bool unquote) {
    return [m, unquote](SimConfig& c, const std::string& v) { c.*m = unquote ? dequote(v) : v; };
}
Setter L(std::vector<float> SimConfig::*m) {
    return [m](SimConfig& c, const std::string& v) { c.*m = float_list(v); };
}
Setter D(float DomainParams::*m) {
    return [m](SimConfig& c, const std::string& v) { c.domain.*m = std::stof(v); };
}
Setter P(float PhysicsParams::*m) {
    return [m](SimConfig& c, const std::string& v) { c.physics.*m = std::stof(v); };
}

const std::unordered_map<std::string, Setter>& schema() {
    static const std::unordered_map<std::string, Setter> table = {
        {"simulation.dt", F(&SimConfig::dt)},
        {"simulation.collision_grid_size_x", I(&SimConfig::collision_grid_size_x)},
        {"simulation.collision_grid_size_y", I(&SimConfig::collision_grid_size_y)},
        {"simulation.collision_grid_size_z", I(&SimConfig::collision_grid_size_z)},
        {"simulation.particles_per_cell", I(&SimConfig::particles_per_cell)},
        {"simulation.threads_per_block", I(&SimConfig::threads_per_block)},
        {"simulation.density_grid_size_x", I(&SimConfig::density_grid_size_x)},
        {"simulation.density_grid_size_y", I(&SimConfig::density_grid_size_y)},
        {"simulation.density_grid_size_z", I(&SimConfig::density_grid_size_z)},
        {"domain.x_min", D(&DomainParams::x_min)},
        {"domain.x_max", D(&DomainParams::x_max)},
        {"domain.y_min", D(&DomainParams::y_min)},
        {"domain.y_max", D(&DomainParams::y_max)},
        {"domain.z_min", D(&DomainParams::z_min)},
        {"domain.z_max", D(&DomainParams::z_max)},
        {"physics.gravity", P(&PhysicsParams::gravity)},
        {"physics.particle_radius", P(&PhysicsParams::particle_radius)},
        {"physics.max_force", P(&PhysicsParams::max_force)},
        {"physics.floor_z", P(&PhysicsParams::floor_z)},
        {"physics.ceiling_z", P(&PhysicsParams::ceiling_z)},
        {"physics.wall_x_min", P(&PhysicsParams::wall_x_min)},
        {"physics.wall_x_max", P(&PhysicsParams::wall_x_max)},
        {"physics.wall_y_min", P(&PhysicsParams::wall_y_min)},
        {"physics.wall_y_max", P(&PhysicsParams::wall_y_max)},
        {"physics.restitution", P(&PhysicsParams::restitution)},
        {"physics.friction", P(&PhysicsParams::friction)},
        {"physics.tangential_model", S(&SimConfig::tangential_model, true)},
        {"physics.tangential_stiffness_ratio", F(&SimConfig::tangential_stiffness_ratio)},
        {"physics.normal_stiffness", F(&SimConfig::normal_stiffness)},
        {"materials.masses", L(&SimConfig::masses)},
        {"initialization.type", S(&SimConfig::init_type, true)},
        {"initialization.state_path", S(&SimConfig::init_state_path, true)},
        {"initialization.x0", F(&SimConfig::init_x0)},
        {"initialization.y0", F(&SimConfig::init_y0)},
        {"initialization.z0", F(&SimConfig::init_z0)},
        {"initialization.vx0", F(&SimConfig::init_vx0)},
        {"initialization.vy0", F(&SimConfig::init_vy0)},
        {"initialization.vz0", F(&SimConfig::init_vz0)},
        {"initialization.jitter", F(&SimConfig::init_jitter)},
        {"initialization.seed", U(&SimConfig::init_seed)},
        {"initialization.perturbation", F(&SimConfig::init_perturbation)},
        {"initialization.perturbation_seed", U(&SimConfig::init_perturbation_seed)},
        {"initialization.cube_length_x", F(&SimConfig::cube_length_x)},
        {"initialization.cube_length_y", F(&SimConfig::cube_length_y)},
        {"initialization.cube_length_z", F(&SimConfig::cube_length_z)},
        {"initialization.two_particle_separation", F(&SimConfig::two_particle_separation)},
        {"initialization.particle_material", I(&SimConfig::particle_material)},
        {"initialization.lattice_spacing", F(&SimConfig::lattice_spacing)},
        {"initialization.perturbation_distribution", S(&SimConfig::perturbation_distribution, true)},
        {"rendering.num_triangles", I(&SimConfig::num_triangles)},
        {"drivers.port", I(&SimConfig::port)},
        {"drivers.steps_per_frame", I(&SimConfig::steps_per_frame)},
        {"drivers.profiling_outer_iters", I(&SimConfig::profiling_outer_iters)},
        {"drivers.profiling_steps_per_iter", I(&SimConfig::profiling_steps_per_iter)},
        {"drivers.kernel_timing", B(&SimConfig::kernel_timing)},
        {"sweep.dt", L(&SimConfig::sweep_dt)},
        {"sweep.max_force", L(&SimConfig::sweep_max_force)},
        {"sweep.restitution", L(&SimConfig::sweep_restitution)},
        {"stability.mode", S(&SimConfig::stability_mode, true)},
        {"stability.sim_time", F(&SimConfig::stability_sim_time)},
        {"stability.acceleration_cutoff", F(&SimConfig::stability_acceleration_cutoff)},
        {"stability.steps_per_steadiness_check", I(&SimConfig::stability_steps_per_steadiness_check)},
        {"model_problem.sim_time", F(&SimConfig::model_problem_sim_time)},
        {"dataset.sim_time", F(&SimConfig::dataset_sim_time)},
        {"dataset.checkpoints", I(&SimConfig::dataset_checkpoints)},
        // The reference stores the solver name verbatim, quotes included
        // (src/sim_config.cpp:143), so a quoted name selects Picard.
        {"time_integration.solver", S(&SimConfig::time_integrator, false)},
        {"time_integration.picard_iterations", I(&SimConfig::picard_iterations)},
    };
    return table;
}

}  // namespace

SimConfig load_config(const std::string& path) {
    SimConfig cfg;
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "load_config: could not open '%s'; using built-in defaults.\n",
                     path.c_str());
        return cfg;
    }
    const auto& table = schema();
    std::string section, line;
    while (std::getline(in, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        line = strip(line);
        if (line.empty()) continue;
        const auto colon = line.find(':');
        if (colon == std::string::npos) {
            std::fprintf(stderr, "load_config: skipping malformed line '%s'\n", line.c_str());
            continue;
        }
        const std::string key = strip(line.substr(0, colon));
        const std::string value = strip(line.substr(colon + 1));
        if (value.empty()) {
            section = key;
            continue;
        }
        const auto it = table.find(section + "." + key);
        if (it == table.end()) {
            std::fprintf(stderr, "load_config: ignoring unknown key '%s.%s'\n",
                         section.c_str(), key.c_str());
            continue;
        }
        it->second(cfg, value);
    }
    return cfg;
}

extern "C" int wd_config_default(wd_config* out) {
    if (!out) return wd::fail(WD_ERR_ARG, "wd_config_default: null output");
    *out = world::to_wd_config(SimConfig{});
    return WD_OK;
}

extern "C" int wd_config_load(const char* path, wd_config* out) {
    if (!path || !out) return wd::fail(WD_ERR_ARG, "wd_config_load: null argument");
    try {
        *out = world::to_wd_config(load_config(path));
    } catch (const std::exception& e) {
        return wd::fail(WD_ERR_ARG, std::string("wd_config_load: ") + e.what());
    }
    return WD_OK;
}
