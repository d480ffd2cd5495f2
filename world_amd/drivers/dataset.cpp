// Batched counterpart of the reference's dataset driver (src/dataset.cpp):
// one rollout per row of a design CSV on top of a base config, the CIC grids
// at K checkpoints written to <out>/grids.bin and one manifest row per
// rollout in <out>/manifest.csv, in the reference's byte layout.  Every row
// is a member of one batch handle, so all rollouts advance together; a row
// that diverges (over-capacity cells or a grain outside the domain) is
// flagged and zero-filled from that checkpoint on, as the reference does.
// --sequential runs the rows one after another instead, one
// ParticleDynamics each, exactly as the reference loop does; both modes
// write the same bytes.
//
// Usage: dataset [config.yaml] [design.csv] [out_dir] [--sequential]

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "driver_util.h"

namespace {

constexpr char kMagic[8] = {'R', 'T', 'P', 'G', 'R', 'D', '0', '1'};
constexpr int kChannels = 4;  // mass, then momentum per axis

struct Design {
    std::vector<std::string> names;
    std::vector<std::vector<float>> rows;
};

// Whitespace and a CR left by CRLF line endings are not part of a cell.
std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string& line) {
    std::vector<std::string> out;
    std::stringstream ss(line);
    std::string cell;
    while (std::getline(ss, cell, ',')) out.push_back(trim(cell));
    return out;
}

Design read_design(const std::string& path) {
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "dataset: could not open design '%s'.\n", path.c_str());
        std::exit(1);
    }
    Design d;
    std::string line;
    bool header = true;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        const std::vector<std::string> cells = split(line);
        if (header) {
            d.names = cells;
            header = false;
            continue;
        }
        std::vector<float> row;
        for (const std::string& c : cells) row.push_back(std::stof(c));
        if (row.size() != d.names.size()) {
            std::fprintf(stderr, "dataset: design row has %zu values, expected %zu\n", row.size(),
                         d.names.size());
            std::exit(1);
        }
        d.rows.push_back(row);
    }
    return d;
}

// The design columns the reference understands (src/dataset.cpp:93-118);
// anything else is an error, never silently ignored.
SimConfig config_for_row(const SimConfig& base, const Design& d, const std::vector<float>& row) {
    SimConfig c = base;
    for (size_t k = 0; k < d.names.size(); ++k) {
        const std::string& name = d.names[k];
        const float v = row[k];
        if (name == "throw_vx") c.init_vx0 = v;
        else if (name == "throw_vy") c.init_vy0 = v;
        else if (name == "throw_vz") c.init_vz0 = v;
        else if (name == "release_z") c.init_z0 = v;
        else if (name == "grain_mass") {
            for (size_t m = 1; m < c.masses.size(); ++m) c.masses[m] = v;  // every grain material
        } else if (name == "restitution") c.physics.restitution = v;
        else if (name == "friction") c.physics.friction = v;
        else {
            std::fprintf(stderr, "dataset: design column '%s' is not a known parameter\n", name.c_str());
            std::exit(1);
        }
    }
    return c;
}

struct Outcome {
    bool diverged = false;
    float time = 0.f;
    int overflow = 0;
    double deposited = 0.0;
};

}  // namespace

int main(int argc, char** argv) {
    const bool sequential = world_drivers::take_flag(argc, argv, "--sequential");
    const std::string config_path = argc > 1 ? argv[1] : "dataset/blob_throw.yaml";
    const std::string design_path = argc > 2 ? argv[2] : "dataset/design.csv";
    const std::string out_dir = argc > 3 ? argv[3] : "dataset";
    const SimConfig base = load_config(config_path);
    const Design design = read_design(design_path);
    const int R = static_cast<int>(design.rows.size());
    const int K = base.dataset_checkpoints;
    if (R == 0 || K <= 0) {
        std::fprintf(stderr, "dataset: nothing to do (%d rollouts, %d checkpoints)\n", R, K);
        return 1;
    }
    const std::vector<float> times = world_drivers::checkpoint_times(base.dataset_sim_time, K);
    world_drivers::ensure_dir(out_dir);

    const int nx = base.density_grid_size_x, ny = base.density_grid_size_y, nz = base.density_grid_size_z;
    const size_t nodes = static_cast<size_t>(nx) * ny * nz;
    const size_t per_grid = static_cast<size_t>(kChannels) * nodes;

    const std::string grid_path = out_dir + "/grids.bin";
    std::FILE* grids = std::fopen(grid_path.c_str(), "wb");
    if (!grids) {
        std::fprintf(stderr, "dataset: could not open '%s' for writing\n", grid_path.c_str());
        return 1;
    }
    std::fwrite(kMagic, 1, 8, grids);
    const int32_t header[7] = {R, K, kChannels, nx, ny, nz, static_cast<int32_t>(design.names.size())};
    std::fwrite(header, sizeof(int32_t), 7, grids);
    std::fwrite(times.data(), sizeof(float), times.size(), grids);
    const long grid_origin = std::ftell(grids);

    std::vector<SimConfig> configs;
    for (int r = 0; r < R; ++r) configs.push_back(config_for_row(base, design, design.rows[r]));
    const int steps = world_drivers::steps_per_checkpoint(base.dataset_sim_time, K, configs[0].dt);
    const std::vector<float> zeros(per_grid, 0.f);
    std::vector<Outcome> out(static_cast<size_t>(R));

    // Grid (r, k) goes to its place in the [rollout][checkpoint] array.
    auto put = [&](int r, int k, const float* g) {
        const long off = grid_origin + static_cast<long>((static_cast<size_t>(r) * K + k) * per_grid * sizeof(float));
        std::fseek(grids, off, SEEK_SET);
        std::fwrite(g, sizeof(float), per_grid, grids);
    };
    auto deposit = [&](const float* g) {
        double s = 0.0;
        for (size_t node = 0; node < nodes; ++node) s += g[node];
        return s;
    };

    if (sequential) {
        for (int r = 0; r < R; ++r) {
            ParticleDynamics sim(configs[r]);
            Outcome& o = out[r];
            for (int k = 0; k < K; ++k) {
                if (!o.diverged) {
                    for (int s = 0; s < steps; ++s) sim.take_step();
                    o.overflow = sim.grid_overflow_count();
                    if (o.overflow > 0 || !sim.is_stable()) o.diverged = true;
                }
                if (o.diverged) {
                    put(r, k, zeros.data());
                    continue;
                }
                sim.compute_density_grid();
                put(r, k, sim.host_density_grid.data());
                if (k == K - 1) o.deposited = deposit(sim.host_density_grid.data());
            }
            o.time = sim.time;
            std::printf("\rrollout %d/%d", r + 1, R);
            std::fflush(stdout);
        }
    } else {
        world_drivers::Batch batch(configs);
        std::vector<float> g;
        std::vector<int> over, stable;
        for (int k = 0; k < K; ++k) {
            batch.step(steps);
            batch.overflow(over);
            batch.stable(stable);
            batch.grids(g);
            for (int r = 0; r < R; ++r) {
                Outcome& o = out[r];
                if (!o.diverged) {
                    for (int s = 0; s < steps; ++s) o.time += configs[r].dt;  // the FP32 clock of a lone run
                    o.overflow = over[r];
                    if (o.overflow > 0 || !stable[r]) o.diverged = true;
                }
                const float* gr = g.data() + static_cast<size_t>(r) * per_grid;
                put(r, k, o.diverged ? zeros.data() : gr);
                if (!o.diverged && k == K - 1) o.deposited = deposit(gr);
            }
            std::printf("\rcheckpoint %d/%d", k + 1, K);
            std::fflush(stdout);
        }
    }
    std::printf("\n");
    std::fclose(grids);

    const std::string manifest_path = out_dir + "/manifest.csv";
    std::FILE* manifest = std::fopen(manifest_path.c_str(), "w");
    if (!manifest) {
        std::fprintf(stderr, "dataset: could not open '%s' for writing\n", manifest_path.c_str());
        return 1;
    }
    std::fprintf(manifest, "rollout");
    for (const std::string& name : design.names) std::fprintf(manifest, ",%s", name.c_str());
    std::fprintf(manifest, ",status,final_time,overflow_cells,deposited_mass\n");
    int ok = 0;
    for (int r = 0; r < R; ++r) {
        const Outcome& o = out[r];
        std::fprintf(manifest, "%d", r);
        for (float v : design.rows[r]) std::fprintf(manifest, ",%.9g", static_cast<double>(v));
        std::fprintf(manifest, ",%s,%.9g,%d,%.9g\n", o.diverged ? "diverged" : "ok",
                     static_cast<double>(o.time), o.overflow, o.deposited);
        ok += o.diverged ? 0 : 1;
    }
    std::fclose(manifest);
    std::printf("Wrote %s (%d rollouts x %d checkpoints x %d channels x %dx%dx%d)\n", grid_path.c_str(), R, K,
                kChannels, nx, ny, nz);
    std::printf("Wrote %s (%d ok, %d diverged)\n", manifest_path.c_str(), ok, R - ok);
    return 0;
}
