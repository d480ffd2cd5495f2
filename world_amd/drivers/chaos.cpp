// Batched counterpart of the reference's chaos driver (src/chaos.cpp):
// member 0 runs unperturbed, members 1..E-1 with perturbation seed 1000 + e,
// and at every checkpoint each member's grain RMS displacement, relative
// mass-field difference and field centre-of-mass shift against member 0 go
// to <out>/divergence.csv; every member's final mass field goes to
// <out>/final_fields.bin (int32 E, int32 nodes, then E x nodes float32).
// All members advance in one batch handle; --sequential runs them one after
// another as the reference does.  Both modes write the same bytes.
//
// Usage: chaos [config.yaml] [members] [out_dir] [--sequential]

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "driver_util.h"

namespace {

// ||rho_e - rho_ref|| / ||rho_ref|| over the mass channel, accumulated as
// the reference does (the reference's own square in FP32, then summed in
// double).
double field_relative_difference(const std::vector<float>& ref, const float* cur, size_t nodes) {
    double num = 0.0, den = 0.0;
    for (size_t node = 0; node < nodes; ++node) {
        const double d = static_cast<double>(cur[node]) - ref[node];
        num += d * d;
        den += ref[node] * ref[node];
    }
    return den > 0.0 ? std::sqrt(num / den) : 0.0;
}

// Centre of mass in node units.
void centre_of_mass(const float* grid, int nx, int ny, int nz, double out[3]) {
    double total = 0.0;
    out[0] = out[1] = out[2] = 0.0;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                const double m = grid[static_cast<size_t>((ix * ny + iy) * nz + iz)];
                total += m;
                out[0] += m * ix;
                out[1] += m * iy;
                out[2] += m * iz;
            }
    if (total > 0.0) {
        for (int a = 0; a < 3; ++a) out[a] /= total;
    }
}

struct Row {
    double rms, field, com;
};

}  // namespace

int main(int argc, char** argv) {
    const bool sequential = world_drivers::take_flag(argc, argv, "--sequential");
    const std::string config_path = argc > 1 ? argv[1] : "chaos/ensemble.yaml";
    const int E = argc > 2 ? std::atoi(argv[2]) : 16;
    const std::string out_dir = argc > 3 ? argv[3] : "chaos";
    const SimConfig base = load_config(config_path);
    if (base.init_perturbation <= 0.0f) {
        std::fprintf(stderr,
                     "chaos: initialization.perturbation is 0, so every member would be identical.\n"
                     "Set it to the perturbation magnitude you want to study.\n");
        return 1;
    }
    if (E < 1) {
        std::fprintf(stderr, "chaos: need at least one member\n");
        return 1;
    }
    const int K = base.dataset_checkpoints;
    const std::vector<float> times = world_drivers::checkpoint_times(base.dataset_sim_time, K);
    world_drivers::ensure_dir(out_dir);

    std::vector<SimConfig> configs(static_cast<size_t>(E), base);
    configs[0].init_perturbation = 0.0f;  // member 0 is the unperturbed reference
    for (int e = 1; e < E; ++e) configs[e].init_perturbation_seed = static_cast<unsigned int>(1000 + e);

    const int nx = base.density_grid_size_x, ny = base.density_grid_size_y, nz = base.density_grid_size_z;
    const size_t nodes = static_cast<size_t>(nx) * ny * nz;
    const int steps = world_drivers::steps_per_checkpoint(base.dataset_sim_time, K, configs[0].dt);

    std::vector<std::vector<float>> ref_pos(static_cast<size_t>(K)), ref_field(static_cast<size_t>(K));
    std::vector<std::vector<Row>> rows(static_cast<size_t>(E), std::vector<Row>(static_cast<size_t>(K)));
    std::vector<std::vector<float>> final_field(static_cast<size_t>(E));
    int n = 0;

    // One member's checkpoint: positions (3 floats per grain) and mass field.
    auto record = [&](int e, int k, const float* pos, const float* field) {
        if (e == 0) {
            ref_pos[k].assign(pos, pos + 3 * static_cast<size_t>(n));
            ref_field[k].assign(field, field + nodes);
        } else {
            double sum_sq = 0.0;
            for (int i = 0; i < n; ++i) {
                double sq = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double d = static_cast<double>(pos[3 * i + a]) - ref_pos[k][3 * i + a];
                    sq += d * d;
                }
                sum_sq += sq;
            }
            Row& r = rows[e][k];
            r.rms = std::sqrt(sum_sq / n);
            r.field = field_relative_difference(ref_field[k], field, nodes);
            double cm[3], cr[3];
            centre_of_mass(field, nx, ny, nz, cm);
            centre_of_mass(ref_field[k].data(), nx, ny, nz, cr);
            double c2 = 0.0;
            for (int a = 0; a < 3; ++a) c2 += (cm[a] - cr[a]) * (cm[a] - cr[a]);
            r.com = std::sqrt(c2);
        }
        if (k == K - 1) final_field[e].assign(field, field + nodes);
    };

    std::vector<float> pos;
    if (sequential) {
        for (int e = 0; e < E; ++e) {
            ParticleDynamics sim(configs[e]);
            if (e == 0) n = sim.n;
            if (sim.n != n) {
                std::fprintf(stderr, "chaos: member %d has %d grains, reference had %d\n", e, sim.n, n);
                return 1;
            }
            for (int k = 0; k < K; ++k) {
                for (int s = 0; s < steps; ++s) sim.take_step();
                sim.unpack_state();
                sim.compute_density_grid();
                record(e, k, sim.positions.data(), sim.host_density_grid.data());
            }
        }
    } else {
        world_drivers::Batch batch(configs);
        n = batch.count(0);
        for (int e = 1; e < E; ++e) {
            if (batch.count(e) != n) {
                std::fprintf(stderr, "chaos: member %d has %d grains, reference had %d\n", e, batch.count(e), n);
                return 1;
            }
        }
        std::vector<float> st, g;
        for (int k = 0; k < K; ++k) {
            batch.step(steps);
            batch.state(st);
            batch.grids(g);
            for (int e = 0; e < E; ++e) {
                pos.resize(3 * static_cast<size_t>(n));
                const float* s = st.data() + 6 * static_cast<size_t>(batch.first(e));
                for (int i = 0; i < n; ++i)
                    for (int a = 0; a < 3; ++a) pos[3 * i + a] = s[6 * static_cast<size_t>(i) + a];
                record(e, k, pos.data(), g.data() + static_cast<size_t>(e) * 4 * nodes);
            }
        }
    }
    std::printf("reference: %d grains, %d checkpoints, %d steps each, grid %dx%dx%d\n", n, K, steps, nx, ny, nz);

    const std::string path = out_dir + "/divergence.csv";
    std::FILE* out = std::fopen(path.c_str(), "w");
    const std::string fields_path = out_dir + "/final_fields.bin";
    std::FILE* fields = std::fopen(fields_path.c_str(), "wb");
    if (!out || !fields) {
        std::fprintf(stderr, "chaos: could not open the outputs in '%s'\n", out_dir.c_str());
        return 1;
    }
    std::fprintf(out, "member,checkpoint,time,grain_rms_displacement,field_relative_difference,com_displacement\n");
    const int32_t header[2] = {E, static_cast<int32_t>(nodes)};
    std::fwrite(header, sizeof(int32_t), 2, fields);
    std::fwrite(final_field[0].data(), sizeof(float), nodes, fields);
    for (int e = 1; e < E; ++e) {
        for (int k = 0; k < K; ++k) {
            const Row& r = rows[e][k];
            std::fprintf(out, "%d,%d,%.9g,%.9g,%.9g,%.9g\n", e, k, static_cast<double>(times[k]), r.rms, r.field, r.com);
        }
        std::fwrite(final_field[e].data(), sizeof(float), nodes, fields);
    }
    std::fclose(out);
    std::fclose(fields);
    std::printf("Wrote %s (%d perturbed members, perturbation = %.3g x radius)\n", path.c_str(), E - 1,
                static_cast<double>(base.init_perturbation));
    return 0;
}
