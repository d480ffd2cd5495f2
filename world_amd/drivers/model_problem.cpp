// Counterpart of the reference's model_problem driver (src/model_problem.cpp):
// one small simulation per (dt, max_force, restitution) combination of the
// config's "sweep:" axes, every step's full state recorded, and all combos
// written to <out>/history.dat as gnuplot "index" blocks (a "# combo" header
// each, two blank lines between them), in the reference's text layout.  A
// plot script for particle 0's height is written beside it and handed to
// gnuplot when one is installed.
//
// Usage: model_problem [config.yaml] [out_dir]

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "driver_util.h"

namespace {

struct Combo {
    float dt, max_force, restitution;
    std::vector<float> time;
    std::vector<std::vector<float>> state;
};

Combo run_combo(const SimConfig& base, float dt, float max_force, float restitution) {
    SimConfig c = base;
    c.physics.max_force = max_force;
    c.physics.restitution = restitution;
    ParticleDynamics sim(c);
    sim.dt = dt;
    const int steps = static_cast<int>(std::lround(static_cast<double>(base.model_problem_sim_time) /
                                                   static_cast<double>(dt)));
    Combo combo{dt, max_force, restitution, {}, {}};
    auto row = [&]() {
        combo.time.push_back(sim.time);
        combo.state.emplace_back(sim.host_state.data(), sim.host_state.data() + kStateStride * sim.n);
    };
    row();  // the initial condition (the constructor unpacked it)
    for (int s = 0; s < steps; ++s) {
        sim.take_step();
        sim.unpack_state();
        row();
    }
    const int over = sim.grid_overflow_count();
    if (over > 0) {
        std::fprintf(stderr,
                     "model_problem: dt=%.6g max_force=%.6g restitution=%.6g -- spatial grid overflowed "
                     "particles_per_cell capacity in %d cell-step(s); simulation has diverged.\n",
                     static_cast<double>(dt), static_cast<double>(max_force), static_cast<double>(restitution),
                     over);
    }
    return combo;
}

void write_history(const std::string& path, int n, const std::vector<Combo>& combos) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Could not open '%s' for writing.\n", path.c_str());
        return;
    }
    std::fprintf(f, "# step time");
    for (int i = 0; i < n; ++i) std::fprintf(f, " x%d y%d z%d vx%d vy%d vz%d", i, i, i, i, i, i);
    std::fprintf(f, "\n");
    for (size_t c = 0; c < combos.size(); ++c) {
        const Combo& cb = combos[c];
        std::fprintf(f, "# combo %zu: dt=%.9g max_force=%.9g restitution=%.9g\n", c, static_cast<double>(cb.dt),
                     static_cast<double>(cb.max_force), static_cast<double>(cb.restitution));
        for (size_t s = 0; s < cb.time.size(); ++s) {
            std::fprintf(f, "%zu %.9g", s, static_cast<double>(cb.time[s]));
            for (float v : cb.state[s]) std::fprintf(f, " %.9g", static_cast<double>(v));
            std::fprintf(f, "\n");
        }
        if (c + 1 < combos.size()) std::fprintf(f, "\n\n");  // gnuplot's index separator
    }
    std::fclose(f);
}

void write_plot(const std::string& script, const std::string& png, const std::string& data,
                const std::vector<Combo>& combos, float contact_z) {
    static const char* palette[] = {"purple", "#1f77b4", "#ff7f0e", "#2ca02c", "#8c564b",
                                    "#e377c2", "#7f7f7f", "#bcbd22", "#17becf", "#000000"};
    std::FILE* f = std::fopen(script.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "set terminal pngcairo size 1000,750 noenhanced\nset output \"%s\"\n", png.c_str());
    std::fprintf(f, "set title \"Particle 0 trajectory\"\nset xlabel \"time (s)\"\nset ylabel \"z position\"\n");
    std::fprintf(f, "set key outside right\nplot ");
    for (size_t c = 0; c < combos.size(); ++c) {
        std::fprintf(f, "%s\"%s\" index %zu using 2:5 with linespoints lw 2 lc rgb \"%s\" pt 7 ps 0.5 "
                        "title \"dt=%.4g, max_force=%.4g, e=%.3g\", \\\n",
                     c == 0 ? "" : "     ", data.c_str(), c, palette[c % 10], static_cast<double>(combos[c].dt),
                     static_cast<double>(combos[c].max_force), static_cast<double>(combos[c].restitution));
    }
    std::fprintf(f, "     %.9g with lines lc rgb \"red\" dashtype 2 lw 2 title \"floor contact\"\n",
                 static_cast<double>(contact_z));
    std::fclose(f);
    if (std::system(("gnuplot \"" + script + "\" 2>/dev/null").c_str()) != 0) {
        std::fprintf(stderr, "Note: gnuplot did not run; '%s' holds the plot script.\n", script.c_str());
    }
}

}  // namespace

int main(int argc, char** argv) {
    const std::string config_path = argc > 1 ? argv[1] : "model_problems/single_particle.yaml";
    const std::string out_dir = argc > 2 ? argv[2] : "model_problems";
    const SimConfig config = load_config(config_path);
    const std::vector<float> dts = config.sweep_dt.empty() ? std::vector<float>{config.dt} : config.sweep_dt;
    const std::vector<float> forces =
            config.sweep_max_force.empty() ? std::vector<float>{config.physics.max_force} : config.sweep_max_force;
    const std::vector<float> rests = config.sweep_restitution.empty()
                                             ? std::vector<float>{config.physics.restitution}
                                             : config.sweep_restitution;
    std::vector<Combo> combos;
    for (float dt : dts)
        for (float f : forces)
            for (float e : rests) combos.push_back(run_combo(config, dt, f, e));

    world_drivers::ensure_dir(out_dir);
    const std::string data = out_dir + "/history.dat";
    const int n = static_cast<int>(combos[0].state[0].size() / kStateStride);
    write_history(data, n, combos);
    write_plot(out_dir + "/trajectory_plot.gp", out_dir + "/trajectory_plot.png", data, combos,
               config.physics.floor_z + config.physics.particle_radius);
    size_t rows = 0;
    for (const Combo& c : combos) rows += c.time.size();
    std::printf("Ran %zu combo(s), %zu total step(s), to '%s'.\n", combos.size(), rows, data.c_str());
    return 0;
}
