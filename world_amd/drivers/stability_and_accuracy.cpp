// Counterpart of the reference's stability sweep (src/stability_and_accuracy.cpp):
// every (dt, max_force) pair of the config's "sweep:" axes runs in its own
// process, so that a run that dies (grid overflow, a device error) costs only
// that cell of the table.  A run is unstable (U) if its final state is
// outside the domain or not finite, or if its process fails; in
// "steady_state" mode it is steady (S) once max |a| drops below
// acceleration_cutoff * g (checked every steps_per_steadiness_check steps)
// and timed out (T) if that never happens within sim_time; in "fixed_time"
// mode a run that ends in bounds is S.  Results go to
// <out>/stability_results.dat ("dt max_force status energy" rows, status
// 0 = U, 1 = S, 2 = T) and a gnuplot script for the colored grid, run when
// gnuplot is installed.
//
// The parent never touches the GPU: it reads the config and starts this
// program again, in child mode, once per pair:
//   stability_and_accuracy --dt <dt> --max_force <f> --out <file> --config <yaml>
//
// Usage: stability_and_accuracy [config.yaml] [out_dir]

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include "driver_util.h"

namespace {

enum Outcome { kUnstable = 0, kSteady = 1, kTimedOut = 2 };

const char* letter(int s) { return s == kSteady ? "S" : (s == kTimedOut ? "T" : "U"); }

// A run whose grid overflowed particles_per_cell has diverged: it ends its
// process with one message, and the parent counts it unstable.
void exit_if_overflowed(const ParticleDynamics& sim, float dt, float max_force) {
    const int over = sim.grid_overflow_count();
    if (over > 0) {
        std::fprintf(stderr,
                     "stability_and_accuracy: dt=%.6g max_force=%.6g -- spatial grid overflowed "
                     "particles_per_cell capacity in %d cell-step(s); simulation has diverged. Exiting.\n",
                     static_cast<double>(dt), static_cast<double>(max_force), over);
        std::exit(1);
    }
}

// Child mode: one pair, "<energy> <status>" to `out`.  Stability is judged
// on the final state only (a transient excursion the walls push back is no
// failure); overflow and steadiness are checked every
// steps_per_steadiness_check steps, so the loop queues steps without a host
// sync in between.
int run_child(float dt, float max_force, const std::string& out, SimConfig config) {
    config.physics.max_force = max_force;
    ParticleDynamics sim(config);
    sim.dt = dt;
    const int steps = static_cast<int>(std::lround(static_cast<double>(config.stability_sim_time) /
                                                   static_cast<double>(dt)));
    const bool steady_mode = config.stability_mode == "steady_state";
    const int every = std::max(1, config.stability_steps_per_steadiness_check);
    const float cutoff = config.physics.gravity * config.stability_acceleration_cutoff;
    bool settled = false;
    for (int s = 0; s < steps; ++s) {
        sim.take_step();
        if (steady_mode && (s + 1) % every == 0) {
            exit_if_overflowed(sim, dt, max_force);
            if (sim.compute_max_acceleration() < cutoff) {
                settled = true;
                break;
            }
        }
    }
    exit_if_overflowed(sim, dt, max_force);
    int status = kSteady;
    if (!sim.is_stable()) status = kUnstable;
    else if (steady_mode) status = settled ? kSteady : kTimedOut;
    // The energy is reported for unstable runs too (NaN / inf when the state
    // blew up).
    const float energy = sim.compute_total_energy();
    std::FILE* f = std::fopen(out.c_str(), "w");
    if (!f) return 1;
    std::fprintf(f, "%.9g %d\n", static_cast<double>(energy), status);
    std::fclose(f);
    return 0;
}

struct Cell {
    int status = kUnstable;
    float energy = 0.f;
};

std::string self_path(const char* argv0) {
    char buf[4096];
    const ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (len <= 0) return argv0;
    buf[len] = '\0';
    return buf;
}

// One pair in a child process; a child that fails or leaves no result is
// unstable.
Cell run_pair(const std::string& self, float dt, float max_force, const std::string& config_path) {
    Cell cell;
    char tmpl[] = "/tmp/stability_and_accuracy_XXXXXX";
    const int fd = mkstemp(tmpl);
    if (fd < 0) {
        std::perror("mkstemp");
        std::exit(1);
    }
    close(fd);
    char dt_s[64], mf_s[64];
    std::snprintf(dt_s, sizeof(dt_s), "%.9g", static_cast<double>(dt));
    std::snprintf(mf_s, sizeof(mf_s), "%.9g", static_cast<double>(max_force));
    const pid_t pid = fork();
    if (pid < 0) {
        std::perror("fork");
        std::exit(1);
    }
    if (pid == 0) {
        // Keep a failing child's messages out of the live table.
        if (std::FILE* null = std::fopen("/dev/null", "w")) dup2(fileno(null), STDERR_FILENO);
        execl(self.c_str(), self.c_str(), "--dt", dt_s, "--max_force", mf_s, "--out", tmpl, "--config",
              config_path.c_str(), static_cast<char*>(nullptr));
        _exit(127);
    }
    int ws = 0;
    waitpid(pid, &ws, 0);
    if (WIFEXITED(ws) && WEXITSTATUS(ws) == 0) {
        if (std::FILE* in = std::fopen(tmpl, "r")) {
            int s = 0;
            if (std::fscanf(in, "%f %d", &cell.energy, &s) == 2 && s >= kUnstable && s <= kTimedOut) {
                cell.status = s;
            }
            std::fclose(in);
        }
    }
    std::remove(tmpl);
    return cell;
}

void write_results(const std::string& path, const std::vector<float>& dts, const std::vector<float>& forces,
                   const std::vector<std::vector<Cell>>& cells) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "Could not open '%s' for writing.\n", path.c_str());
        return;
    }
    std::fprintf(f, "# dt max_force status(0=unstable,1=steady,2=timed_out) energy\n");
    for (size_t i = 0; i < dts.size(); ++i) {
        for (size_t j = 0; j < forces.size(); ++j) {
            std::fprintf(f, "%.9g %.9g %d %.9g\n", static_cast<double>(dts[i]), static_cast<double>(forces[j]),
                         cells[i][j].status, static_cast<double>(cells[i][j].energy));
        }
    }
    std::fclose(f);
}

// A colored rectangle per pair on an index grid (the swept values need not
// be evenly spaced), tics labelled with the values.
void write_plot(const std::string& script, const std::string& png, const std::vector<float>& dts,
                const std::vector<float>& forces, const std::vector<std::vector<Cell>>& cells) {
    std::FILE* f = std::fopen(script.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "set terminal pngcairo size 900,700\nset output \"%s\"\n", png.c_str());
    std::fprintf(f, "set title \"Stability sweep: dt vs max_force\" noenhanced\n");
    std::fprintf(f, "set xlabel \"dt\" noenhanced\nset ylabel \"max_force\" noenhanced\nunset key\n");
    std::fprintf(f, "set xrange [-0.5:%g]\nset yrange [-0.5:%g]\n", static_cast<double>(dts.size()) - 0.5,
                 static_cast<double>(forces.size()) - 0.5);
    std::fprintf(f, "set xtics (");
    for (size_t i = 0; i < dts.size(); ++i) std::fprintf(f, "%s\"%g\" %zu", i ? ", " : "", static_cast<double>(dts[i]), i);
    std::fprintf(f, ") rotate by -45\nset ytics (");
    for (size_t j = 0; j < forces.size(); ++j) {
        std::fprintf(f, "%s\"%g\" %zu", j ? ", " : "", static_cast<double>(forces[j]), j);
    }
    std::fprintf(f, ")\n");
    static const char* color[] = {"red", "green", "yellow"};
    int obj = 1;
    for (size_t i = 0; i < dts.size(); ++i) {
        for (size_t j = 0; j < forces.size(); ++j) {
            std::fprintf(f, "set object %d rectangle from %g,%g to %g,%g fc rgb \"%s\" fs solid 1.0 noborder\n", obj++,
                         static_cast<double>(i) - 0.5, static_cast<double>(j) - 0.5, static_cast<double>(i) + 0.5,
                         static_cast<double>(j) + 0.5, color[cells[i][j].status]);
        }
    }
    std::fprintf(f, "plot NaN notitle\n");
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 9 && !std::strcmp(argv[1], "--dt") && !std::strcmp(argv[3], "--max_force") &&
        !std::strcmp(argv[5], "--out") && !std::strcmp(argv[7], "--config")) {
        return run_child(std::strtof(argv[2], nullptr), std::strtof(argv[4], nullptr), argv[6], load_config(argv[8]));
    }
    const std::string config_path = argc > 1 ? argv[1] : "stability/config.yaml";
    const std::string out_dir = argc > 2 ? argv[2] : "stability";
    const SimConfig config = load_config(config_path);
    const std::string self = self_path(argv[0]);
    const std::vector<float> dts = config.sweep_dt.empty() ? std::vector<float>{config.dt} : config.sweep_dt;
    const std::vector<float> forces =
            config.sweep_max_force.empty() ? std::vector<float>{config.physics.max_force} : config.sweep_max_force;

    std::printf("Legend: U = unstable, S = steady, T = timed out (bounded, never settled)\n\n%-12s", "dt \\ max_f");
    for (float mf : forces) std::printf("%-10g", static_cast<double>(mf));
    std::printf("\n");
    std::fflush(stdout);
    std::vector<std::vector<Cell>> cells(dts.size(), std::vector<Cell>(forces.size()));
    for (size_t i = 0; i < dts.size(); ++i) {
        std::printf("%-12g", static_cast<double>(dts[i]));
        std::fflush(stdout);
        for (size_t j = 0; j < forces.size(); ++j) {
            cells[i][j] = run_pair(self, dts[i], forces[j], config_path);
            std::printf("%-10s", letter(cells[i][j].status));
            std::fflush(stdout);
        }
        std::printf("\n");
    }
    std::printf("\n");

    world_drivers::ensure_dir(out_dir);
    const std::string data = out_dir + "/stability_results.dat";
    const std::string script = out_dir + "/stability_plot.gp";
    write_results(data, dts, forces, cells);
    write_plot(script, out_dir + "/stability_plot.png", dts, forces, cells);
    if (std::system(("gnuplot \"" + script + "\" 2>/dev/null").c_str()) != 0) {
        std::fprintf(stderr, "Note: gnuplot did not run; '%s' holds the plot script.\n", script.c_str());
    }
    std::printf("Wrote sweep data to '%s'.\n", data.c_str());
    return 0;
}
