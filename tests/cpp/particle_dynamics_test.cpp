// The reference's GoogleTest cases (test/particle_dynamics_test.cpp:16-285)
// re-expressed against the drop-in C++ class in include/world, with a small
// self-contained harness (GoogleTest is not available offline).  Each case
// uses the reference's configuration and tolerance.  Exit status 0 iff every
// case passed; one line per case on stdout.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "world/particle_dynamics.h"

namespace {

int g_failures = 0;
bool g_case_failed = false;

#define EXPECT_TRUE(c) check_((c), #c, __LINE__)
#define EXPECT_EQ(a, b) check_((a) == (b), #a " == " #b, __LINE__)
#define EXPECT_GT(a, b) check_((a) > (b), #a " > " #b, __LINE__)
#define EXPECT_GE(a, b) check_((a) >= (b), #a " >= " #b, __LINE__)
#define EXPECT_LE(a, b) check_((a) <= (b), #a " <= " #b, __LINE__)
#define EXPECT_LT(a, b) check_((a) < (b), #a " < " #b, __LINE__)
#define EXPECT_NEAR(a, b, tol) check_(std::fabs((double)(a) - (double)(b)) <= (double)(tol), \
                                      #a " ~ " #b, __LINE__)

void check_(bool ok, const char* what, int line) {
    if (!ok) {
        std::printf("    failed at line %d: %s\n", line, what);
        g_case_failed = true;
    }
}

void run(const char* name, const std::function<void()>& body) {
    g_case_failed = false;
    body();
    std::printf("%s %s\n", g_case_failed ? "FAIL" : "PASS", name);
    if (g_case_failed) ++g_failures;
}

}  // namespace

int main() {
    run("RunsStablyForFixedIterations", [] {
        ParticleDynamics sim;
        for (int i = 0; i < 100; ++i) sim.take_step();
        EXPECT_TRUE(sim.is_stable());
    });

    run("FreeFallMatchesKinematics", [] {
        ParticleDynamics sim;
        for (int i = 0; i < sim.n; ++i)
            for (int a = 0; a < kDim; ++a) sim.host_state[kStateStride * i + kDim + a] = 0.0f;
        sim.device_state.copy_from_host(sim.host_state);
        const float z0 = sim.host_state[2];
        const int steps = 100;
        for (int i = 0; i < steps; ++i) sim.take_step();
        sim.unpack_state();
        const float expected = -9.81f * sim.dt * sim.dt * static_cast<float>(steps) *
                               static_cast<float>(steps + 1) / 2.0f;
        EXPECT_NEAR(sim.positions[2] - z0, expected, 1e-6f);
    });

    run("ParticleDoesNotPenetrateFloor", [] {
        ParticleDynamics sim;
        const float start[kStateStride] = {0.5f, 0.5f, -0.5f, 0.0f, 0.0f, 0.0f};
        for (int a = 0; a < kStateStride; ++a) sim.host_state[a] = start[a];
        sim.device_state.copy_from_host(sim.host_state);
        for (int i = 0; i < 20000; ++i) sim.take_step();
        sim.unpack_state();
        EXPECT_GE(sim.positions[2], -1.0f + 0.01f - 1e-3f);
        EXPECT_TRUE(sim.is_stable());
    });

    run("RepeatedRunsAreDeterministic", [] {
        auto once = [](std::vector<float>& out) {
            ParticleDynamics sim;
            for (int i = 0; i < 50; ++i) sim.take_step();
            sim.unpack_state();
            out = sim.positions;
        };
        std::vector<float> a, b;
        once(a);
        once(b);
        EXPECT_EQ(a.size(), b.size());
        bool same = a.size() == b.size();
        for (size_t i = 0; same && i < a.size(); ++i) same = (a[i] == b[i]);
        EXPECT_TRUE(same);
    });

    run("GridOverflowIsCountedAndStaysInBounds", [] {
        SimConfig config;
        config.collision_grid_size_x = 1;
        config.collision_grid_size_y = 1;
        config.collision_grid_size_z = 1;
        config.particles_per_cell = 2;
        ParticleDynamics sim(config);
        EXPECT_GT(sim.n, config.particles_per_cell);
        EXPECT_EQ(sim.grid_overflow_count(), 0);
        sim.take_step();
        EXPECT_GT(sim.grid_overflow_count(), 0);
        sim.unpack_state();
        for (float c : sim.positions) EXPECT_TRUE(std::isfinite(c));
    });

    run("TotalEnergyNeverIncreases", [] {
        ParticleDynamics sim;
        float previous = sim.compute_total_energy();
        const float initial = previous;
        EXPECT_GT(initial, 0.0f);
        const float tolerance = 1e-3f * initial;
        for (int interval = 0; interval < 12; ++interval) {
            for (int step = 0; step < 100; ++step) sim.take_step();
            const float current = sim.compute_total_energy();
            EXPECT_LE(current, previous + tolerance);
            previous = current;
        }
        EXPECT_LT(previous, initial);
    });

    run("DensityGridConservesTotalMass", [] {
        SimConfig config;
        config.init_jitter = 0.3f;
        ParticleDynamics sim(config);
        for (int step = 0; step < 250; ++step) sim.take_step();
        sim.compute_density_grid();
        double deposited = 0.0;
        for (size_t node = 0; node < sim.density_grid_nodes(); ++node) deposited += sim.host_density_grid[node];
        sim.unpack_state();
        double expected = 0.0;
        for (int i = 0; i < sim.n; ++i) expected += sim.host_mass[sim.host_material[i]];
        EXPECT_GT(expected, 0.0);
        EXPECT_NEAR(deposited, expected, 1e-4 * expected);
    });

    run("DensityGridPutsGrainOnNodeAtSingleNode", [] {
        SimConfig config;
        config.init_type = "single_particle";
        config.density_grid_size_x = config.density_grid_size_y = config.density_grid_size_z = 5;
        config.init_x0 = -0.5f;
        config.init_y0 = 0.0f;
        config.init_z0 = 0.5f;
        ParticleDynamics sim(config);
        sim.compute_density_grid();
        const int expected_node = (1 * 5 + 2) * 5 + 3;
        const float m = sim.host_mass[config.particle_material];
        for (size_t node = 0; node < sim.density_grid_nodes(); ++node) {
            if (static_cast<int>(node) == expected_node) EXPECT_NEAR(sim.host_density_grid[node], m, 1e-6f * m);
            else EXPECT_NEAR(sim.host_density_grid[node], 0.0f, 1e-6f * m);
        }
    });

    run("DensityGridSplitsCellCenterEightWays", [] {
        SimConfig config;
        config.init_type = "single_particle";
        config.density_grid_size_x = config.density_grid_size_y = config.density_grid_size_z = 5;
        config.init_x0 = -0.75f;
        config.init_y0 = -0.25f;
        config.init_z0 = 0.25f;
        ParticleDynamics sim(config);
        sim.compute_density_grid();
        const float m = sim.host_mass[config.particle_material];
        int nonzero = 0;
        float total = 0.0f;
        for (size_t node = 0; node < sim.density_grid_nodes(); ++node) {
            const float value = sim.host_density_grid[node];
            total += value;
            if (value > 1e-9f) {
                ++nonzero;
                EXPECT_NEAR(value, m / 8.0f, 1e-5f * m);
            }
        }
        EXPECT_EQ(nonzero, 8);
        EXPECT_NEAR(total, m, 1e-5f * m);
    });

    run("DensityGridMomentumMatchesMassTimesVelocity", [] {
        SimConfig config;
        config.init_type = "single_particle";
        config.init_vx0 = 2.0f;
        config.init_vy0 = -3.0f;
        config.init_vz0 = 0.5f;
        ParticleDynamics sim(config);
        sim.compute_density_grid();
        const size_t nodes = sim.density_grid_nodes();
        const float m = sim.host_mass[config.particle_material];
        const float v[kDim] = {2.0f, -3.0f, 0.5f};
        for (int a = 0; a < kDim; ++a) {
            double momentum = 0.0;
            for (size_t node = 0; node < nodes; ++node) momentum += sim.host_density_grid[(a + 1) * nodes + node];
            EXPECT_NEAR(momentum, static_cast<double>(m) * v[a],
                        1e-4 * std::fabs(static_cast<double>(m) * v[a]) + 1e-9);
        }
    });

    run("CubeGrainsReceiveInitialVelocity", [] {
        SimConfig config;
        config.init_vx0 = 1.5f;
        config.init_vy0 = -0.5f;
        config.init_vz0 = 0.25f;
        config.physics.gravity = 0.0f;
        ParticleDynamics sim(config);
        EXPECT_GT(sim.n, 1);
        const float expected[kDim] = {1.5f, -0.5f, 0.25f};
        for (int i = 0; i < sim.n; ++i)
            for (int a = 0; a < kDim; ++a) EXPECT_EQ(sim.host_state[kStateStride * i + kDim + a], expected[a]);
        std::vector<float> before(sim.positions.begin(), sim.positions.end());
        const int steps = 50;
        for (int step = 0; step < steps; ++step) sim.take_step();
        sim.unpack_state();
        const float elapsed = sim.dt * static_cast<float>(steps);
        for (int a = 0; a < kDim; ++a) {
            const float moved = sim.positions[a] - before[a];
            EXPECT_NEAR(moved, expected[a] * elapsed, 1e-4f * std::fabs(expected[a] * elapsed) + 1e-6f);
        }
    });

    run("ConfigLoaderAndDtAssignment", [] {
        // Drivers assign sim.dt after construction (src/stability_and_accuracy.cpp:71).
        SimConfig config = load_config("/nonexistent/path.yaml");
        config.physics.gravity = 9.81f;
        ParticleDynamics sim(config);
        sim.dt = 2e-4f;
        for (int i = 0; i < 10; ++i) sim.take_step();
        EXPECT_NEAR(sim.time, 10 * 2e-4f, 1e-9f);
    });

    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
