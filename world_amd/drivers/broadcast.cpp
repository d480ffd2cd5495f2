// Counterpart of the reference's interactive driver (src/broadcast.cpp): a
// WebSocket server streaming the simulation to the browser client
// (client/world.js).  Text message "initialize" (re)builds the simulation
// and answers with six binary messages: int32 n; int32[3] collision grid
// size; int32 num_triangles; float particle_radius; float[6] domain bounds
// (x_min, x_max, y_min, y_max, z_min, z_max); float[3n] positions.  Text
// message "run" advances steps_per_frame steps and answers with one binary
// message float[3n + 6]: the positions, then sim time, real-time ratio and
// the milliseconds of neighbor search, force pass and step of the frame and
// of the state unpack.  A grid overflow ends the process (exit 1), as in
// the reference.  Clients are served one at a time, each in turn; every
// connection shares the one simulation.  drivers.port 0 takes a free port;
// the port is printed either way.
//
// Usage: broadcast [config.yaml]

#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "particle_dynamics.h"
#include "sim_config.h"
#include "ws_server.h"

namespace {

template <typename T>
bool send_pod(world_ws::Connection& c, const T* data, size_t count) {
    return c.send_binary(data, count * sizeof(T));
}

bool send_initial(world_ws::Connection& c, ParticleDynamics& sim) {
    const int32_t n = sim.n;
    const int32_t grid[3] = {sim.collision_grid_size_x, sim.collision_grid_size_y, sim.collision_grid_size_z};
    const int32_t triangles = sim.config.num_triangles;
    const float radius = sim.config.physics.particle_radius;
    const float bounds[6] = {sim.config.domain.x_min, sim.config.domain.x_max, sim.config.domain.y_min,
                             sim.config.domain.y_max, sim.config.domain.z_min, sim.config.domain.z_max};
    return send_pod(c, &n, 1) && send_pod(c, grid, 3) && send_pod(c, &triangles, 1) && send_pod(c, &radius, 1) &&
           send_pod(c, bounds, 6) && send_pod(c, sim.positions.data(), sim.positions.size());
}

bool send_frame(world_ws::Connection& c, ParticleDynamics& sim, int steps_per_frame) {
    // The stage clocks accumulate over the simulation's life: a frame's
    // times are differences.
    const float find0 = sim.find_neighbors_wct();
    const float rhs0 = sim.compute_rhs_wct();
    const float step0 = sim.take_step_wct();
    for (int s = 0; s < steps_per_frame; ++s) sim.take_step();
    // One host sync per frame, not per step.
    const int overflowed = sim.grid_overflow_count();
    if (overflowed > 0) {
        std::cerr << "world: collision grid overflowed particles_per_cell capacity in " << overflowed
                  << " cell-step(s) (dt=" << sim.dt << "); simulation has diverged. Exiting." << std::endl;
        std::exit(1);
    }
    const float t_find = sim.find_neighbors_wct() - find0;
    const float t_rhs = sim.compute_rhs_wct() - rhs0;
    const float t_step = sim.take_step_wct() - step0;
    sim.unpack_state();
    std::vector<float> payload(sim.positions);
    payload.push_back(sim.time);
    payload.push_back(sim.real_time_ratio);
    payload.push_back(t_find);
    payload.push_back(t_rhs);
    payload.push_back(t_step);
    payload.push_back(sim.time_unpack_state);
    return send_pod(c, payload.data(), payload.size());
}

}  // namespace

int main(int argc, char** argv) {
    const std::string config_path = argc > 1 ? argv[1] : "config.yaml";
    const SimConfig config = load_config(config_path);
    std::unique_ptr<ParticleDynamics> sim;

    world_ws::Listener listener(config.port);
    if (!listener.ok()) {
        std::cerr << "broadcast: cannot listen on port " << config.port << std::endl;
        return 1;
    }
    std::cout << "Listening on port " << listener.port() << std::endl;
    for (;;) {
        const int fd = listener.accept_fd();
        if (fd < 0) {
            std::perror("accept");
            return 1;
        }
        world_ws::Connection conn(fd);
        if (!conn.handshake()) continue;
        std::string message;
        world_ws::Opcode kind = world_ws::kText;
        while (conn.read_message(message, kind)) {
            std::cout << "Received message: " << message << std::endl;
            if (kind != world_ws::kText) continue;  // the reference acts on text messages only
            if (message == "initialize") {
                sim.reset();
                sim = std::make_unique<ParticleDynamics>(config);
                if (!send_initial(conn, *sim)) break;
            } else if (message == "run") {
                if (!sim) {  // before any "initialize": nothing to advance
                    std::cout << "Unknown message: " << message << " (no simulation yet)" << std::endl;
                    continue;
                }
                if (!send_frame(conn, *sim, config.steps_per_frame)) break;
            } else {
                std::cout << "Unknown message: " << message << std::endl;
            }
        }
    }
}
