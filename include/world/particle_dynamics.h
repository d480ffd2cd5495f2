// particle_dynamics.h -- drop-in C++ class for the reference's ParticleDynamics
// (src/particle_dynamics.h:17-141), implemented over the C ABI of
// libworld_amd.so (include/world_amd.h).
//
// Same public members and methods, same semantics:
//   * host_state / host_material / host_mass / positions / host_density_grid
//     are host arrays the caller may read and write;
//   * device_state.copy_from_host(host_state) pushes a host edit (the way the
//     reference's tests perturb a running simulation);
//   * dt is a public member re-read by every take_step();
//   * time accumulates dt in FP32, one step at a time;
//   * GPU failures abort with a message, as the reference's CUDA_CHECK does.
// Header-only: link with -lworld_amd.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../world_amd.h"
#include "sim_config.h"

namespace world {

inline void check(int rc, const char* what) {
    if (rc != WD_OK) {
        std::fprintf(stderr, "world_amd error in %s: %s\n", what, wd_last_error());
        std::abort();
    }
}

}  // namespace world

template <typename T>
class DeviceVector;

// Host buffer of T (reference: src/host_vector.h:10-56).  Move-only.
template <typename T>
class HostVector {
public:
    HostVector() = default;
    explicit HostVector(size_t n) : v_(n) {}
    HostVector(const HostVector&) = delete;
    HostVector& operator=(const HostVector&) = delete;
    HostVector(HostVector&&) noexcept = default;
    HostVector& operator=(HostVector&&) noexcept = default;

    T& operator[](size_t i) { return v_[i]; }
    const T& operator[](size_t i) const { return v_[i]; }
    T* data() noexcept { return v_.data(); }
    const T* data() const noexcept { return v_.data(); }
    size_t size() const noexcept { return v_.size(); }

    inline void copy_from_device(const DeviceVector<T>& device);

private:
    std::vector<T> v_;
};

// Device-side array of the engine, addressed through the handle that owns it
// (reference: src/device_vector.h:13-67).  copy_from_host writes the array
// the engine holds for this member; the engine keeps grains in cell order
// internally, so there is no raw device pointer to hand out.
template <typename T>
class DeviceVector {
public:
    enum Kind { kState, kMaterial, kDensityGrid };
    DeviceVector() = default;
    DeviceVector(wd_sim** sim, Kind kind, size_t n) : sim_(sim), kind_(kind), size_(n) {}

    size_t size() const noexcept { return size_; }

    void copy_from_host(const HostVector<T>& host) {
        if (host.size() != size_) {
            std::fprintf(stderr, "DeviceVector::copy_from_host: size %zu != %zu\n", host.size(), size_);
            std::abort();
        }
        if (kind_ == kState) {
            world::check(wd_set_state(*sim_, reinterpret_cast<const float*>(host.data())),
                         "device_state.copy_from_host");
        } else if (kind_ == kMaterial) {
            world::check(wd_set_material(*sim_, reinterpret_cast<const int*>(host.data())),
                         "device_material.copy_from_host");
        } else {
            std::fprintf(stderr, "DeviceVector::copy_from_host: read-only member\n");
            std::abort();
        }
    }

    void read_into(T* out) const {
        if (kind_ == kState) {
            world::check(wd_get_state(*sim_, reinterpret_cast<float*>(out)), "copy_from_device");
        } else if (kind_ == kMaterial) {
            world::check(wd_get_material(*sim_, reinterpret_cast<int*>(out)), "copy_from_device");
        } else {
            world::check(wd_density_grid(*sim_, reinterpret_cast<float*>(out)), "copy_from_device");
        }
    }

private:
    wd_sim** sim_ = nullptr;
    Kind kind_ = kState;
    size_t size_ = 0;
};

template <typename T>
inline void HostVector<T>::copy_from_device(const DeviceVector<T>& device) {
    if (device.size() != size()) v_.resize(device.size());
    device.read_into(data());
}

class ParticleDynamics {
public:
    SimConfig config;

    float time = 0.f;
    float last_time = 0.f;
    std::chrono::steady_clock::time_point last_wall_clock_time;
    float real_time_ratio = 0.f;
    int n = 0;
    float dt = 0.f;
    int collision_grid_size_x = 0;
    int collision_grid_size_y = 0;
    int collision_grid_size_z = 0;
    int particles_per_cell = 0;

    std::vector<float> positions;
    HostVector<float> host_state;
    DeviceVector<float> device_state;
    HostVector<int> host_material;
    DeviceVector<int> device_material;
    HostVector<float> host_mass;

    HostVector<float> host_density_grid;
    int density_grid_size_x = 0;
    int density_grid_size_y = 0;
    int density_grid_size_z = 0;

    float time_unpack_state = 0.f;

    explicit ParticleDynamics(const SimConfig& config_ = SimConfig{}) : config(config_) {
        wd_ = world::to_wd_config(config);
        world::check(wd_create(&wd_, &sim_), "ParticleDynamics");
        world::check(wd_num_grains(sim_, &n), "ParticleDynamics");
        collision_grid_size_x = config.collision_grid_size_x;
        collision_grid_size_y = config.collision_grid_size_y;
        collision_grid_size_z = config.collision_grid_size_z;
        particles_per_cell = config.particles_per_cell;
        dt = config.dt;
        dt_sent_ = dt;
        density_grid_size_x = config.density_grid_size_x;
        density_grid_size_y = config.density_grid_size_y;
        density_grid_size_z = config.density_grid_size_z;
        host_mass = HostVector<float>(config.masses.size());
        for (size_t i = 0; i < config.masses.size(); ++i) host_mass[i] = config.masses[i];
        bind(n);
        host_material.copy_from_device(device_material);
        last_wall_clock_time = std::chrono::steady_clock::now();
        unpack_state();
    }

    ~ParticleDynamics() {
        if (sim_) wd_destroy(sim_);
    }
    ParticleDynamics(const ParticleDynamics&) = delete;
    ParticleDynamics& operator=(const ParticleDynamics&) = delete;

    // Rebuilds the simulation with new_n grains at rest at the origin
    // (reference: src/particle_dynamics.cu:97-106 reallocates the arrays).
    void resize(const int new_n) {
        std::vector<float> st(static_cast<size_t>(kStateStride) * new_n, 0.f);
        std::vector<int> mat(static_cast<size_t>(new_n), config.particle_material);
        recreate(new_n, st.data(), mat.data());
    }

    void unpack_state() {
        const auto t0 = std::chrono::steady_clock::now();
        host_state.copy_from_device(device_state);
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < kDim; ++a) positions[kDim * i + a] = host_state[kStateStride * i + a];
        }
        time_unpack_state = std::chrono::duration<float, std::milli>(
                std::chrono::steady_clock::now() - t0).count();
        real_time_ratio = get_real_time_ratio();
    }

    float get_real_time_ratio() {
        const auto now = std::chrono::steady_clock::now();
        const double wall = std::chrono::duration<double>(now - last_wall_clock_time).count();
        const float ratio = wall > 0.0 ? (time - last_time) / static_cast<float>(wall) : 0.f;
        last_time = time;
        last_wall_clock_time = now;
        return ratio;
    }

    float find_neighbors_wct() const { return stage(WD_STAGE_FIND_NEIGHBORS); }
    float compute_rhs_wct() const { return stage(WD_STAGE_COMPUTE_RHS); }
    float take_step_wct() const { return stage(WD_STAGE_TAKE_STEP); }
    float density_grid_wct() const { return stage(WD_STAGE_DENSITY_GRID); }

    // Host-side initializers (src/particle_dynamics.cu:128-235): they rebuild
    // the simulation with the chosen initial condition.
    void initialize_to_two_particles(const float x0, const float y0, const float z0) {
        reinit(WD_INIT_TWO_PARTICLES, x0, y0, z0);
    }
    void initialize_to_cube(const float x0, const float y0, const float z0) {
        reinit(WD_INIT_CUBE, x0, y0, z0);
    }
    void initialize_to_single_particle(const float x0, const float y0, const float z0) {
        reinit(WD_INIT_SINGLE_PARTICLE, x0, y0, z0);
    }
    void initialize_from_file(const std::string& path) {
        wd_config c = wd_;
        c.init_type = WD_INIT_FILE;
        std::snprintf(c.init_state_path, WD_PATH_MAX, "%s", path.c_str());
        int m = 0;
        if (wd_config_grain_count(&c, &m) != WD_OK) {
            std::fprintf(stderr, "initialize_from_file: %s\n", wd_last_error());
            std::exit(1);
        }
        std::vector<float> st(static_cast<size_t>(kStateStride) * m);
        std::vector<int> mat(static_cast<size_t>(m));
        if (wd_initial_state(&c, st.data(), mat.data(), m) != WD_OK) {
            std::fprintf(stderr, "initialize_from_file: %s\n", wd_last_error());
            std::exit(1);
        }
        recreate(m, st.data(), mat.data());
    }

    void take_step() {
        if (dt != dt_sent_) {
            world::check(wd_set_dt(sim_, dt), "take_step");
            dt_sent_ = dt;
        }
        world::check(wd_step(sim_, 1), "take_step");
        time += dt;
    }

    float compute_total_energy() {
        float e = 0.f;
        world::check(wd_total_energy(sim_, &e), "compute_total_energy");
        return e;
    }

    size_t compute_density_grid() {
        world::check(wd_density_grid(sim_, host_density_grid.data()), "compute_density_grid");
        return host_density_grid.size();
    }

    size_t density_grid_nodes() const {
        return static_cast<size_t>(density_grid_size_x) * static_cast<size_t>(density_grid_size_y) *
               static_cast<size_t>(density_grid_size_z);
    }

    // Copies the state to host_state and checks it is finite and in the box.
    bool is_stable() {
        host_state.copy_from_device(device_state);
        const DomainParams& d = config.domain;
        const float lo[kDim] = {d.x_min, d.y_min, d.z_min};
        const float hi[kDim] = {d.x_max, d.y_max, d.z_max};
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < kDim; ++a) {
                const float p = host_state[kStateStride * i + a];
                if (!std::isfinite(p) || p < lo[a] || p > hi[a]) return false;
            }
        }
        return true;
    }

    float compute_max_acceleration() {
        float a = 0.f;
        world::check(wd_max_accel(sim_, &a), "compute_max_acceleration");
        return a;
    }

    int grid_overflow_count() const {
        int c = 0;
        world::check(wd_overflow_count(sim_, &c), "grid_overflow_count");
        return c;
    }

    // The underlying C-ABI handle, for extensions (energy log, CIC log).
    wd_sim* handle() const { return sim_; }

private:
    wd_config wd_{};
    wd_sim* sim_ = nullptr;
    float dt_sent_ = 0.f;

    void bind(int count) {
        n = count;
        positions.assign(static_cast<size_t>(kDim) * n, 0.f);
        host_state = HostVector<float>(static_cast<size_t>(kStateStride) * n);
        host_material = HostVector<int>(static_cast<size_t>(n));
        device_state = DeviceVector<float>(&sim_, DeviceVector<float>::kState,
                                           static_cast<size_t>(kStateStride) * n);
        device_material = DeviceVector<int>(&sim_, DeviceVector<int>::kMaterial, static_cast<size_t>(n));
        host_density_grid = HostVector<float>(4 * density_grid_nodes());
    }

    void recreate(int count, const float* st, const int* mat) {
        wd_sim* fresh = nullptr;
        world::check(wd_create_from_state(&wd_, count, st, mat, &fresh), "ParticleDynamics");
        if (sim_) wd_destroy(sim_);
        sim_ = fresh;
        dt_sent_ = wd_.dt;
        bind(count);
        for (int i = 0; i < count; ++i) host_material[i] = mat[i];
        for (size_t k = 0; k < host_state.size(); ++k) host_state[k] = st[k];
        unpack_state();
    }

    void reinit(int type, float x0, float y0, float z0) {
        wd_config c = wd_;
        c.init_type = type;
        c.init_x0 = x0;
        c.init_y0 = y0;
        c.init_z0 = z0;
        int m = 0;
        world::check(wd_config_grain_count(&c, &m), "initialize");
        std::vector<float> st(static_cast<size_t>(kStateStride) * m);
        std::vector<int> mat(static_cast<size_t>(m));
        world::check(wd_initial_state(&c, st.data(), mat.data(), m), "initialize");
        recreate(m, st.data(), mat.data());
    }

    float stage(int k) const {
        float ms[4] = {0.f, 0.f, 0.f, 0.f};
        world::check(wd_stage_ms(sim_, ms), "stage timing");
        return ms[k];
    }
};
