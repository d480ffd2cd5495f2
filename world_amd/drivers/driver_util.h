// Shared pieces of the batched driver counterparts (dataset, chaos,
// model_problem): checked C-ABI calls, a batch handle, and the checkpoint
// schedule the reference drivers use.
#pragma once

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "particle_dynamics.h"
#include "sim_config.h"

namespace world_drivers {

inline void ensure_dir(const std::string& path) {
    if (mkdir(path.c_str(), 0755) != 0 && errno != EEXIST) {
        std::perror(("mkdir " + path).c_str());
        std::exit(1);
    }
}

// Checkpoint k is at (k + 1) / K * total (src/dataset.cpp:141-148,
// src/chaos.cpp:94-99).
inline std::vector<float> checkpoint_times(float total, int k_count) {
    std::vector<float> t(static_cast<size_t>(k_count));
    for (int k = 0; k < k_count; ++k) {
        t[k] = total * static_cast<float>(k + 1) / static_cast<float>(k_count);
    }
    return t;
}

// Steps between checkpoints, rounded as the reference drivers round it
// (src/dataset.cpp:190-192, src/chaos.cpp:116-117).
inline int steps_per_checkpoint(float total, int k_count, float dt) {
    const int s = static_cast<int>(total / (static_cast<float>(k_count) * dt) + 0.5f);
    return s > 1 ? s : 1;
}

// One engine handle advancing several members together.
class Batch {
public:
    explicit Batch(const std::vector<SimConfig>& configs) {
        std::vector<wd_config> c;
        c.reserve(configs.size());
        for (const SimConfig& s : configs) c.push_back(world::to_wd_config(s));
        world::check(wd_batch_create(c.data(), static_cast<int>(c.size()), &sim_), "wd_batch_create");
        members_ = static_cast<int>(c.size());
        world::check(wd_num_grains(sim_, &n_), "wd_num_grains");
        first_.resize(members_);
        count_.resize(members_);
        for (int m = 0; m < members_; ++m) {
            world::check(wd_batch_member_range(sim_, m, &first_[m], &count_[m]), "wd_batch_member_range");
        }
        nodes_ = static_cast<size_t>(configs[0].density_grid_size_x) * configs[0].density_grid_size_y *
                 configs[0].density_grid_size_z;
    }
    ~Batch() {
        if (sim_) wd_destroy(sim_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    int members() const { return members_; }
    int first(int m) const { return first_[m]; }
    int count(int m) const { return count_[m]; }
    size_t nodes() const { return nodes_; }

    void step(int k) { world::check(wd_step(sim_, k), "wd_step"); }
    // State of every member back to back (6 floats per grain).
    void state(std::vector<float>& out) {
        out.resize(static_cast<size_t>(6) * n_);
        if (n_) world::check(wd_get_state(sim_, out.data()), "wd_get_state");
    }
    // Every member's CIC grid, [members][4][nodes].
    void grids(std::vector<float>& out) {
        out.resize(static_cast<size_t>(members_) * 4 * nodes_);
        world::check(wd_batch_density_grid(sim_, -1, out.data()), "wd_batch_density_grid");
    }
    void overflow(std::vector<int>& out) {
        out.resize(members_);
        world::check(wd_batch_overflow_counts(sim_, out.data()), "wd_batch_overflow_counts");
    }
    void stable(std::vector<int>& out) {
        out.resize(members_);
        world::check(wd_batch_is_stable(sim_, out.data()), "wd_batch_is_stable");
    }

private:
    wd_sim* sim_ = nullptr;
    int members_ = 0, n_ = 0;
    size_t nodes_ = 0;
    std::vector<int> first_, count_;
};

// True when argv holds the flag (and removes it).
inline bool take_flag(int& argc, char** argv, const char* flag) {
    for (int i = 1; i < argc; ++i) {
        if (std::string(argv[i]) == flag) {
            for (int j = i; j + 1 < argc; ++j) argv[j] = argv[j + 1];
            --argc;
            return true;
        }
    }
    return false;
}

}  // namespace world_drivers
