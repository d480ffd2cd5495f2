/*
 * world_amd.h -- C ABI of the MI355X (gfx950) granular-dynamics engine.
 *
 * One handle owns one simulation: its device buffers, one HIP stream and the
 * captured step graph.  Every entry point returns 0 on success and a negative
 * status on failure; wd_last_error() then returns a message for the calling
 * thread.  Host buffers are owned by the caller.  All arrays crossing this
 * boundary are in ORIGINAL grain order (the order the initializer produced or
 * the caller last wrote), even though the engine keeps grains in cell order
 * internally.
 *
 * State layout is the reference's flat FP32 [x, y, z, vx, vy, vz] per grain
 * with z up (src/sim_config.h:5-15).  The C++ class façade in
 * include/world/particle_dynamics.h drives these functions; the ctypes layer
 * in world_amd/_capi.py binds the same symbols for Python.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef WORLD_AMD_H
#define WORLD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WD_OK 0
#define WD_ERR_ARG (-1)      /* bad argument or configuration */
#define WD_ERR_HIP (-2)      /* HIP runtime failure */
#define WD_ERR_IO (-3)       /* file could not be read */
#define WD_ERR_STATE (-4)    /* call not valid in the handle's current state */
#define WD_ERR_DEVICE (-5)   /* a kernel reported an internal failure */

#define WD_MAX_MATERIALS 16
#define WD_PATH_MAX 512

/* Stage indices of wd_stage_ms (reference *_wct getters,
 * src/particle_dynamics.h:77-80). */
#define WD_STAGE_FIND_NEIGHBORS 0
#define WD_STAGE_COMPUTE_RHS 1
#define WD_STAGE_TAKE_STEP 2
#define WD_STAGE_DENSITY_GRID 3

enum wd_init_type {
    WD_INIT_CUBE = 0,
    WD_INIT_TWO_PARTICLES = 1,
    WD_INIT_SINGLE_PARTICLE = 2,
    WD_INIT_FILE = 3
};

enum wd_integrator { WD_SEMI_IMPLICIT_EULER = 0, WD_BACKWARD_EULER_PICARD = 1 };

/* Extension: tangential contact model.  VISCOUS is the reference's stateless
 * regularized Coulomb law (src/compute_rhs_kernel.cu:66-78); SPRING_HISTORY
 * adds a per-contact tangential spring whose displacement persists between
 * steps and is capped at the Coulomb limit. */
enum wd_tangential_model { WD_TANGENTIAL_VISCOUS = 0, WD_TANGENTIAL_SPRING_HISTORY = 1 };

/* Extension: distribution of initialization.perturbation. */
enum wd_perturbation_distribution { WD_PERTURB_UNIFORM = 0, WD_PERTURB_GAUSSIAN = 1 };

/*
 * Plain-data mirror of the reference SimConfig (src/sim_config.h:64-200),
 * restricted to what the solver reads, plus opt-in extension keys whose
 * defaults reproduce the reference.  Field names follow the YAML keys.
 */
typedef struct wd_config {
    /* simulation: */
    float dt;
    int collision_grid_size_x, collision_grid_size_y, collision_grid_size_z;
    int particles_per_cell;
    int threads_per_block;
    int density_grid_size_x, density_grid_size_y, density_grid_size_z;
    /* domain: */
    float x_min, x_max, y_min, y_max, z_min, z_max;
    /* physics: (DomainParams / PhysicsParams, src/sim_config.h:28-61) */
    float gravity, particle_radius, max_force;
    float floor_z, ceiling_z, wall_x_min, wall_x_max, wall_y_min, wall_y_max;
    float restitution, friction;
    /* physics extensions */
    int tangential_model;             /* wd_tangential_model, default VISCOUS */
    float tangential_stiffness_ratio; /* k_t / k_n for SPRING_HISTORY, default 2/7 */
    float normal_stiffness;           /* <= 0: k = max_force / particle_radius (reference) */
    /* materials: */
    int num_masses;
    float masses[WD_MAX_MATERIALS];
    /* initialization: */
    int init_type;                    /* wd_init_type */
    char init_state_path[WD_PATH_MAX];
    float init_x0, init_y0, init_z0;
    float init_vx0, init_vy0, init_vz0;
    float init_jitter;
    unsigned int init_seed;
    float init_perturbation;
    unsigned int init_perturbation_seed;
    float cube_length_x, cube_length_y, cube_length_z;
    float two_particle_separation;
    int particle_material;
    float lattice_spacing;            /* extension, default 1.01 (diameters) */
    int perturbation_distribution;    /* extension, wd_perturbation_distribution */
    /* time_integration: */
    int time_integrator;              /* wd_integrator */
    int picard_iterations;
    /* drivers: */
    int kernel_timing;
} wd_config;

typedef struct wd_sim wd_sim;

/* Message for the last failing call on this thread ("" if none). */
const char* wd_last_error(void);

/* Library version string. */
const char* wd_version(void);

/* sizeof(wd_config) as compiled into the library, so foreign-language
 * bindings can verify their struct layout. */
size_t wd_config_size(void);

/* Reference SimConfig{} defaults (src/sim_config.h:64-200). */
int wd_config_default(wd_config* out);

/* YAML-subset loader with the reference's semantics: starts from defaults,
 * overrides present keys, warns on unknown keys, returns defaults for a
 * missing file (load_config, src/sim_config.cpp:153-195). */
int wd_config_load(const char* path, wd_config* out);

/* Number of grains the configured initializer produces, without touching the
 * GPU (initialize_to_*, src/particle_dynamics.cu:128-235). */
int wd_config_grain_count(const wd_config* cfg, int* n_out);

/* Host-side initializer: writes the initial state (6n floats) and material
 * (n ints) exactly as the constructor would (src/particle_dynamics.cu:20-33).
 * `capacity` is the number of grains the buffers hold. */
int wd_initial_state(const wd_config* cfg, float* state_out, int* material_out, int capacity);

/* Constructs a simulation: initializes on the host, uploads, sizes every
 * buffer (ParticleDynamics::ParticleDynamics, src/particle_dynamics.cu:10-94).
 * A handle holds at most 268,435,455 (2^28 - 1) grains; more is WD_ERR_ARG. */
int wd_create(const wd_config* cfg, wd_sim** out);

/* Constructs a simulation from caller-supplied grains instead of the
 * configured initializer (state: 6n floats, material: n ids, original
 * order); the file initializer's path (src/particle_dynamics.cu:212-235)
 * and ParticleDynamics::resize (src/particle_dynamics.cu:97-106) use it. */
int wd_create_from_state(const wd_config* cfg, int n, const float* state, const int* material,
                         wd_sim** out);

/* Releases the handle (ParticleDynamics::~ParticleDynamics). */
int wd_destroy(wd_sim* sim);

int wd_num_grains(const wd_sim* sim, int* n_out);

/* Copies the state to the host, original order (unpack_state /
 * host_state.copy_from_device, src/particle_dynamics.cu:109-124). */
int wd_get_state(wd_sim* sim, float* state_out);

/* Overwrites the state from the host, original order
 * (device_state.copy_from_host, src/device_vector.h:59-61). */
int wd_set_state(wd_sim* sim, const float* state_in);

/* Per-grain material ids (host_material / device_material). */
int wd_get_material(wd_sim* sim, int* material_out);
int wd_set_material(wd_sim* sim, const int* material_in);

/* Time step read by the next wd_step (public ParticleDynamics::dt, read per
 * step at src/particle_dynamics.cu:268,276). */
int wd_set_dt(wd_sim* sim, float dt);
int wd_get_time(const wd_sim* sim, float* time_out);
int wd_set_time(wd_sim* sim, float time);

/* Advances nsteps steps of the configured integrator
 * (ParticleDynamics::take_step, src/particle_dynamics.cu:264-281).
 * Asynchronous: returns once the work is queued on the handle's stream. */
int wd_step(wd_sim* sim, int nsteps);

/* Waits for all queued work on the handle's stream. */
int wd_sync(wd_sim* sim);

/* Cloud-in-cell mass + momentum deposit, channel-major
 * [4][nx][ny][nz] float (compute_density_grid,
 * src/particle_dynamics.cu:289-293; src/density_grid_kernel.cu:25-106).
 * `out` receives 4 * density nodes floats. */
int wd_density_grid(wd_sim* sim, float* out);

/* Total kinetic + gravitational + elastic energy of the current state
 * (compute_total_energy, src/particle_dynamics.cu:296-308;
 * src/energy_kernel.cu:22-100). */
int wd_total_energy(wd_sim* sim, float* energy_out);

/* Extension: per-step energy log.  When capacity > 0, every subsequent step
 * records the total energy of the state it starts from (the value
 * compute_total_energy would return at that point), fused into the force pass.
 * wd_energy_log_read copies the first `count` entries recorded since the log
 * was enabled and returns how many exist in *recorded. */
int wd_energy_log_enable(wd_sim* sim, int capacity);
int wd_energy_log_read(wd_sim* sim, double* out, int count, int* recorded);

/* Extension: device-resident CIC checkpoints (the dataset driver's K
 * recorded grids, src/dataset.cpp:201-214, without a host round trip per
 * checkpoint).  wd_density_log_enable reserves `capacity` grids;
 * wd_density_log_record queues one deposit of the current state into the
 * next slot without waiting; wd_density_log_read copies the first `count`
 * recorded grids (4 * nodes floats each) and reports how many exist. */
int wd_density_log_enable(wd_sim* sim, int capacity);
int wd_density_log_record(wd_sim* sim);
int wd_density_log_read(wd_sim* sim, float* out, int count, int* recorded);

/* Measurement: runs nsteps steps with one HIP event pair around every kernel
 * launch on the handle's stream and writes the mean duration (ms) of each
 * kernel of a step, in launch order, to kernel_ms[0..*count).  `names`
 * (optional) receives a comma-separated list of the kernel names.
 * The state advances by nsteps steps. */
int wd_profile_step(wd_sim* sim, int nsteps, float* kernel_ms, int capacity, int* count,
                    char* names, size_t names_capacity);

/* Max |acceleration| from the last force pass (compute_max_acceleration,
 * src/particle_dynamics.cu:328-344). */
int wd_max_accel(wd_sim* sim, float* out);

/* Finite and inside the domain box (is_stable, src/particle_dynamics.cu:311-325). */
int wd_is_stable(wd_sim* sim, int* stable_out);

/* Lifetime count of (cell, neighbor pass) over-capacity incidents
 * (grid_overflow_count / FindNeighborsKernel::overflow_count,
 * src/find_neighbors_kernel.cu:230-234). */
int wd_overflow_count(wd_sim* sim, int* count_out);

/* Lifetime per-stage milliseconds (find_neighbors_wct, compute_rhs_wct,
 * take_step_wct, density_grid_wct; src/particle_dynamics.cu:350-358).
 * Waits for queued work.  The force pass integrates in the same kernel, so
 * its time is reported under COMPUTE_RHS and TAKE_STEP stays 0. */
int wd_stage_ms(wd_sim* sim, float out[4]);

/* ---- batched ensemble (extension) ---------------------------------------
 * The reference runs ensemble members one after another, each in its own
 * ParticleDynamics, and rejects members of unequal grain counts
 * (src/chaos.cpp:163-171; the dataset driver's row loop, src/dataset.cpp:
 * 189-232).  A batch handle advances E members with one launch sequence per
 * step.  Members may differ in their initialization (grain counts, seeds,
 * perturbation, initial velocity), physics.friction, physics.restitution and
 * the material masses; every other field must match.
 * Each member evolves exactly as it would alone: the same cell order, the
 * same force sums and therefore the same bits.
 *
 * Every single-simulation entry point also accepts a batch handle and then
 * works on the concatenation of the members: state, material and history
 * arrays hold member 0's grains, then member 1's, ... (grain ids are global,
 * offset by wd_batch_member_range's `first`); wd_total_energy and the energy
 * log sum over all members; wd_max_accel is the maximum over all members;
 * wd_overflow_count counts over-capacity cells of every member;
 * wd_density_grid returns member 0's grid; each density-log record holds the
 * E grids back to back. */
int wd_batch_create(const wd_config* configs, int members, wd_sim** sim_out);
int wd_batch_members(const wd_sim* sim, int* members_out);
/* Grains of member m are [first, first + count) of the concatenated arrays. */
int wd_batch_member_range(const wd_sim* sim, int member, int* first_out, int* count_out);
/* CIC grid (4 * nodes floats, src/particle_dynamics.cu:289-293) of one
 * member, or of all members back to back when member is -1. */
int wd_batch_density_grid(wd_sim* sim, int member, float* grid_out);
/* Per member: the lifetime over-capacity cell count
 * (grid_overflow_count, src/find_neighbors_kernel.cu:230-234) and whether
 * every grain is finite and inside the domain (is_stable,
 * src/particle_dynamics.cu:311-325).  One int per member each. */
int wd_batch_overflow_counts(wd_sim* sim, int* counts_out);
int wd_batch_is_stable(wd_sim* sim, int* stable_out);

/* ---- parameters and the discrete adjoint (extension) ---------------------
 * The reference has no differentiation.  These entry points give gradients
 * of a function of the state at the end of a window of semi-implicit Euler
 * steps with the spring_history tangential model, with respect to the
 * window's initial positions and velocities, the friction coefficient and
 * the normal stiffness.  The adjoint is the exact derivative of the forward
 * step with every contact set and clamp / stick-slip branch frozen to the
 * forward's (src/compute_rhs_kernel.cu:81-212 and
 * src/semi_implicit_euler_kernel.cu:4-16 are the functions differentiated,
 * with the tangential spring of this engine in place of
 * src/compute_rhs_kernel.cu:66-78). */

/* physics.friction of one member (0 for a single simulation) and
 * physics.normal_stiffness (all members); the damping constants follow k_n
 * (src/compute_rhs_kernel.cu:41-47).  Take effect at the next step. */
int wd_set_friction(wd_sim* sim, int member, float mu);
int wd_set_normal_stiffness(wd_sim* sim, float k_n);

/* Advances nsteps steps from the current state (as wd_step) and keeps a
 * checkpoint of the state and its springs every checkpoint_every steps.
 * The backward pass holds per-step slots for checkpoint_every + 1 steps
 * (cell-sorted state, cell starts, springs: about 420 B per grain each),
 * twice when the window has more than one segment (the next segment is
 * recomputed beside the current one's adjoints). */
int wd_adjoint_record(wd_sim* sim, int nsteps, int checkpoint_every);

/* Given the cotangent of the window's final CIC grids (members x 4 x nodes
 * floats, the layout of wd_batch_density_grid(sim, -1, ...); may be null) and
 * of its final state (6 floats per grain, original order; may be null),
 * writes dL/d(initial state) (6 floats per grain) and, per member,
 * {dL/dfriction, dL/dnormal_stiffness} (2 doubles per member).  Recomputes
 * each segment from its checkpoint; leaves the handle at the window's end.
 * Fails if a cell held more than particles_per_cell grains in the window. */
int wd_adjoint_backward(wd_sim* sim, const float* grid_cotangent, const float* state_cotangent,
                        float* grad_state0, double* grad_params);

/* ---- device buffers (extension; the reference has no counterpart) ---------
 * wd_set_state, wd_batch_density_grid and wd_adjoint_backward with the
 * caller's buffers in device memory of the handle's GPU (torch tensors, say)
 * instead of host memory, for a differentiable caller whose data stays on the
 * GPU; grad_params stays a host array.  Each call waits for the device first
 * (the caller's producers, on any stream) and returns with its results
 * complete.  WD_ERR_ARG for a pointer that is not device memory of the
 * handle's GPU. */
int wd_set_state_device(wd_sim* sim, const float* state_in);
int wd_batch_density_grid_device(wd_sim* sim, int member, float* grid_out);
int wd_adjoint_backward_device(wd_sim* sim, const float* grid_cotangent, const float* state_cotangent,
                               float* grad_state0, double* grad_params);

/* ---- launch configuration (extension) -----------------------------------
 * The force pass has six variants: one lane per grain with 64, 128 or 256
 * grains per block (lanes 0), or a one-wave block whose contacts are spread
 * over its lanes with 1, 2 or 4 lanes per grain (block 64, lanes 1 / 2 / 4).
 * By default the first wd_step of at least 18 steps tunes it online (three
 * round-robin rounds over the candidates, timed on the engine's stream); the
 * results are bit-identical for every choice.  The reference fixes
 * threads_per_block from the config (src/kernel.cu:31-38, blocks =
 * ceil(n / threads_per_block)). */
int wd_set_force_variant(wd_sim* sim, int block, int lanes); /* (0, 0): autotune */
int wd_get_force_variant(const wd_sim* sim, int* block_out, int* lanes_out); /* (0, 0) until tuned */
int wd_set_force_block(wd_sim* sim, int block); /* 0: autotune, or 64 / 128 / 256 with lanes 0 */
int wd_get_force_block(const wd_sim* sim, int* block_out); /* 0 until tuned */

/* The cell scan's tile (cells per tile of the single-pass scan of the cell
 * counts): 1,024, 2,048 or 4,096, tuned online on the same steps as the
 * force pass's variant (the cell starts do not depend on it); 0 returns it
 * to the tuner.  The reference scans nothing: it compacts occupied cells
 * with atomics (compact_cells_kernel, src/find_neighbors_kernel.cu:61-70). */
int wd_set_scan_tile(wd_sim* sim, int cells);
int wd_get_scan_tile(const wd_sim* sim, int* cells_out);

/* How a step bins its grains into cells.  Scatter: the count pass gives each
 * grain a slot in its cell, a scatter pass writes the ids in cell order and
 * the reorder ranks a grain among its cell's ids.  Buckets: the count pass
 * writes the id into its cell's bucket (8 per cell, the rest to a spill
 * list), and the reorder ranks from there: one launch fewer, scattered
 * writes.  The cell order, and every result, is the same either way; the
 * tuner times both on its steps (buckets are only allocated when they fit in
 * 1 GiB).  The reference bins with its own atomics-and-fill kernels
 * (src/find_neighbors_kernel.cu:43-161, 190-223). */
#define WD_BINNING_TUNED 0
#define WD_BINNING_SCATTER 1
#define WD_BINNING_BUCKETS 2
int wd_set_binning(wd_sim* sim, int mode);
int wd_get_binning(const wd_sim* sim, int* mode_out); /* WD_BINNING_SCATTER or _BUCKETS */

/* Fused steps (on by default for a single simulation; ensembles and the
 * kernel-timing mode run the classic sequence).  A step is then three
 * launches (four with scatter binning) instead of six: the force pass also
 * counts its grains into the next step's cells, on a grid of the last
 * occupied box widened by one cell per side; the cell scan's last block
 * reduces the pass's block partials (box, max |a|, energy, contacts) while
 * the others scan; the reorder completes the bin.  A grain that moves more
 * than one cell in a step is binned exactly by a one-block fallback inside
 * the scan; wd_fused_misses counts those steps.  Every result is the same bit
 * for bit either way; 0 selects the classic sequence.  The reference runs
 * its neighbour kernels, force kernel and integrator as separate launches
 * with a synchronization after each (src/particle_dynamics.cu:264-281,
 * src/find_neighbors_kernel.cu:190-223). */
int wd_set_fused_steps(wd_sim* sim, int on);

/* Graph replay of steps (on by default): 0 launches every kernel of every
 * step eagerly on the handle's stream (same kernels, same results), for
 * tools that follow kernel launches one by one.  The reference launches
 * eagerly and synchronizes after each kernel (src/kernel.cu:36). */
int wd_set_graph_replay(wd_sim* sim, int on);
int wd_get_fused_steps(const wd_sim* sim, int* on_out); /* 1 when steps run fused */
int wd_fused_misses(wd_sim* sim, unsigned long long* misses_out);

/* Extension, opt-in: collision cells coarser than the configured grid (the
 * reference sizes cells from the config, collision_grid_size_*,
 * src/sim_config.h:64-200).  Every axis gets floor(G / scale) cells, so a
 * cell is at least `scale` times the configured one; the neighbour search
 * still finds every contact (cells only grow) but visits them in another
 * order, so the FP32 force sums differ from the configured grid's in the
 * last bits: results then match the reference's FP32 arithmetic on the
 * coarser grid bit for bit, and the forces of the configured grid within
 * the FP32 rounding of their sums (1e-5 of the summed term magnitudes).
 * A coarser cell holds more grains: particles_per_cell may then be
 * exceeded.  wd_tune_cell_scale times `steps` real steps on each of the
 * scales 1, 1.25 and 1.5 (the first of each warming up) and keeps the
 * fastest whose steps overflowed no cell; it advances the simulation by
 * 3 * steps steps.  Scale 1 (the default) is the configured grid. */
int wd_set_cell_scale(wd_sim* sim, float scale);
int wd_get_cell_scale(const wd_sim* sim, float* scale_out, int* grid_out /* 3 ints, may be null */);
int wd_tune_cell_scale(wd_sim* sim, int steps, float* chosen_out);

/* Completed force passes and the contacts (wall and pair, counted from each
 * side) summed over them, since construction: contacts / (passes * n) is the
 * mean contact count per grain that the force pass's history traffic scales
 * with.  Extension; no counterpart in the reference. */
int wd_contact_stats(wd_sim* sim, unsigned long long* passes_out, unsigned long long* contacts_out);

/* ---- parity / inspection entry points ---------------------------------- */

/* Extension, tangential_model spring_history: the stored contact springs in
 * original grain order, WD_HISTORY_SLOTS per grain: partner[i*S + k] is the
 * partner's original index (walls: -1..-6 for x-lower, x-upper, y-lower,
 * y-upper, floor, ceiling), xi[(i*S + k)*3 + a] the spring displacement,
 * count[i] the number of valid slots.  The reference's friction is stateless
 * (src/compute_rhs_kernel.cu:50-65); this has no counterpart there. */
#define WD_HISTORY_SLOTS 24
int wd_get_history(wd_sim* sim, int* partner, float* xi, int* count);

/* Runs one neighbor pass on the current state and returns the cell order:
 * ids_out[p] = original index of the grain at cell-sorted position p, and
 * keys_out[p] = its flat collision cell (cx*Gy + cy)*Gz + cz
 * (flat_cell / particle_cell, src/find_neighbors_kernel.cu:28-40). */
int wd_cell_order(wd_sim* sim, int* ids_out, uint32_t* keys_out);

/* Extension with no counterpart in the reference: wd_cell_order for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same outputs as
 * wd_cell_order, staged through a fixed device array of
 * WD_NOALLOC_MAX_GRAINS ids and keys (8 MiB per GPU the library runs on,
 * held until the process ends) instead of per-call hipMalloc / hipFree.
 * Returns WD_ERR_ARG for a handle of more grains.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (threads, or
 * handles on one GPU driven concurrently) race. */
#define WD_NOALLOC_MAX_GRAINS (1 << 20)
int wd_cell_order_noalloc(wd_sim* sim, int* ids_out, uint32_t* keys_out);

/* Neighbor rows in the reference's flat format: n rows of
 * 27 * particles_per_cell ints, original indices, -1 terminated
 * (device_neighbors, src/particle_dynamics.h:48-52;
 * find_neighbors_kernel, src/find_neighbors_kernel.cu:119-161).
 * `capacity` is the number of ints `out` holds. */
int wd_neighbor_rows(wd_sim* sim, int* out, size_t capacity);

/* One force evaluation at the current state without integrating:
 * rhs = [v, F/m] per grain, original order (device_rhs written by
 * compute_rhs_kernel, src/compute_rhs_kernel.cu:81-212). */
int wd_rhs(wd_sim* sim, float* rhs_out);

#ifdef __cplusplus
}
#endif

#endif /* WORLD_AMD_H */
