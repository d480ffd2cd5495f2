// Host side of the engine: one wd_sim per simulation, its device buffers,
// its HIP stream, graph replay of the step, and the C-ABI entry points.
//
// Construction, state access and diagnostics follow the reference solver
// object (src/particle_dynamics.cu:10-360); the step pipeline is the one
// described in kernels.hip.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <array>
#include <vector>

#include "../../include/world_amd.h"
#include "capi_util.h"
#include "kernels.h"

// The kernels are compiled in this translation unit.
#include "kernels.hip"
#include "adjoint.hip"

using namespace wd;

static_assert(WD_HISTORY_SLOTS == kHist, "history slot count differs between ABI and kernels");

namespace {

#define WD_HIP(expr)                                                                        \
    do {                                                                                    \
        const hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                             \
            return wd::fail(WD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                   \
    } while (0)

// Host evaluation of the restitution -> dashpot relation
// (src/compute_rhs_kernel.cu:41-47), in the reference's FP32 order.
float damping(float e, float k, float m) {
    const float pi = 3.14159265358979323846f;
    if (e <= 0.f) return 2.f * std::sqrt(k * m);
    if (e >= 1.f) return 0.f;
    const float ln_e = std::log(e);
    return -2.f * std::sqrt(m * k) * ln_e / std::sqrt(pi * pi + ln_e * ln_e);
}

struct TimedStage {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

struct wd_sim {
    wd_config cfg{};
    int n = 0;
    int device = 0;               // the HIP device current when the handle was created
    float cell_scale = 1.f;       // collision cells this many times the configured ones (per axis, at least)
    float time = 0.f;
    hipStream_t stream = nullptr;
    Params hp{};
    Params* dp = nullptr;
    Ctl* ctl = nullptr;
    float4 *Tp = nullptr, *Tv = nullptr, *Sp = nullptr, *Sv = nullptr;
    float4* NP[2] = {nullptr, nullptr};
    float4* NV[2] = {nullptr, nullptr};
    float4* H[2] = {nullptr, nullptr};  // contact-history generations [kHist][n], by T position
    int* Hc[2] = {nullptr, nullptr};
    int* hidx[2] = {nullptr, nullptr};  // spring index of each cell-sorted grain (two Picard buffers)
    int hidx_cur = 0;                   // hidx buffer the last bin wrote
    uint32_t *key = nullptr, *slot = nullptr, *count = nullptr, *start = nullptr;
    int* tmp_id = nullptr;
    // Bucket binning (no scatter pass): kBucket ids per cell and two spill
    // lists of n (cell, id) entries, allocated unless the buckets would pass
    // 1 GiB.  Which binning runs is tuned online with the scan tile (the
    // bucket writes are scattered: they pay at 32K grains, not at 1M); the
    // cell order is the same either way.
    uint32_t* bucket = nullptr;
    uint2* spill = nullptr;
    bool bucket_ok = false;
    bool bucket_bin = false;
    bool bin_fixed = false;       // set by wd_set_binning
    double bin_ms[2] = {};        // last tuning's bin time (count, scatter, reorder) per mode
    unsigned long long* status = nullptr;
    size_t scan_cells = 0, ntiles_max = 0;  // cells allocated (whole largest tiles); most tiles
    size_t scan_cells_used = 0;              // cells + 1 of the largest possible sub-grids
    float* st6 = nullptr;
    int* mat_d = nullptr;  // original-order (member << 4 | material) codes
    // ensemble members (1 for a single simulation)
    int members = 1;
    std::vector<int> goff;             // grain id offsets, members + 1
    std::vector<wd_config> mcfg;       // per-member configuration
    int* d_mgoff = nullptr;
    int* d_mbox = nullptr;
    uint32_t* d_mcell = nullptr;
    // Fused steps (a lone simulation): the pending and next grids of the
    // force pass's count (6 box ints and 2 cell offsets each).
    int* d_pbox = nullptr;
    uint32_t* d_pcell = nullptr;
    int* d_nbox = nullptr;
    uint32_t* d_ncell = nullptr;
    // A fused step is [force pass that also counts its grains into the
    // pending grid, cell scan that also reduces the pass's partials,
    // (scatter,) reorder]: three or four launches instead of six.  It needs
    // the state binned with the current grid on entry and leaves it so;
    // fused_ready records that (a classic bin restores it).
    bool fused_on = true;         // wd_set_fused_steps
    bool fused_ready = false;
    FinSlice* fin_slices = nullptr;  // the scan's slices of the partials, combined by the reorder
    float* d_mmu = nullptr;
    unsigned* d_mscale = nullptr;
    float* d_mtab = nullptr;                 // per-member material tables (members > 1)
    unsigned long long* d_movf = nullptr;    // per-member over-capacity cells
    size_t nodes = 0;                  // CIC nodes per member channel
    int nblocks = 0;
    BlockPart* parts = nullptr;
    BlockMembers* pmem = nullptr;   // member part of the partials (ensembles only)
    double* energy_log = nullptr;
    double* ewave = nullptr;      // per-kEnergyGroup-grain energy sums of the last energy pass
    int force_var = -1;           // index into kForceVariants; -1 until the autotuner has chosen
    int scan_var = 1;             // index into kScanItemsVariants (2,048-cell tiles until tuned)
    bool scan_fixed = false;      // set by wd_set_scan_tile
    bool force_fixed = false;     // set by wd_set_force_block / wd_set_force_variant
    long long steps_run = 0;      // steps taken by this handle
    int* bin_box_out = nullptr;   // where k_count copies the boxes it bins with (adjoint recompute)
    uint32_t* bin_cell_out = nullptr;
    long long next_tune = 0;      // step count at which the autotuner runs (again)
    double tune_ms[sizeof(kForceVariants) / sizeof(kForceVariants[0])] = {};  // last tuning's time per variant (force pass + its reduction)
    double scan_ms[3] = {};       // last tuning's cell-scan time per tile size
    int energy_cap = 0;
    bool energy_on = false;
    unsigned long long* cic_acc = nullptr;
    float* cic_out = nullptr;
    size_t cic_count = 0;
    float* rhs_buf = nullptr;
    float* cic_log = nullptr;
    int cic_log_cap = 0, cic_log_count = 0;
    struct ProfRec {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<ProfRec>* prof = nullptr;  // set while wd_profile_step runs
    // Step graphs of 1, 2, 4, 8 and 16 steps: a call of k steps replays
    // k / 16 sixteen-step graphs and one graph per set bit of k mod 16, so a
    // short call costs at most five launches however it is split.
    static constexpr int kGraphs = 5;
    hipGraphExec_t graph_pow[kGraphs] = {};
    hipGraphExec_t graph_cic_log = nullptr;  // one logged deposit (captured with the log buffer)
    bool timing = false;
    bool graphs_on = true;        // wd_set_graph_replay: off launches every step's kernels eagerly
    float stage_ms[4] = {0.f, 0.f, 0.f, 0.f};
    std::vector<TimedStage> pending;
    std::vector<hipEvent_t> spare;

    // Discrete adjoint (extension): checkpoints of the recorded window and
    // the cotangent buffers, all owned by `owned`.
    struct Snap {
        float4* P = nullptr;  // positions by original id
        float4* V = nullptr;  // velocities by original id
        float4* H = nullptr;  // springs [kHist][n] by original id
        int* C = nullptr;     // spring counts by original id
    };
    // The bin of one recomputed step and the springs it read: what the
    // adjoint of that step consumes.
    struct BinSlot {
        float4 *Sp = nullptr, *Sv = nullptr;
        uint32_t* start = nullptr;
        int* hidx = nullptr;
        float4* H = nullptr;  // springs read by the step [kHist][n]
        int* Hc = nullptr;
        int* mbox = nullptr;
        uint32_t* mcell = nullptr;
    };
    struct AdjointState {
        int nsteps = 0, every = 0, nseg = 0;
        unsigned long long overflow0 = 0;
        std::vector<Snap> ckpt;
        // Slot sets of every + 1 slots (the last receives the final springs):
        // one set, or two when the backward sweep overlaps the recompute of
        // the next segment with the contact adjoints of the current one.
        std::vector<BinSlot> slot;
        int sets = 1;
        BinSlot& at(int set, int k) { return slot[static_cast<size_t>(set) * (every + 1) + k]; }
        hipStream_t rstream = nullptr;       // recompute stream of the overlapped sweep
        hipEvent_t ev_start = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_swp[2] = {nullptr, nullptr};
        float4 *lx = nullptr, *lv = nullptr, *own_x = nullptr, *own_v = nullptr;
        float4 *lx2 = nullptr, *lv2 = nullptr;  // dL/d(state) of a segment's last step, in its cell order
        float4 *lxi[2] = {nullptr, nullptr}, *cg_d = nullptr, *cg_v = nullptr;
        int* cg_cnt = nullptr;
        int* cg_p = nullptr;  // partners of the stored pair contacts, grain-major
        double2* gparam = nullptr;
        float *grid_cot = nullptr, *state_cot = nullptr;
        std::vector<void*> owned;
        // Captured segments (recompute + contact adjoints), keyed by
        // {steps, spring parity, lxi buffer, force variant}; with two slot
        // sets, recomputes keyed by {steps, parity, set, variant} and
        // contact adjoints by {steps, lxi buffer, set, -1}.
        std::vector<std::pair<std::array<int, 4>, hipGraphExec_t>> seg_graphs, rec_graphs, swp_graphs;
    };
    AdjointState* adj = nullptr;

    void drop_adjoint() {
        if (!adj) return;
        for (auto& g : adj->seg_graphs) (void)hipGraphExecDestroy(g.second);
        for (auto& g : adj->rec_graphs) (void)hipGraphExecDestroy(g.second);
        for (auto& g : adj->swp_graphs) (void)hipGraphExecDestroy(g.second);
        for (hipEvent_t e : {adj->ev_start, adj->ev_rec[0], adj->ev_rec[1], adj->ev_swp[0], adj->ev_swp[1]}) {
            if (e) (void)hipEventDestroy(e);
        }
        if (adj->rstream) (void)hipStreamDestroy(adj->rstream);
        for (void* b : adj->owned) (void)hipFree(b);
        delete adj;
        adj = nullptr;
    }

    ~wd_sim() { release(); }

    void release() {
        drop_adjoint();
        for (auto& g : graph_pow) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
        if (graph_cic_log) (void)hipGraphExecDestroy(graph_cic_log);
        graph_cic_log = nullptr;
        for (auto& t : pending) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        pending.clear();
        for (auto e : spare) (void)hipEventDestroy(e);
        spare.clear();
        void* bufs[] = {dp, ctl, Tp, Tv, Sp, Sv, NP[0], NP[1], NV[0], NV[1], H[0], H[1], Hc[0],
                        Hc[1], key, slot, count,
                        start, tmp_id, bucket, spill, status, st6, mat_d, parts, energy_log, cic_acc,
                        cic_out, rhs_buf, cic_log, d_mgoff, d_mbox, d_mcell, d_mmu, d_mscale, ewave,
                        hidx[0], hidx[1], d_mtab, d_movf, pmem, d_pbox, d_pcell, d_nbox, d_ncell,
                        fin_slices};
        for (void* b : bufs) {
            if (b) (void)hipFree(b);
        }
        dp = nullptr; ctl = nullptr; Tp = Tv = Sp = Sv = nullptr;
        NP[0] = NP[1] = NV[0] = NV[1] = nullptr;
        H[0] = H[1] = nullptr;
        Hc[0] = Hc[1] = nullptr;
        hidx[0] = hidx[1] = nullptr;
        key = slot = count = start = nullptr; tmp_id = nullptr; status = nullptr;
        bucket = nullptr; spill = nullptr;
        st6 = nullptr; mat_d = nullptr; parts = nullptr; pmem = nullptr; energy_log = nullptr; ewave = nullptr;
        cic_acc = nullptr; cic_out = nullptr; rhs_buf = nullptr; cic_log = nullptr;
        d_mgoff = nullptr; d_mbox = nullptr; d_mcell = nullptr; d_mmu = nullptr;
        d_mscale = nullptr; d_mtab = nullptr; d_movf = nullptr;
        d_pbox = nullptr; d_pcell = nullptr; d_nbox = nullptr; d_ncell = nullptr;
        fin_slices = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    bool picard() const { return cfg.time_integrator == WD_BACKWARD_EULER_PICARD; }
    bool history() const { return cfg.tangential_model == WD_TANGENTIAL_SPRING_HISTORY; }

    hipEvent_t take_event() {
        if (!spare.empty()) {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }

    // Opens a timed stage on the stream (only when kernel timing is on).
    hipEvent_t stage_begin() {
        if (!timing) return nullptr;
        hipEvent_t a = take_event();
        (void)hipEventRecord(a, stream);
        return a;
    }
    void stage_end(int stage, hipEvent_t a) {
        if (!timing || !a) return;
        hipEvent_t b = take_event();
        (void)hipEventRecord(b, stream);
        pending.push_back({stage, a, b});
        if (pending.size() > 8192) resolve_timing();
    }
    void resolve_timing() {
        if (pending.empty()) return;
        (void)hipStreamSynchronize(stream);
        for (auto& t : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) stage_ms[t.stage] += ms;
            spare.push_back(t.a);
            spare.push_back(t.b);
        }
        pending.clear();
    }

    // One launch on the engine's stream.  When profiling, the launch carries
    // a start / stop event pair that the runtime stamps at the kernel's own
    // start and end (hipExtLaunchKernelGGL), so the pair times the kernel
    // alone, without the dispatch gap around it.
    template <typename K, typename... A>
    void go(const char* name, K kernel, dim3 grid, dim3 block, A... args) {
        go_lds(name, kernel, grid, block, 0u, args...);
    }
    // The same with `lds` bytes of dynamic LDS.
    template <typename K, typename... A>
    void go_lds(const char* name, K kernel, dim3 grid, dim3 block, unsigned lds, A... args) {
        if (!prof) {
            hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
            return;
        }
        hipEvent_t a = take_event(), b = take_event();
        hipExtLaunchKernelGGL(kernel, grid, block, lds, stream, a, b, 0, args...);
        prof->push_back({name, a, b});
    }

    // One neighbor pass: T (any order) -> S (cell order).  NMODE as k_reorder.
    // fin_flags >= 0: a fused step's bin, whose count the force pass made
    // (into the pending grid) and whose scan also reduces that pass's
    // partials with these k_finalize flags (fin_parts partials of fin_gpp
    // grains).
    // step_bin: a classic bin that starts a step (or primes fused steps): it
    // replaces a fused step's bin of the same state, whose pending
    // over-capacity count k_count then clears.
    void enqueue_bin(int nmode, const float4* nin_p, const float4* nin_v, float4* nout_p,
                     float4* nout_v, int fin_flags = -1, int fin_parts = 0, int fin_gpp = 0,
                     bool step_bin = false) {
        const bool fused = fin_flags >= 0;
        // The per-grain bin passes run one-wave blocks: at 32K grains they
        // then spread over every CU, at 1M they are ~2% faster than 256-thread
        // blocks.
        const int bbs = 64;
        const int gb = grain_blocks(n, bbs);
        if (!fused) {
            if (bucket_bin) {
                go("k_count", k_count<true>, dim3(gb), dim3(bbs), (const Params*)dp, ctl, (const float4*)Tp,
                   (const float4*)Tv, (const uint32_t*)key, slot, count, n, bin_box_out, bin_cell_out, bucket,
                   spill, step_bin ? 1 : 0);
            } else {
                go("k_count", k_count<false>, dim3(gb), dim3(bbs), (const Params*)dp, ctl, (const float4*)Tp,
                   (const float4*)Tv, (const uint32_t*)key, slot, count, n, bin_box_out, bin_cell_out, bucket,
                   spill, step_bin ? 1 : 0);
            }
        }
        // Tiles of the chosen size covering every cell the handle can hold,
        // on up to 1,024 blocks; blocks past the current box's tiles return.
        // A fused step's scan has kFinSliceBlocks more blocks, which reduce
        // the force pass's partials in slices (combined by the reorder).
        const size_t tile = static_cast<size_t>(kScanThreads) * kScanItemsVariants[scan_var];
        const int scan_grid = static_cast<int>(std::min<size_t>((scan_cells_used + tile - 1) / tile, 1024));
        ScanFuse fz{};
        ReorderFuse rf{};
        if (fused) {
            const bool with_e = (fin_flags & (kFinEnergyLog | kFinEnergyLast)) != 0;
            fz.on = 1;
            fz.parts = parts;
            fz.nparts = fin_parts;
            fz.ewave = ewave;
            fz.ng = (n + kEnergyGroup - 1) / kEnergyGroup;
            fz.with_e = with_e ? 1 : 0;
            fz.slices = fin_slices;
            fz.bio = bin_io();
            fz.key = key;
            fz.Tp = Tp;
            fz.n = n;
            rf.on = 1;
            rf.slices = fin_slices;
            rf.energy_log = energy_log;
            rf.flags = fin_flags;
            rf.box_out = bin_box_out;
            rf.cell_out = bin_cell_out;
            (void)fin_gpp;
        }
        auto scan = [&](auto kern) {
            go("k_scan", kern, dim3(scan_grid + (fused ? kFinSliceBlocks : 0)), dim3(kScanThreads),
               (const Params*)dp, ctl, count, start, status, fz);
        };
        if (scan_var == 0) scan(k_scan<4>);
        else if (scan_var == 1) scan(k_scan<8>);
        else scan(k_scan<16>);
        if (!bucket_bin) {
            go("k_scatter", k_scatter, dim3(gb), dim3(bbs), (const Params*)dp, (const uint32_t*)key,
               (const uint32_t*)slot, (const float4*)Tp, (const float4*)Tv, start, tmp_id, n, fused ? 1 : 0);
        }
        // Picard iterations after the first carry the spring index of the
        // step's first bin through every re-sort (nmode 2).
        const int hout = nmode == 2 ? hidx_cur ^ 1 : 0;
        const Carry carry{nin_p, nin_v, nout_p, nout_v, hidx[hidx_cur], history() ? hidx[hout] : nullptr};
        hidx_cur = hout;
        auto reorder = [&](auto kern) {
            go("k_reorder", kern, dim3(gb), dim3(bbs), (const Params*)dp, ctl,
               (const uint32_t*)key, (const uint32_t*)start, (const int*)tmp_id, (const uint32_t*)bucket,
               (const uint2*)spill, (const float4*)Tp, (const float4*)Tv, Sp, Sv, carry, n, rf);
        };
        if (bucket_bin) {
            if (nmode == 0) reorder(k_reorder<0, true>);
            else if (nmode == 1) reorder(k_reorder<1, true>);
            else reorder(k_reorder<2, true>);
        } else {
            if (nmode == 0) reorder(k_reorder<0, false>);
            else if (nmode == 1) reorder(k_reorder<1, false>);
            else reorder(k_reorder<2, false>);
        }
    }

    void enqueue_bin_current(bool step_bin = false) {
        enqueue_bin(0, nullptr, nullptr, nullptr, nullptr, -1, 0, 0, step_bin);
    }

    // The count a fused step's force pass makes for the next bin.
    BinIO bin_io() const { return BinIO{count, slot, bucket, spill, bucket_bin ? 2 : 1}; }

    template <int MODE, bool ENERGY, int BS>
    void launch_force_bs(const float4* np, const float4* nv, float* rhs, BinIO bio) {
        const HistoryIO hio{{H[0], H[1]}, {Hc[0], Hc[1]}, hidx[hidx_cur]};
        auto force = [&](auto kern) {
            go("k_force", kern, dim3(grain_blocks(n, BS)), dim3(BS), (const Params*)dp, ctl,
               (const float4*)Sp, (const float4*)Sv, np, nv, (const uint32_t*)start, Tp, Tv, key, rhs,
               parts, ewave, hio, n, bio);
        };
        if (history()) force(k_force<MODE, ENERGY, true, BS>);
        else force(k_force<MODE, ENERGY, false, BS>);
    }

    template <int MODE, bool ENERGY, int L>
    void launch_force_q(const float4* np, const float4* nv, BinIO bio) {
        const HistoryIO hio{{H[0], H[1]}, {Hc[0], Hc[1]}, hidx[hidx_cur]};
        // One lane per grain keeps the damping table of the handle's
        // materials in dynamic LDS, read when there is no per-member table.
        const unsigned lds = L == 1 && hp.mtab == nullptr
                ? static_cast<unsigned>(sizeof(float) * hp.nmat * hp.nmat) : 0u;
        auto force = [&](auto kern) {
            go_lds("k_force", kern, dim3(grain_blocks(n, 64 / L)), dim3(64), lds, (const Params*)dp, ctl,
                   (const float4*)Sp, (const float4*)Sv, np, nv, (const uint32_t*)start, Tp, Tv, key, parts,
                   ewave, hio, n, bio);
        };
        if (history()) force(k_force_q<MODE, ENERGY, true, L>);
        else force(k_force_q<MODE, ENERGY, false, L>);
    }

    // The variant the force pass runs with: the tuned or pinned one, else
    // 256 grains per block, one lane each.
    ForceVariant variant() const { return kForceVariants[force_var < 0 ? 2 : force_var]; }

    // last_pass: the pass completes a step (its springs become current).
    // fused: the pass counts its grains for the next bin and its partials are
    // reduced by that bin's scan: the k_finalize flags are returned in
    // *fin (with the partial count and grains per partial), nothing is
    // launched for them here.
    template <int MODE, bool ENERGY>
    void launch_force(const float4* np, const float4* nv, float* rhs, bool last_pass = true,
                      bool fused = false, int* fin = nullptr) {
        const ForceVariant fv = MODE == kModeRHS ? ForceVariant{kBlock, 0} : variant();
        const int bs = variant_grains(fv);
        const int gb = grain_blocks(n, bs);
        const BinIO bio = fused && MODE != kModeRHS ? bin_io() : BinIO{nullptr, nullptr, nullptr, nullptr, 0};
        if constexpr (MODE == kModeRHS) {
            launch_force_bs<MODE, ENERGY, 256>(np, nv, rhs, bio);
        } else {
            if (fv.lanes == 1) launch_force_q<MODE, ENERGY, 1>(np, nv, bio);
            else if (fv.lanes == 2) launch_force_q<MODE, ENERGY, 2>(np, nv, bio);
            else if (fv.lanes == 4) launch_force_q<MODE, ENERGY, 4>(np, nv, bio);
            else if (fv.block == 64) launch_force_bs<MODE, ENERGY, 64>(np, nv, rhs, bio);
            else if (fv.block == 128) launch_force_bs<MODE, ENERGY, 128>(np, nv, rhs, bio);
            else launch_force_bs<MODE, ENERGY, 256>(np, nv, rhs, bio);
        }
        int flags = 0;
        if (MODE != kModeRHS) {
            flags = kFinBox | kFinAccel | (ENERGY ? kFinEnergyLog : 0);
            if (history() && last_pass) flags |= kFinFlipHistory;
        } else if (ENERGY) {
            flags = kFinEnergyLast;
        }
        if (fused && MODE != kModeRHS) {
            fin[0] = flags | kFinPending;
            fin[1] = gb;
            fin[2] = bs;
            return;
        }
        if (flags) finalize(gb, flags, bs);
    }

    // gpp: grains per block partial of the pass being reduced.
    void finalize(int nparts, int flags, int gpp = kBlock) {
        go("k_finalize", k_finalize, dim3(1), dim3(kFinalizeBlock), (const Params*)dp, ctl,
           (const BlockPart*)parts, nparts, gpp, (const uint32_t*)key, (const float4*)Tp,
           (const float4*)Tv, (const double*)ewave, energy_log, flags);
    }

    // CIC deposit of the current state into `out` (4 * nodes floats), queued;
    // with `log` set, into the density log's next slot, counted on the
    // device (so that one captured graph serves every record).
    void enqueue_cic(float* out, bool log = false) {
        const int zb = static_cast<int>(std::min<size_t>((cic_count + kBlock - 1) / kBlock, 4096));
        unsigned* const next = log ? &ctl->cic_log_next : nullptr;
        go("k_cic_zero", k_cic_zero, dim3(zb), dim3(kBlock), cic_acc, cic_count, next);
        if (n > 0) {
            const int sb = (n + kBlock - 1) / kBlock;
            go("k_cic_scale", k_cic_scale, dim3(sb), dim3(kBlock), (const Params*)dp, parts,
               (const float4*)Tp, (const float4*)Tv, n);
            finalize(sb, kFinCicScale);
            // one-wave blocks: a block's node box is small and the grid
            // spreads over every CU (32K grains: 48 -> 33 us per deposit with
            // its four other kernels; 1M: 319 -> 196 us)
            go("k_cic_deposit", k_cic_deposit, dim3(grain_blocks(n, 64)), dim3(64), (const Params*)dp,
               (const float4*)Tp, (const float4*)Tv, cic_acc, n);
        }
        go("k_cic_convert", k_cic_convert, dim3(zb), dim3(kBlock), (const Params*)dp,
           (const unsigned long long*)cic_acc, out, cic_count, (const unsigned*)next);
    }

    // One step the classic way: bin the state (count, scan, scatter,
    // reorder), then the force pass and the reduction of its partials.
    void enqueue_step() {
        if (n == 0) return;
        if (!picard()) {
            hipEvent_t a = stage_begin();
            enqueue_bin_current(true);
            stage_end(WD_STAGE_FIND_NEIGHBORS, a);
            a = stage_begin();
            if (energy_on) launch_force<kModeSIE, true>(nullptr, nullptr, nullptr);
            else launch_force<kModeSIE, false>(nullptr, nullptr, nullptr);
            stage_end(WD_STAGE_COMPUTE_RHS, a);
            return;
        }
        // Picard: x^n is carried along every re-sort; every iteration starts
        // the springs from step n's generation, and only the last iteration's
        // springs become current.
        for (int k = 0; k < cfg.picard_iterations; ++k) {
            const int b = k & 1, pb = (k - 1) & 1;
            const bool last = k + 1 == cfg.picard_iterations;
            hipEvent_t a = stage_begin();
            if (k == 0) enqueue_bin(1, nullptr, nullptr, NP[0], NV[0], -1, 0, 0, true);
            else enqueue_bin(2, NP[pb], NV[pb], NP[b], NV[b]);
            stage_end(WD_STAGE_FIND_NEIGHBORS, a);
            a = stage_begin();
            if (energy_on && k == 0) launch_force<kModeBEP, true>(NP[b], NV[b], nullptr, last);
            else launch_force<kModeBEP, false>(NP[b], NV[b], nullptr, last);
            stage_end(WD_STAGE_COMPUTE_RHS, a);
        }
    }

    // Fused steps: a lone simulation, stepping through graphs or the tuner
    // (the kernel-timing mode keeps the classic order of its stages).
    bool use_fused() const { return fused_on && members == 1 && n > 0 && !timing; }

    // Makes the state binned with the current grid, as a fused step expects
    // on entry (for Picard, with x^n in the base buffer NP[0]).
    void prime_fused() {
        if (fused_ready) return;
        if (picard()) enqueue_bin(1, nullptr, nullptr, NP[0], NV[0], -1, 0, 0, true);
        else enqueue_bin_current(true);
        fused_ready = true;
    }

    // One fused step from a binned state: the force pass on S counts its
    // output T into the pending grid; T is binned by scan (+ the reduction of
    // the pass's partials) and reorder.  Picard: every iteration is such a
    // pass and bin; the bins carry x^n along (the last one copies the new
    // state into NP[0], the next step's base).
    void enqueue_step_fused() {
        if (n == 0) return;
        int fin[3] = {0, 0, 0};
        if (!picard()) {
            if (energy_on) launch_force<kModeSIE, true>(nullptr, nullptr, nullptr, true, true, fin);
            else launch_force<kModeSIE, false>(nullptr, nullptr, nullptr, true, true, fin);
            enqueue_bin(0, nullptr, nullptr, nullptr, nullptr, fin[0], fin[1], fin[2]);
            return;
        }
        for (int k = 0; k < cfg.picard_iterations; ++k) {
            const int b = k & 1;
            const bool last = k + 1 == cfg.picard_iterations;
            if (energy_on && k == 0) launch_force<kModeBEP, true>(NP[b], NV[b], nullptr, last, true, fin);
            else launch_force<kModeBEP, false>(NP[b], NV[b], nullptr, last, true, fin);
            if (last) enqueue_bin(1, nullptr, nullptr, NP[0], NV[0], fin[0], fin[1], fin[2]);
            else enqueue_bin(2, NP[b], NV[b], NP[b ^ 1], NV[b ^ 1], fin[0], fin[1], fin[2]);
        }
    }

    // A fused semi-implicit Euler step whose bin writes where `redirect`
    // points the bin's outputs once the force pass is enqueued (the
    // adjoint's recompute: the next step's slot).
    template <class Redirect>
    void enqueue_step_fused_into(Redirect redirect) {
        int fin[3] = {0, 0, 0};
        if (energy_on) launch_force<kModeSIE, true>(nullptr, nullptr, nullptr, true, true, fin);
        else launch_force<kModeSIE, false>(nullptr, nullptr, nullptr, true, true, fin);
        redirect();
        enqueue_bin(0, nullptr, nullptr, nullptr, nullptr, fin[0], fin[1], fin[2]);
    }

    // The step the handle runs now (fused steps need a primed state).
    void enqueue_any_step() {
        if (use_fused()) enqueue_step_fused();
        else enqueue_step();
    }

    void drop_graphs() {
        for (auto& g : graph_pow) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }

    int capture(int steps, hipGraphExec_t* out) {
        hipGraph_t g = nullptr;
        WD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < steps; ++s) enqueue_any_step();
        WD_HIP(hipStreamEndCapture(stream, &g));
        const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(WD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return WD_OK;
    }

    // Online autotuning of the force pass's variant (block size, lanes per
    // grain): steps run eagerly, cycling through the candidates round-robin
    // (so that the evolving state biases none of them), with the force
    // kernels and the reductions of their partials (more partials for fewer
    // grains per block) timed by kernel-stamped events; the first round warms
    // up, the fastest of the rest is kept.  These are real steps of the
    // simulation (the results do not depend on the variant).  A tuning spans
    // as many calls as it needs: a caller that advances one step per call
    // (the C++ class's take_step) is tuned over its first 18 calls.  The best
    // variant moves as contacts build up (a settling pile has several times
    // the contacts of the released lattice), so tuning runs again after 256,
    // 2048, 16384, ... steps.
    static constexpr int kCand = static_cast<int>(sizeof(kForceVariants) / sizeof(kForceVariants[0]));
    struct Tuning {
        int pos = -1;        // next (round, candidate) position; -1: no tuning in progress
        int rounds = 0;
        bool first = false;
        int before = -1;     // variant the graphs were captured with
        int scan_before = 1;
        bool bin_before = false;
        int scans = 0;       // tuning steps taken (each also times one scan tile size and bin mode)
        bool cand[kCand] = {};
        double best[kCand] = {};
        double scan_best[3] = {};
        double bin_best[2] = {};
        // Fused steps: the force pass also counts for the bin (bucket or slot
        // writes) and the scan also reduces the pass's partials, so the
        // variant, scan tile and binning are not timed apart.  The variants
        // are timed with the tile and binning held (whole steps), then the
        // six (tile, binning) pairs with the chosen variant: `combos` more
        // steps (two passes the first time, the first warming up).
        bool fused = false;
        int combos = 0;
        double combo_best[6] = {};
    } tuning;

    void begin_tuning() {
        Tuning& t = tuning;
        // The first tuning warms every variant up for a round and keeps the
        // best of two more; a later one times each remaining candidate once
        // and drops those that were more than 1.5x the best last time.
        t.first = force_var < 0 || tune_ms[0] == 0.0;
        t.rounds = t.first ? 3 : 1;
        double last_best = 1e30;
        for (int c = 0; c < kCand; ++c) last_best = std::min(last_best, tune_ms[c]);
        for (int c = 0; c < kCand; ++c) {
            t.cand[c] = t.first || tune_ms[c] <= 1.5 * last_best;
            t.best[c] = 1e30;
        }
        t.before = force_var;
        t.scan_before = scan_var;
        t.bin_before = bucket_bin;
        t.scans = 0;
        for (int v = 0; v < 3; ++v) t.scan_best[v] = 1e30;
        for (int v = 0; v < 2; ++v) t.bin_best[v] = 1e30;
        t.fused = use_fused();
        const bool pairs = !scan_fixed || (bucket_ok && !bin_fixed);
        t.combos = t.fused && pairs ? (t.first ? 12 : 6) : 0;
        for (int v = 0; v < 6; ++v) t.combo_best[v] = 1e30;
        t.pos = 0;
    }

    int autotune(int& nsteps) {
        if (n == 0) return WD_OK;
        if (tuning.pos < 0) {
            if (steps_run < next_tune) return WD_OK;
            begin_tuning();
        }
        Tuning& t = tuning;
        const int total = t.rounds * kCand;
        // One eager step with kernel-stamped events; returns the times of
        // the force pass (+ k_finalize), the scan and the bin passes.
        auto timed_step = [&](double& force, double& scan, double& bin) -> int {
            std::vector<ProfRec> recs;
            if (use_fused()) prime_fused();
            prof = &recs;
            enqueue_any_step();
            prof = nullptr;
            WD_HIP(hipStreamSynchronize(stream));
            force = scan = bin = 0.0;
            for (auto& rec : recs) {
                float e = 0.f;
                const bool ok = hipEventElapsedTime(&e, rec.a, rec.b) == hipSuccess;
                // the force pass and the reduction of its block partials
                if (ok && (std::strcmp(rec.name, "k_force") == 0 || std::strcmp(rec.name, "k_finalize") == 0)) {
                    force += e;
                }
                if (ok && std::strcmp(rec.name, "k_scan") == 0) scan += e;
                if (ok && (std::strcmp(rec.name, "k_count") == 0 || std::strcmp(rec.name, "k_scatter") == 0 ||
                           std::strcmp(rec.name, "k_reorder") == 0)) {
                    bin += e;
                }
                spare.push_back(rec.a);
                spare.push_back(rec.b);
            }
            time += hp.dt;
            --nsteps;
            ++steps_run;
            return WD_OK;
        };
        while (nsteps > 0 && t.pos < total) {
            const int r = t.pos / kCand, c = t.pos % kCand;
            ++t.pos;
            if (!t.cand[c]) continue;
            force_var = c;
            int sv = scan_var, bv = bucket_bin ? 1 : 0;
            bool scan_timed = false, bin_timed = false;
            if (!t.fused) {
                // The cell scan's tile size is tuned on the same steps, cycling
                // independently of the force variant (its time does not depend
                // on it); its first round warms up as the force pass's does.
                sv = t.scans % 3;
                scan_timed = !scan_fixed && (t.scans >= 3 || !t.first);
                if (!scan_fixed) scan_var = sv;
                // and the binning, alternating (with buckets on odd steps)
                bv = t.scans % 2;
                bin_timed = bucket_ok && !bin_fixed && (t.scans >= 2 || !t.first);
                if (bucket_ok && !bin_fixed) bucket_bin = bv == 1;
            }
            ++t.scans;
            double ms = 0.0, scan = 0.0, bin = 0.0;
            const int rc = timed_step(ms, scan, bin);
            if (rc != WD_OK) return rc;
            if (t.fused) ms += scan + bin;  // whole steps
            if (r > 0 || !t.first) t.best[c] = std::min(t.best[c], ms);
            if (scan_timed) t.scan_best[sv] = std::min(t.scan_best[sv], scan);
            if (bin_timed) t.bin_best[bv] = std::min(t.bin_best[bv], bin);
        }
        while (t.pos < total && !t.cand[t.pos % kCand]) ++t.pos;
        if (t.pos < total) return check_device();  // continues on the next call
        if (t.pos == total) {
            int best = -1;
            for (int c = 0; c < kCand; ++c) {
                if (t.cand[c] && (best < 0 || t.best[c] < t.best[best])) best = c;
            }
            for (int c = 0; c < kCand; ++c) tune_ms[c] = t.best[c];
            force_var = best;
        }
        // fused: the (scan tile, binning) pairs with the chosen variant
        while (nsteps > 0 && t.pos < total + t.combos) {
            const int q = t.pos - total, combo = q % 6;
            ++t.pos;
            const int sv = combo % 3, bv = combo / 3;
            if ((scan_fixed && sv != scan_var) || ((!bucket_ok || bin_fixed) && bv != (bucket_bin ? 1 : 0))) {
                continue;  // a pair the handle cannot take
            }
            scan_var = sv;
            bucket_bin = bv == 1;
            double ms = 0.0, scan = 0.0, bin = 0.0;
            const int rc = timed_step(ms, scan, bin);
            if (rc != WD_OK) return rc;
            if (!t.first || q >= 6) t.combo_best[combo] = std::min(t.combo_best[combo], ms + scan + bin);
        }
        if (t.pos < total + t.combos) return check_device();  // continues on the next call
        if (t.combos > 0) {
            int cb = -1;
            for (int v = 0; v < 6; ++v) {
                if (t.combo_best[v] < 1e30 && (cb < 0 || t.combo_best[v] < t.combo_best[cb])) cb = v;
            }
            if (cb >= 0) {
                scan_var = cb % 3;
                bucket_bin = cb / 3 == 1;
                for (int v = 0; v < 3; ++v) scan_ms[v] = std::min(t.combo_best[v], t.combo_best[v + 3]);
                for (int v = 0; v < 2; ++v) {
                    bin_ms[v] = std::min(t.combo_best[3 * v], std::min(t.combo_best[3 * v + 1], t.combo_best[3 * v + 2]));
                }
            } else {
                scan_var = t.scan_before;
                bucket_bin = t.bin_before;
            }
        }
        const int best = force_var;
        if (!t.fused && !scan_fixed) {
            int sb = t.scan_before;
            for (int v = 0; v < 3; ++v) {
                if (t.scan_best[v] < 1e30 && t.scan_best[v] < t.scan_best[sb]) sb = v;
            }
            for (int v = 0; v < 3; ++v) scan_ms[v] = t.scan_best[v];
            scan_var = sb;
        }
        if (!t.fused && bucket_ok && !bin_fixed) {
            for (int v = 0; v < 2; ++v) bin_ms[v] = t.bin_best[v];
            const int cur = t.bin_before ? 1 : 0;
            const int other = 1 - cur;
            bucket_bin = (t.bin_best[other] < t.bin_best[cur] ? other : cur) == 1;
        }
        next_tune = next_tune == 0 ? 256 : 8 * next_tune;
        // graphs captured with another variant, tile size or binning
        if (best != t.before || scan_var != t.scan_before || bucket_bin != t.bin_before) drop_graphs();
        t.pos = -1;
        return check_device();
    }

    int step(int nsteps) {
        if (!force_fixed) {
            const int rc = autotune(nsteps);
            if (rc != WD_OK) return rc;
        }
        if (n > 0 && nsteps > 0) {
            if (use_fused()) prime_fused();
            else fused_ready = false;  // a classic step leaves S unbinned
            if (timing) {
                for (int s = 0; s < nsteps; ++s) enqueue_step();
            } else if (!graphs_on) {
                for (int s = 0; s < nsteps; ++s) enqueue_any_step();
            } else {
                // graphs are captured the first time a call needs them
                auto graph = [&](int b, hipGraphExec_t* out) -> int {
                    if (!graph_pow[b]) {
                        const int rc = capture(1 << b, &graph_pow[b]);
                        if (rc != WD_OK) return rc;
                    }
                    *out = graph_pow[b];
                    return WD_OK;
                };
                hipGraphExec_t g = nullptr;
                int left = nsteps;
                if (left >= (1 << (kGraphs - 1))) {
                    int rc = graph(kGraphs - 1, &g);
                    if (rc != WD_OK) return rc;
                    for (; left >= (1 << (kGraphs - 1)); left -= 1 << (kGraphs - 1)) {
                        WD_HIP(hipGraphLaunch(g, stream));
                    }
                }
                for (int b = kGraphs - 2; b >= 0; --b) {
                    if (!(left & (1 << b))) continue;
                    int rc = graph(b, &g);
                    if (rc != WD_OK) return rc;
                    WD_HIP(hipGraphLaunch(g, stream));
                }
            }
            WD_HIP(hipGetLastError());
        }
        // The reference accumulates time in FP32, one dt per step.
        for (int s = 0; s < nsteps; ++s) time += hp.dt;
        steps_run += nsteps;
        return WD_OK;
    }

    int check_device() {
        int err = 0;
        WD_HIP(hipMemcpy(&err, &ctl->error, sizeof(int), hipMemcpyDeviceToHost));
        if (err == kAdjErrContacts) {
            (void)hipMemset(&ctl->error, 0, sizeof(int));
            return fail(WD_ERR_DEVICE, "adjoint: a grain has more pair contacts than the adjoint keeps (32)");
        }
        if (err) return fail(WD_ERR_DEVICE, "device: scan look-back wait exceeded its bound");
        return WD_OK;
    }

    int sync() {
        WD_HIP(hipStreamSynchronize(stream));
        return check_device();
    }
};

namespace {

template <typename T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    WD_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
    return WD_OK;
}

int validate(const wd_config& c) {
    if (c.collision_grid_size_x < 1 || c.collision_grid_size_y < 1 || c.collision_grid_size_z < 1) {
        return fail(WD_ERR_ARG, "collision grid sizes must be positive");
    }
    if (c.collision_grid_size_x > 2046 || c.collision_grid_size_y > 2046 ||
        c.collision_grid_size_z > 1022) {
        return fail(WD_ERR_ARG, "collision grid exceeds 2046 x 2046 x 1022 cells");
    }
    const double cells = (c.collision_grid_size_x + 2.0) * (c.collision_grid_size_y + 2.0) *
                         (c.collision_grid_size_z + 2.0);
    if (cells > 1.0e9) return fail(WD_ERR_ARG, "collision grid exceeds 1e9 cells");
    if (c.particles_per_cell < 1) return fail(WD_ERR_ARG, "particles_per_cell must be positive");
    if (c.density_grid_size_x < 1 || c.density_grid_size_y < 1 || c.density_grid_size_z < 1) {
        return fail(WD_ERR_ARG, "density grid sizes must be positive");
    }
    if (c.num_masses < 1 || c.num_masses > WD_MAX_MATERIALS) {
        return fail(WD_ERR_ARG, "materials.masses must hold 1..16 entries");
    }
    if (c.particle_material < 0 || c.particle_material >= c.num_masses) {
        return fail(WD_ERR_ARG, "initialization.particle_material has no entry in materials.masses");
    }
    if (c.tangential_model != WD_TANGENTIAL_VISCOUS && c.tangential_model != WD_TANGENTIAL_SPRING_HISTORY) {
        return fail(WD_ERR_ARG, "unknown tangential_model");
    }
    if (c.tangential_model == WD_TANGENTIAL_SPRING_HISTORY && !(c.tangential_stiffness_ratio > 0.f)) {
        return fail(WD_ERR_ARG, "tangential_stiffness_ratio must be positive for spring_history");
    }
    return WD_OK;
}

// The pair test of the contact law, (sqrtf(s2) - r) - r < 0 in FP32, is
// monotone in s2 (correctly rounded sqrt and subtraction), so it holds
// exactly below one FP32 value: the smallest s >= 0 for which it fails, found
// by bisection over the bit patterns of the non-negative floats (ordered as
// integers).  0 when nothing touches (r <= 0).
float touch_threshold(float r) {
    auto touching = [r](uint32_t bits) {
        float s;
        std::memcpy(&s, &bits, sizeof(s));
        return (std::sqrt(s) - r) - r < 0.f;
    };
    uint32_t lo = 0u, hi = 0x7F800000u;  // +inf never touches
    if (!touching(lo)) return 0.f;
    while (hi - lo > 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (touching(mid)) lo = mid;
        else hi = mid;
    }
    float s;
    std::memcpy(&s, &hi, sizeof(s));
    return s;
}

void fill_params(const wd_config& c, int n, Params& P) {
    std::memset(&P, 0, sizeof(P));
    P.n = n;
    P.gx = c.collision_grid_size_x;
    P.gy = c.collision_grid_size_y;
    P.gz = c.collision_grid_size_z;
    P.kcap = c.particles_per_cell;
    P.nmat = c.num_masses;
    P.lo[0] = c.x_min; P.lo[1] = c.y_min; P.lo[2] = c.z_min;
    P.hi[0] = c.x_max; P.hi[1] = c.y_max; P.hi[2] = c.z_max;
    P.wlo[0] = c.wall_x_min; P.wlo[1] = c.wall_y_min; P.wlo[2] = c.floor_z;
    P.whi[0] = c.wall_x_max; P.whi[1] = c.wall_y_max; P.whi[2] = c.ceiling_z;
    P.gravity = c.gravity;
    P.radius = c.particle_radius;
    P.max_force = c.max_force;
    float k = c.max_force / c.particle_radius;
    if (c.normal_stiffness > 0.f) {
        k = c.normal_stiffness;
        P.kneg = -c.normal_stiffness;
        P.clamp_delta = c.max_force / c.normal_stiffness;
    } else {
        P.kneg = -c.max_force / c.particle_radius;
        P.clamp_delta = c.particle_radius;
    }
    P.spring_k = k;
    // (2r)^2 with a 1e-5 relative margin: a pair whose FP32 squared distance
    // exceeds it has norm - r - r > 0 after rounding, so its spring, dashpot,
    // friction and energy terms are all exactly zero.
    P.cut2 = static_cast<float>(4.0 * c.particle_radius * static_cast<double>(c.particle_radius) * (1.0 + 1e-5));
    P.touch2 = touch_threshold(c.particle_radius);
    P.mu = c.friction;
    P.dt = c.dt;
    P.history = c.tangential_model == WD_TANGENTIAL_SPRING_HISTORY ? 1 : 0;
    P.kt = c.tangential_stiffness_ratio * k;
    P.inv_kt = 1.f / P.kt;
    P.dgx = c.density_grid_size_x;
    P.dgy = c.density_grid_size_y;
    P.dgz = c.density_grid_size_z;
    for (int i = 0; i < kMaxMat; ++i) {
        const float mi = i < c.num_masses ? c.masses[i] : 0.f;
        P.mass[i] = mi;
        P.c_wall[i] = damping(c.restitution, k, mi);
        for (int j = 0; j < kMaxMat; ++j) {
            const float mj = j < c.num_masses ? c.masses[j] : 0.f;
            const float msum = mi + mj;
            const float mred = (msum > 0.f) ? (mi * mj) / msum : 0.f;
            P.c_pair[i * kMaxMat + j] = damping(c.restitution, k, mred);
        }
    }
}

int upload_params(wd_sim* s) {
    const Params& P = s->hp;
    // every kernel dereferences these unconditionally (the member partials
    // only in a batch): refuse to publish a table that would fault
    if (!P.mgoff || !P.mbox || !P.mcell || !P.mmu || !P.mscale || !P.movf ||
        (P.members > 1 && (!P.pmem || !P.mtab))) {
        return fail(WD_ERR_STATE, "internal: solver constants lost a device buffer");
    }
    WD_HIP(hipMemcpy(s->dp, &s->hp, sizeof(Params), hipMemcpyHostToDevice));
    return WD_OK;
}

// Each member's mass and damping tables, formed by the same code as a single
// simulation's (members > 1 only).
int upload_member_tables(wd_sim* s) {
    if (!s->d_mtab) return WD_OK;
    std::vector<float> tab(static_cast<size_t>(kMatTab) * s->members);
    for (int m = 0; m < s->members; ++m) {
        Params q;
        fill_params(s->mcfg[m], s->n, q);
        float* t = tab.data() + static_cast<size_t>(kMatTab) * m;
        std::memcpy(t, q.mass, sizeof(float) * kMaxMat);
        std::memcpy(t + kMaxMat, q.c_wall, sizeof(float) * kMaxMat);
        std::memcpy(t + 2 * kMaxMat, q.c_pair, sizeof(float) * kMaxMat * kMaxMat);
    }
    WD_HIP(hipMemcpy(s->d_mtab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
    return WD_OK;
}

// Recomputes the derived constants from s->cfg (keeping the current dt and
// every member table and buffer the handle owns) and uploads them; graphs
// read them through dp.
// The collision grid of a handle: the configured one, or coarsened so that
// every cell is at least `scale` times the configured cell along each axis.
void grid_dims(const wd_config& c, float scale, int g[3]) {
    const int cfg[3] = {c.collision_grid_size_x, c.collision_grid_size_y, c.collision_grid_size_z};
    for (int a = 0; a < 3; ++a) {
        g[a] = scale <= 1.f ? cfg[a] : std::max(1, static_cast<int>(std::floor(cfg[a] / scale)));
    }
}

int refresh_params(wd_sim* s) {
    Params keep = s->hp;
    fill_params(s->cfg, s->n, s->hp);
    int g[3];
    grid_dims(s->cfg, s->cell_scale, g);
    s->hp.gx = g[0];
    s->hp.gy = g[1];
    s->hp.gz = g[2];
    s->hp.dt = keep.dt;
    s->hp.members = keep.members;
    s->hp.mgoff = keep.mgoff;
    s->hp.mbox = keep.mbox;
    s->hp.mcell = keep.mcell;
    s->hp.pbox = keep.pbox;
    s->hp.pcell = keep.pcell;
    s->hp.nbox = keep.nbox;
    s->hp.ncell = keep.ncell;
    s->hp.mmu = keep.mmu;
    s->hp.mscale = keep.mscale;
    s->hp.movf = keep.movf;
    s->hp.mtab = keep.mtab;
    s->hp.pmem = keep.pmem;  // member partials of a batch: k_finalize reads them
    const int rc = upload_member_tables(s);
    if (rc != WD_OK) return rc;
    return upload_params(s);
}

int allocate(wd_sim* s) {
    const int n = s->n;
    const wd_config& c = s->cfg;
    int rc;
#define WD_TRY(x) do { rc = (x); if (rc != WD_OK) return rc; } while (0)
    WD_TRY(dalloc(&s->dp, 1));
    WD_TRY(dalloc(&s->ctl, 1));
    WD_TRY(dalloc(&s->Tp, n)); WD_TRY(dalloc(&s->Tv, n));
    WD_TRY(dalloc(&s->Sp, n)); WD_TRY(dalloc(&s->Sv, n));
    if (s->picard()) {
        for (int b = 0; b < 2; ++b) { WD_TRY(dalloc(&s->NP[b], n)); WD_TRY(dalloc(&s->NV[b], n)); }
    }
    if (s->history()) {
        const size_t hs = static_cast<size_t>(kHist) * n;
        for (int b = 0; b < 2; ++b) {
            WD_TRY(dalloc(&s->H[b], hs));
            WD_TRY(dalloc(&s->Hc[b], n));
            WD_TRY(dalloc(&s->hidx[b], n));
        }
    }
    WD_TRY(dalloc(&s->key, n)); WD_TRY(dalloc(&s->slot, n)); WD_TRY(dalloc(&s->tmp_id, n));
    const size_t cells = static_cast<size_t>(c.collision_grid_size_x + 2) *
                         (c.collision_grid_size_y + 2) * (c.collision_grid_size_z + 2) *
                         static_cast<size_t>(s->members);
    s->scan_cells_used = cells + 1;
    s->ntiles_max = (cells + 1 + kScanTileMin - 1) / kScanTileMin;
    s->scan_cells = (cells + 1 + kScanTileMax - 1) / kScanTileMax * kScanTileMax;
    WD_TRY(dalloc(&s->count, s->scan_cells));
    WD_TRY(dalloc(&s->start, s->scan_cells));
    s->bucket_ok = s->scan_cells * kBucket * sizeof(uint32_t) <= (size_t(1) << 30);
    s->bucket_bin = s->bucket_ok && n <= (1 << 17);  // until tuned
    if (s->bucket_ok) {
        WD_TRY(dalloc(&s->bucket, s->scan_cells * kBucket));
        WD_TRY(dalloc(&s->spill, static_cast<size_t>(2) * std::max(n, 1)));
    }
    WD_TRY(dalloc(&s->status, s->ntiles_max));
    WD_TRY(dalloc(&s->st6, static_cast<size_t>(6) * n));
    WD_TRY(dalloc(&s->mat_d, n));
    s->nblocks = grain_blocks(n, 16);  // the most partials any pass writes (k_force_q, 4 lanes)
    WD_TRY(dalloc(&s->parts, s->nblocks));
    if (s->members > 1) WD_TRY(dalloc(&s->pmem, s->nblocks));
    WD_TRY(dalloc(&s->ewave, static_cast<size_t>((n + kEnergyGroup - 1) / kEnergyGroup)));
    s->nodes = static_cast<size_t>(c.density_grid_size_x) * c.density_grid_size_y * c.density_grid_size_z;
    s->cic_count = static_cast<size_t>(4) * s->nodes * static_cast<size_t>(s->members);
    const int E = s->members;
    WD_TRY(dalloc(&s->d_mgoff, E + 1));
    WD_TRY(dalloc(&s->d_mbox, 6 * E));
    WD_TRY(dalloc(&s->d_mcell, E + 1));
    WD_TRY(dalloc(&s->fin_slices, kFinSlices));
    WD_TRY(dalloc(&s->d_pbox, 6));
    WD_TRY(dalloc(&s->d_pcell, 2));
    WD_TRY(dalloc(&s->d_nbox, 6));
    WD_TRY(dalloc(&s->d_ncell, 2));
    WD_HIP(hipMemset(s->d_pcell, 0, sizeof(uint32_t) * 2));
    WD_HIP(hipMemset(s->d_ncell, 0, sizeof(uint32_t) * 2));
    WD_TRY(dalloc(&s->d_mmu, E));
    WD_TRY(dalloc(&s->d_mscale, E));
    WD_TRY(dalloc(&s->d_movf, E));
    WD_HIP(hipMemset(s->d_movf, 0, sizeof(unsigned long long) * E));
    if (E > 1) WD_TRY(dalloc(&s->d_mtab, static_cast<size_t>(kMatTab) * E));
    std::vector<float> mus(E);
    for (int m = 0; m < E; ++m) mus[m] = s->mcfg[m].friction;
    WD_HIP(hipMemcpy(s->d_mgoff, s->goff.data(), sizeof(int) * (E + 1), hipMemcpyHostToDevice));
    WD_HIP(hipMemcpy(s->d_mmu, mus.data(), sizeof(float) * E, hipMemcpyHostToDevice));
    WD_HIP(hipMemset(s->d_mcell, 0, sizeof(uint32_t) * (E + 1)));
    WD_HIP(hipMemset(s->d_mscale, 0, sizeof(unsigned) * E));
    s->hp.members = E;
    s->hp.mgoff = s->d_mgoff;
    s->hp.mbox = s->d_mbox;
    s->hp.mcell = s->d_mcell;
    s->hp.pbox = s->d_pbox;
    s->hp.pcell = s->d_pcell;
    s->hp.nbox = s->d_nbox;
    s->hp.ncell = s->d_ncell;
    s->hp.mmu = s->d_mmu;
    s->hp.mscale = s->d_mscale;
    s->hp.movf = s->d_movf;
    s->hp.pmem = s->pmem;
    s->hp.mtab = s->d_mtab;
    WD_TRY(upload_member_tables(s));
    WD_TRY(dalloc(&s->cic_acc, s->cic_count));  // int64 fixed point, all members
    WD_TRY(dalloc(&s->cic_out, s->cic_count));
#undef WD_TRY
    WD_HIP(hipMemset(s->ctl, 0, sizeof(Ctl)));
    WD_HIP(hipMemset(s->count, 0, s->scan_cells * sizeof(uint32_t)));
    WD_HIP(hipMemset(s->status, 0, s->ntiles_max * sizeof(unsigned long long)));

    return upload_params(s);
}

// Uploads an original-order state (and optionally materials) and rebuilds
// the working layout.
int ingest(wd_sim* s, const float* st, const int* mat, bool st_on_device = false) {
    s->fused_ready = false;
    if (s->n == 0) return WD_OK;
    WD_HIP(hipStreamSynchronize(s->stream));
    WD_HIP(hipMemcpy(s->st6, st, sizeof(float) * 6 * static_cast<size_t>(s->n),
                     st_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    if (mat) {
        std::vector<int> code(static_cast<size_t>(s->n));
        for (int m = 0; m < s->members; ++m) {
            for (int i = s->goff[m]; i < s->goff[m + 1]; ++i) code[i] = (m << 4) | mat[i];
        }
        WD_HIP(hipMemcpy(s->mat_d, code.data(), sizeof(int) * code.size(), hipMemcpyHostToDevice));
    }
    const int gb = grain_blocks(s->n);
    hipLaunchKernelGGL(k_ingest, dim3(gb), dim3(kBlock), 0, s->stream, s->dp, s->parts, s->st6,
                       s->mat_d, s->Tp, s->Tv, s->key, s->Hc[0], s->Hc[1], s->n);
    s->finalize(gb, kFinBox);
    WD_HIP(hipGetLastError());
    return s->sync();
}

int export_state(wd_sim* s) {
    if (s->n == 0) return WD_OK;
    const int blocks = (s->n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_export, dim3(blocks), dim3(kBlock), 0, s->stream, s->Tp, s->Tv, s->st6,
                       s->mat_d, s->n);
    WD_HIP(hipGetLastError());
    return s->sync();
}

}  // namespace

#define WD_CHECK_SIM(s) \
    if (!(s)) return fail(WD_ERR_ARG, std::string(__func__) + ": null handle")

namespace {

// Fields every ensemble member must share (the grid, the box, the spring and
// the integrator); members may differ in their initial grains, friction,
// restitution and material masses.
int check_compatible(const wd_config& a, const wd_config& b) {
    const bool same =
            a.dt == b.dt && a.collision_grid_size_x == b.collision_grid_size_x &&
            a.collision_grid_size_y == b.collision_grid_size_y &&
            a.collision_grid_size_z == b.collision_grid_size_z &&
            a.particles_per_cell == b.particles_per_cell &&
            a.density_grid_size_x == b.density_grid_size_x &&
            a.density_grid_size_y == b.density_grid_size_y &&
            a.density_grid_size_z == b.density_grid_size_z && a.x_min == b.x_min &&
            a.x_max == b.x_max && a.y_min == b.y_min && a.y_max == b.y_max && a.z_min == b.z_min &&
            a.z_max == b.z_max && a.gravity == b.gravity && a.particle_radius == b.particle_radius &&
            a.max_force == b.max_force && a.floor_z == b.floor_z && a.ceiling_z == b.ceiling_z &&
            a.wall_x_min == b.wall_x_min && a.wall_x_max == b.wall_x_max &&
            a.wall_y_min == b.wall_y_min && a.wall_y_max == b.wall_y_max &&
            a.tangential_model == b.tangential_model &&
            a.tangential_stiffness_ratio == b.tangential_stiffness_ratio &&
            a.normal_stiffness == b.normal_stiffness && a.num_masses == b.num_masses &&
            a.time_integrator == b.time_integrator && a.picard_iterations == b.picard_iterations;
    return same ? WD_OK
                : fail(WD_ERR_ARG, "ensemble members differ in more than initial state, friction, "
                                   "restitution and material masses");
}

int create_with(const wd_config* cfgs, int members, const std::vector<int>& goff, const float* st,
                const int* mat, wd_sim** out) {
    *out = nullptr;
    for (int m = 0; m < members; ++m) {
        int rc = validate(cfgs[m]);
        if (rc == WD_OK && m > 0) rc = check_compatible(cfgs[0], cfgs[m]);
        if (rc != WD_OK) return rc;
        if (members > 1 && goff[m + 1] <= goff[m]) return fail(WD_ERR_ARG, "ensemble member without grains");
    }
    const int n = goff[members];
    if (members > (1 << 26)) return fail(WD_ERR_ARG, "too many ensemble members");
    // Kernels address grain arrays with 32-bit byte offsets.
    if (n >= (1 << 28)) return fail(WD_ERR_ARG, "at most 268,435,455 grains per handle");
    for (int m = 0; m < members; ++m) {
        for (int i = goff[m]; i < goff[m + 1]; ++i) {
            if (mat[i] < 0 || mat[i] >= cfgs[m].num_masses) return fail(WD_ERR_ARG, "material id out of range");
        }
    }
    wd_sim* s = new wd_sim();
    s->cfg = cfgs[0];
    s->n = n;
    s->members = members;
    s->goff = goff;
    s->mcfg.assign(cfgs, cfgs + members);
    s->timing = cfgs[0].kernel_timing != 0;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(WD_ERR_HIP, "wd_create: hipGetDevice failed");
    }
    fill_params(cfgs[0], n, s->hp);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(WD_ERR_HIP, "wd_create: hipStreamCreate failed");
    }
    int rc = allocate(s);
    if (rc == WD_OK) rc = ingest(s, st, mat);
    if (rc != WD_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return WD_OK;
}

}  // namespace

extern "C" int wd_create(const wd_config* cfg, wd_sim** out) {
    if (!cfg || !out) return fail(WD_ERR_ARG, "wd_create: null argument");
    *out = nullptr;
    int n = 0;
    int rc = wd_config_grain_count(cfg, &n);
    if (rc != WD_OK) return rc;
    std::vector<float> st(static_cast<size_t>(6) * std::max(n, 1));
    std::vector<int> mat(static_cast<size_t>(std::max(n, 1)));
    rc = wd_initial_state(cfg, st.data(), mat.data(), n);
    if (rc != WD_OK) return rc;
    return create_with(cfg, 1, {0, n}, st.data(), mat.data(), out);
}

extern "C" int wd_create_from_state(const wd_config* cfg, int n, const float* state,
                                    const int* material, wd_sim** out) {
    if (!cfg || !out || n < 0 || (n > 0 && (!state || !material))) {
        return fail(WD_ERR_ARG, "wd_create_from_state: bad argument");
    }
    return create_with(cfg, 1, {0, n}, state, material, out);
}

extern "C" int wd_destroy(wd_sim* s) {
    if (s) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        delete s;
    }
    return WD_OK;
}

extern "C" int wd_num_grains(const wd_sim* s, int* n_out) {
    WD_CHECK_SIM(s);
    if (!n_out) return fail(WD_ERR_ARG, "wd_num_grains: null output");
    *n_out = s->n;
    return WD_OK;
}

extern "C" int wd_get_state(wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_get_state: null output");
    int rc = export_state(s);
    if (rc != WD_OK) return rc;
    if (s->n) WD_HIP(hipMemcpy(out, s->st6, sizeof(float) * 6 * static_cast<size_t>(s->n), hipMemcpyDeviceToHost));
    return WD_OK;
}

extern "C" int wd_set_state(wd_sim* s, const float* in) {
    WD_CHECK_SIM(s);
    if (!in) return fail(WD_ERR_ARG, "wd_set_state: null input");
    return ingest(s, in, nullptr);
}

extern "C" int wd_get_material(wd_sim* s, int* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_get_material: null output");
    int rc = export_state(s);
    if (rc != WD_OK) return rc;
    if (s->n) WD_HIP(hipMemcpy(out, s->mat_d, sizeof(int) * static_cast<size_t>(s->n), hipMemcpyDeviceToHost));
    for (int i = 0; i < s->n; ++i) out[i] &= (kMaxMat - 1);
    return WD_OK;
}

extern "C" int wd_set_material(wd_sim* s, const int* in) {
    WD_CHECK_SIM(s);
    if (!in) return fail(WD_ERR_ARG, "wd_set_material: null input");
    s->fused_ready = false;  // the binned copy holds the old material codes
    for (int i = 0; i < s->n; ++i) {
        if (in[i] < 0 || in[i] >= s->cfg.num_masses) return fail(WD_ERR_ARG, "wd_set_material: material id out of range");
    }
    if (s->n == 0) return WD_OK;
    WD_HIP(hipStreamSynchronize(s->stream));
    std::vector<int> code(static_cast<size_t>(s->n));
    for (int m = 0; m < s->members; ++m) {
        for (int i = s->goff[m]; i < s->goff[m + 1]; ++i) code[i] = (m << 4) | in[i];
    }
    WD_HIP(hipMemcpy(s->mat_d, code.data(), sizeof(int) * code.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_set_material, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                       s->Tp, s->Tv, s->mat_d, s->n);
    WD_HIP(hipGetLastError());
    return s->sync();
}

extern "C" int wd_set_dt(wd_sim* s, float dt) {
    WD_CHECK_SIM(s);
    if (s->hp.dt == dt) return WD_OK;
    WD_HIP(hipStreamSynchronize(s->stream));
    s->hp.dt = dt;
    s->cfg.dt = dt;
    return upload_params(s);
}

extern "C" int wd_get_time(const wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_get_time: null output");
    *out = s->time;
    return WD_OK;
}

extern "C" int wd_set_time(wd_sim* s, float t) {
    WD_CHECK_SIM(s);
    s->time = t;
    return WD_OK;
}

extern "C" int wd_step(wd_sim* s, int nsteps) {
    WD_CHECK_SIM(s);
    if (nsteps < 0) return fail(WD_ERR_ARG, "wd_step: negative step count");
    return s->step(nsteps);
}

extern "C" int wd_sync(wd_sim* s) {
    WD_CHECK_SIM(s);
    return s->sync();
}

extern "C" int wd_density_grid(wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_density_grid: null output");
    hipEvent_t a = s->stage_begin();
    s->enqueue_cic(s->cic_out);
    s->stage_end(WD_STAGE_DENSITY_GRID, a);
    WD_HIP(hipGetLastError());
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    WD_HIP(hipMemcpy(out, s->cic_out, sizeof(float) * 4 * s->nodes, hipMemcpyDeviceToHost));
    return WD_OK;
}

extern "C" int wd_total_energy(wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_total_energy: null output");
    if (s->n == 0) {
        *out = 0.f;
        return WD_OK;
    }
    hipEvent_t a = s->stage_begin();
    s->enqueue_bin_current();
    s->stage_end(WD_STAGE_FIND_NEIGHBORS, a);
    s->launch_force<kModeRHS, true>(nullptr, nullptr, nullptr);
    WD_HIP(hipGetLastError());
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    double e = 0.0;
    WD_HIP(hipMemcpy(&e, &s->ctl->energy_last, sizeof(double), hipMemcpyDeviceToHost));
    *out = static_cast<float>(e);
    return WD_OK;
}

extern "C" int wd_energy_log_enable(wd_sim* s, int capacity) {
    WD_CHECK_SIM(s);
    if (capacity < 0) return fail(WD_ERR_ARG, "wd_energy_log_enable: negative capacity");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    if (s->energy_log) {
        (void)hipFree(s->energy_log);
        s->energy_log = nullptr;
    }
    s->energy_cap = capacity;
    if (capacity > 0) {
        rc = dalloc(&s->energy_log, capacity);
        if (rc != WD_OK) return rc;
    }
    const int zero = 0;
    WD_HIP(hipMemcpy(&s->ctl->energy_count, &zero, sizeof(int), hipMemcpyHostToDevice));
    WD_HIP(hipMemcpy(&s->ctl->energy_capacity, &capacity, sizeof(int), hipMemcpyHostToDevice));
    const bool on = capacity > 0;
    if (on != s->energy_on) {
        s->energy_on = on;
        s->drop_graphs();
    }
    if (on) s->drop_graphs();  // the log pointer is baked into captured launches
    return WD_OK;
}

extern "C" int wd_energy_log_read(wd_sim* s, double* out, int count, int* recorded) {
    WD_CHECK_SIM(s);
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    int have = 0;
    WD_HIP(hipMemcpy(&have, &s->ctl->energy_count, sizeof(int), hipMemcpyDeviceToHost));
    if (recorded) *recorded = have;
    const int m = std::min(count, have);
    if (m > 0 && out) WD_HIP(hipMemcpy(out, s->energy_log, sizeof(double) * m, hipMemcpyDeviceToHost));
    return WD_OK;
}

extern "C" int wd_max_accel(wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_max_accel: null output");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    unsigned bits = 0;
    WD_HIP(hipMemcpy(&bits, &s->ctl->accel_last, sizeof(unsigned), hipMemcpyDeviceToHost));
    std::memcpy(out, &bits, sizeof(float));
    return WD_OK;
}

extern "C" int wd_is_stable(wd_sim* s, int* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_is_stable: null output");
    std::vector<float> st(static_cast<size_t>(6) * s->n);
    int rc = s->n ? wd_get_state(s, st.data()) : WD_OK;
    if (rc != WD_OK) return rc;
    const float lo[3] = {s->cfg.x_min, s->cfg.y_min, s->cfg.z_min};
    const float hi[3] = {s->cfg.x_max, s->cfg.y_max, s->cfg.z_max};
    *out = 1;
    for (int i = 0; i < s->n && *out; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float p = st[6 * static_cast<size_t>(i) + a];
            if (!std::isfinite(p) || p < lo[a] || p > hi[a]) {
                *out = 0;
                break;
            }
        }
    }
    return WD_OK;
}

extern "C" int wd_overflow_count(wd_sim* s, int* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_overflow_count: null output");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    unsigned long long v = 0;
    WD_HIP(hipMemcpy(&v, &s->ctl->overflow_lifetime, sizeof(v), hipMemcpyDeviceToHost));
    *out = static_cast<int>(std::min<unsigned long long>(v, 0x7FFFFFFFull));
    return WD_OK;
}

extern "C" int wd_stage_ms(wd_sim* s, float out[4]) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_stage_ms: null output");
    s->resolve_timing();
    for (int k = 0; k < 4; ++k) out[k] = s->stage_ms[k];
    return WD_OK;
}

extern "C" int wd_cell_order(wd_sim* s, int* ids_out, uint32_t* keys_out) {
    WD_CHECK_SIM(s);
    if (!ids_out || !keys_out) return fail(WD_ERR_ARG, "wd_cell_order: null output");
    if (s->n == 0) return WD_OK;
    s->enqueue_bin_current();
    int* ids = nullptr;
    uint32_t* keys = nullptr;
    int rc = dalloc(&ids, s->n);
    if (rc == WD_OK) rc = dalloc(&keys, s->n);
    if (rc == WD_OK) {
        hipLaunchKernelGGL(k_export_order, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                           s->dp, s->Sp, ids, keys, s->n);
        rc = s->sync();
    }
    if (rc == WD_OK && (hipMemcpy(ids_out, ids, sizeof(int) * s->n, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(keys_out, keys, sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost) != hipSuccess)) {
        rc = fail(WD_ERR_HIP, "wd_cell_order: copy failed");
    }
    if (ids) (void)hipFree(ids);
    if (keys) (void)hipFree(keys);
    return rc;
}

// ---------------------------------------------------------------------------
// Cell order without allocating: the staging arrays of wd_cell_order live in
// a fixed device array instead of per-call hipMalloc / hipFree.

namespace {

// ids in the first half, keys in the second; every call on a device shares
// that device's copy.
__device__ __align__(256) uint32_t scratch[2 * static_cast<size_t>(WD_NOALLOC_MAX_GRAINS)];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
uint32_t* scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (`device`), looked
// up once per device: in host code the array's name is only a stand-in.
int scratch_on(int device, uint32_t** out) {
    if (device < 0 || device >= kMaxDevices) return fail(WD_ERR_ARG, "scratch: device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            return fail(WD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = scratch_ptr[device];
    return WD_OK;
}

}  // namespace

extern "C" int wd_cell_order_noalloc(wd_sim* s, int* ids_out, uint32_t* keys_out) {
    WD_CHECK_SIM(s);
    if (!ids_out || !keys_out) return fail(WD_ERR_ARG, "wd_cell_order_noalloc: null output");
    if (s->n > WD_NOALLOC_MAX_GRAINS) {
        return fail(WD_ERR_ARG, "wd_cell_order_noalloc: the handle holds " + std::to_string(s->n) +
                                        " grains, more than the " + std::to_string(WD_NOALLOC_MAX_GRAINS) +
                                        " its fixed staging array holds");
    }
    if (s->n == 0) return WD_OK;
    int prev = 0;
    WD_HIP(hipGetDevice(&prev));
    if (prev != s->device) WD_HIP(hipSetDevice(s->device));
    uint32_t* base = nullptr;
    int rc = scratch_on(s->device, &base);
    if (rc == WD_OK) {
        int* ids = reinterpret_cast<int*>(base);
        uint32_t* keys = base + WD_NOALLOC_MAX_GRAINS;
        s->enqueue_bin_current();
        hipLaunchKernelGGL(k_export_order, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                           s->dp, s->Sp, ids, keys, s->n);
        rc = s->sync();
        if (rc == WD_OK &&
            (hipMemcpy(ids_out, ids, sizeof(int) * s->n, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(keys_out, keys, sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost) != hipSuccess)) {
            rc = fail(WD_ERR_HIP, "wd_cell_order_noalloc: copy failed");
        }
    }
    if (prev != s->device) (void)hipSetDevice(prev);
    return rc;
}

extern "C" int wd_neighbor_rows(wd_sim* s, int* out, size_t capacity) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_neighbor_rows: null output");
    const size_t need = static_cast<size_t>(s->n) * 27 * s->cfg.particles_per_cell;
    if (capacity < need) return fail(WD_ERR_ARG, "wd_neighbor_rows: output too small");
    if (s->n == 0) return WD_OK;
    s->enqueue_bin_current();
    int* rows = nullptr;
    int rc = dalloc(&rows, need);
    if (rc == WD_OK) {
        hipLaunchKernelGGL(k_neighbor_rows, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                           s->dp, s->Sv, s->Sp, s->start, rows, s->n);
        rc = s->sync();
    }
    if (rc == WD_OK && hipMemcpy(out, rows, sizeof(int) * need, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = fail(WD_ERR_HIP, "wd_neighbor_rows: copy failed");
    }
    if (rows) (void)hipFree(rows);
    return rc;
}

extern "C" int wd_rhs(wd_sim* s, float* out) {
    WD_CHECK_SIM(s);
    if (!out) return fail(WD_ERR_ARG, "wd_rhs: null output");
    if (s->n == 0) return WD_OK;
    if (!s->rhs_buf) {
        int rc = dalloc(&s->rhs_buf, static_cast<size_t>(6) * s->n);
        if (rc != WD_OK) return rc;
    }
    s->enqueue_bin_current();
    s->launch_force<kModeRHS, false>(nullptr, nullptr, s->rhs_buf);
    WD_HIP(hipGetLastError());
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    WD_HIP(hipMemcpy(out, s->rhs_buf, sizeof(float) * 6 * static_cast<size_t>(s->n), hipMemcpyDeviceToHost));
    return WD_OK;
}

extern "C" int wd_density_log_enable(wd_sim* s, int capacity) {
    WD_CHECK_SIM(s);
    if (capacity < 0) return fail(WD_ERR_ARG, "wd_density_log_enable: negative capacity");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    WD_HIP(hipStreamSynchronize(s->stream));
    if (s->graph_cic_log) (void)hipGraphExecDestroy(s->graph_cic_log);  // captured with the old buffer
    s->graph_cic_log = nullptr;
    const unsigned zero = 0u;
    WD_HIP(hipMemcpy(&s->ctl->cic_log_next, &zero, sizeof(zero), hipMemcpyHostToDevice));
    if (s->cic_log) {
        (void)hipFree(s->cic_log);
        s->cic_log = nullptr;
    }
    s->cic_log_cap = capacity;
    s->cic_log_count = 0;
    if (capacity > 0) return dalloc(&s->cic_log, s->cic_count * static_cast<size_t>(capacity));
    return WD_OK;
}

extern "C" int wd_density_log_record(wd_sim* s) {
    WD_CHECK_SIM(s);
    if (s->cic_log_count >= s->cic_log_cap) return fail(WD_ERR_STATE, "wd_density_log_record: log full");
    // The slot is taken on the device (ctl->cic_log_next, which equals
    // cic_log_count here), so the deposit's five launches replay from one
    // graph captured at the first record.
    if (s->timing || s->n == 0) {
        s->enqueue_cic(s->cic_log, true);
    } else {
        if (!s->graph_cic_log) {
            hipGraph_t g = nullptr;
            WD_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            s->enqueue_cic(s->cic_log, true);
            WD_HIP(hipStreamEndCapture(s->stream, &g));
            const hipError_t e = hipGraphInstantiate(&s->graph_cic_log, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->graph_cic_log = nullptr;
                return fail(WD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
        }
        WD_HIP(hipGraphLaunch(s->graph_cic_log, s->stream));
    }
    WD_HIP(hipGetLastError());
    s->cic_log_count++;
    return WD_OK;
}

extern "C" int wd_density_log_read(wd_sim* s, float* out, int count, int* recorded) {
    WD_CHECK_SIM(s);
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    if (recorded) *recorded = s->cic_log_count;
    const int m = std::min(count, s->cic_log_count);
    if (m > 0 && out) {
        WD_HIP(hipMemcpy(out, s->cic_log, sizeof(float) * s->cic_count * static_cast<size_t>(m),
                         hipMemcpyDeviceToHost));
    }
    return WD_OK;
}

extern "C" int wd_profile_step(wd_sim* s, int nsteps, float* kernel_ms, int capacity, int* count,
                               char* names, size_t names_capacity) {
    WD_CHECK_SIM(s);
    if (nsteps <= 0 || !kernel_ms || !count) return fail(WD_ERR_ARG, "wd_profile_step: bad argument");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    std::vector<wd_sim::ProfRec> recs;
    const bool timing = s->timing;
    s->timing = false;
    // the steps the handle runs (fused ones from a binned state)
    const bool fused = s->use_fused();
    if (fused) s->prime_fused();
    s->prof = &recs;
    for (int k = 0; k < nsteps; ++k) s->enqueue_any_step();
    s->prof = nullptr;
    if (!fused) s->fused_ready = false;
    s->timing = timing;
    for (int k = 0; k < nsteps; ++k) s->time += s->hp.dt;
    rc = s->sync();
    // Group by position within a step: every step launches the same sequence.
    const size_t per = nsteps > 0 ? recs.size() / static_cast<size_t>(nsteps) : 0;
    std::vector<double> acc(per, 0.0);
    for (size_t r = 0; r < recs.size(); ++r) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, recs[r].a, recs[r].b);
        if (per) acc[r % per] += ms;
        s->spare.push_back(recs[r].a);
        s->spare.push_back(recs[r].b);
    }
    if (rc != WD_OK) return rc;
    *count = static_cast<int>(per);
    std::string list;
    for (size_t k = 0; k < per; ++k) {
        if (static_cast<int>(k) < capacity) kernel_ms[k] = static_cast<float>(acc[k] / nsteps);
        if (k) list += ",";
        list += recs[k].name;
    }
    if (names && names_capacity) {
        std::strncpy(names, list.c_str(), names_capacity - 1);
        names[names_capacity - 1] = '\0';
    }
    return WD_OK;
}

extern "C" int wd_get_history(wd_sim* s, int* partner, float* xi, int* count) {
    WD_CHECK_SIM(s);
    if (!partner || !xi || !count) return fail(WD_ERR_ARG, "wd_get_history: null output");
    if (!s->history()) return fail(WD_ERR_STATE, "wd_get_history: tangential_model is not spring_history");
    if (s->n == 0) return WD_OK;
    int* dp_ = nullptr;
    float* dx = nullptr;
    int* dc = nullptr;
    const size_t slots = static_cast<size_t>(s->n) * kHist;
    int rc = dalloc(&dp_, slots);
    if (rc == WD_OK) rc = dalloc(&dx, 3 * slots);
    if (rc == WD_OK) rc = dalloc(&dc, s->n);
    if (rc == WD_OK) {
        rc = s->sync();
        int par = 0;
        if (rc == WD_OK && hipMemcpy(&par, &s->ctl->hist_parity, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = fail(WD_ERR_HIP, "wd_get_history: parity read failed");
        }
        if (rc == WD_OK) {
            hipLaunchKernelGGL(k_export_history, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               s->stream, s->Tp, s->H[par], s->Hc[par], dp_, dx, dc, s->n);
            rc = s->sync();
        }
    }
    if (rc == WD_OK && (hipMemcpy(partner, dp_, sizeof(int) * slots, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(xi, dx, sizeof(float) * 3 * slots, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(count, dc, sizeof(int) * s->n, hipMemcpyDeviceToHost) != hipSuccess)) {
        rc = fail(WD_ERR_HIP, "wd_get_history: copy failed");
    }
    if (dp_) (void)hipFree(dp_);
    if (dx) (void)hipFree(dx);
    if (dc) (void)hipFree(dc);
    return rc;
}

extern "C" int wd_batch_create(const wd_config* cfgs, int members, wd_sim** out) {
    if (!cfgs || !out || members < 1) return fail(WD_ERR_ARG, "wd_batch_create: bad argument");
    *out = nullptr;
    std::vector<int> goff(static_cast<size_t>(members) + 1, 0);
    for (int m = 0; m < members; ++m) {
        int k = 0;
        const int rc = wd_config_grain_count(&cfgs[m], &k);
        if (rc != WD_OK) return rc;
        goff[m + 1] = goff[m] + k;
    }
    const int n = goff[members];
    std::vector<float> st(static_cast<size_t>(6) * std::max(n, 1));
    std::vector<int> mat(static_cast<size_t>(std::max(n, 1)));
    for (int m = 0; m < members; ++m) {
        const int rc = wd_initial_state(&cfgs[m], st.data() + 6 * static_cast<size_t>(goff[m]),
                                        mat.data() + goff[m], goff[m + 1] - goff[m]);
        if (rc != WD_OK) return rc;
    }
    return create_with(cfgs, members, goff, st.data(), mat.data(), out);
}

extern "C" int wd_batch_members(const wd_sim* s, int* members) {
    WD_CHECK_SIM(s);
    if (!members) return fail(WD_ERR_ARG, "wd_batch_members: null output");
    *members = s->members;
    return WD_OK;
}

extern "C" int wd_batch_member_range(const wd_sim* s, int member, int* first, int* count) {
    WD_CHECK_SIM(s);
    if (!first || !count || member < 0 || member >= s->members) {
        return fail(WD_ERR_ARG, "wd_batch_member_range: bad argument");
    }
    *first = s->goff[member];
    *count = s->goff[member + 1] - s->goff[member];
    return WD_OK;
}

extern "C" int wd_batch_density_grid(wd_sim* s, int member, float* out) {
    WD_CHECK_SIM(s);
    if (!out || member < -1 || member >= s->members) return fail(WD_ERR_ARG, "wd_batch_density_grid: bad argument");
    hipEvent_t a = s->stage_begin();
    s->enqueue_cic(s->cic_out);
    s->stage_end(WD_STAGE_DENSITY_GRID, a);
    WD_HIP(hipGetLastError());
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    const size_t per = 4 * s->nodes;
    if (member < 0) {
        WD_HIP(hipMemcpy(out, s->cic_out, sizeof(float) * s->cic_count, hipMemcpyDeviceToHost));
    } else {
        WD_HIP(hipMemcpy(out, s->cic_out + per * static_cast<size_t>(member), sizeof(float) * per,
                         hipMemcpyDeviceToHost));
    }
    return WD_OK;
}

// ---------------------------------------------------------------------------
// Parameter updates used by gradient-based callers (extension).

extern "C" int wd_set_friction(wd_sim* s, int member, float mu) {
    WD_CHECK_SIM(s);
    if (member < 0 || member >= s->members) return fail(WD_ERR_ARG, "wd_set_friction: bad member");
    WD_HIP(hipStreamSynchronize(s->stream));
    s->mcfg[member].friction = mu;
    if (member == 0) s->cfg.friction = mu;
    WD_HIP(hipMemcpy(s->d_mmu + member, &mu, sizeof(float), hipMemcpyHostToDevice));
    return refresh_params(s);
}

extern "C" int wd_set_normal_stiffness(wd_sim* s, float k_n) {
    WD_CHECK_SIM(s);
    if (!(k_n > 0.f)) return fail(WD_ERR_ARG, "wd_set_normal_stiffness: k_n must be positive");
    WD_HIP(hipStreamSynchronize(s->stream));
    s->cfg.normal_stiffness = k_n;
    for (auto& c : s->mcfg) c.normal_stiffness = k_n;
    return refresh_params(s);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of a window of SIE steps (extension).

namespace {

int adj_buf(wd_sim* s, void** p, size_t bytes) {
    *p = nullptr;
    WD_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
    s->adj->owned.push_back(*p);
    return WD_OK;
}

template <typename T>
int adj_alloc(wd_sim* s, T** p, size_t count) {
    return adj_buf(s, reinterpret_cast<void**>(p), count * sizeof(T));
}

int adj_snap(wd_sim* s, wd_sim::Snap& sn) {
    const size_t n = static_cast<size_t>(s->n);
    int rc = adj_alloc(s, &sn.P, n);
    if (rc == WD_OK) rc = adj_alloc(s, &sn.V, n);
    if (rc == WD_OK) rc = adj_alloc(s, &sn.H, static_cast<size_t>(kHist) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &sn.C, n);
    return rc;
}

HistoryIO hist_io(wd_sim* s) {
    return HistoryIO{{s->H[0], s->H[1]}, {s->Hc[0], s->Hc[1]}, s->hidx[s->hidx_cur]};
}

void snap_save(wd_sim* s, const wd_sim::Snap& sn) {
    hipLaunchKernelGGL(k_snap_save, dim3((s->n + kBlock - 1) / kBlock), dim3(kBlock), 0, s->stream,
                       s->ctl, s->Tp, s->Tv, hist_io(s), sn.P, sn.V, sn.H, sn.C, s->n);
}

void snap_load(wd_sim* s, const wd_sim::Snap& sn) {
    s->fused_ready = false;
    const int gb = grain_blocks(s->n);
    hipLaunchKernelGGL(k_snap_load, dim3(gb), dim3(kBlock), 0, s->stream, s->dp, s->ctl, s->parts,
                       sn.P, sn.V, sn.H, sn.C, s->Tp, s->Tv, s->key, hist_io(s), s->n);
    s->finalize(gb, kFinBox);
}

int read_overflow(wd_sim* s, unsigned long long* out) {
    WD_HIP(hipStreamSynchronize(s->stream));
    WD_HIP(hipMemcpy(out, &s->ctl->overflow_lifetime, sizeof(*out), hipMemcpyDeviceToHost));
    return WD_OK;
}

// The forward window with a checkpoint every `every` steps.
int adj_run_record(wd_sim* s) {
    auto* A = s->adj;
    snap_save(s, A->ckpt[0]);
    for (int c = 1; c <= A->nseg; ++c) {
        const int rc = s->step(std::min(A->every, A->nsteps - (c - 1) * A->every));
        if (rc != WD_OK) return rc;
        snap_save(s, A->ckpt[c]);
    }
    WD_HIP(hipGetLastError());
    return s->sync();
}

// One segment of the backward sweep: recomputes its `steps` steps from
// checkpoint `ck` with the engine's bin and spring buffers redirected into the
// slots (step k bins into slot k and reads its springs from slot k, writing
// the next ones into slot k + 1), then applies the contact adjoints of those
// steps in reverse.  `par` is the device's spring parity at the segment start
// and `cur` the lxi buffer that holds the spring cotangents (in original
// order) at the segment start.  Everything after
// the checkpoint load reads only the slots and fixed buffers, so it is
// captured once per {steps, par, cur, variant} and replayed for the other
// segments (eagerly with kernel timing on).
void adj_recompute_body(wd_sim* s, int set, int steps, int par) {
    auto* A = s->adj;
    auto bin_into = [&](int k) {
        wd_sim::BinSlot& b = A->at(set, k);
        s->bin_box_out = b.mbox;  // the bin keeps the step's grid in the slot
        s->bin_cell_out = b.mcell;
        s->Sp = b.Sp;
        s->Sv = b.Sv;
        s->start = b.start;
        s->hidx[0] = b.hidx;
    };
    auto springs = [&](int k) {
        s->H[par] = A->at(set, k).H;
        s->Hc[par] = A->at(set, k).Hc;
        s->H[par ^ 1] = A->at(set, k + 1).H;
        s->Hc[par ^ 1] = A->at(set, k + 1).Hc;
    };
    if (s->use_fused()) {
        // Fused steps (the classic sequence's bits in three launches): the
        // checkpoint's state is binned into slot 0; step k's force pass then
        // reads slot k and its bin fills slot k + 1.
        bin_into(0);
        s->enqueue_bin_current(true);
        for (int k = 0; k < steps; ++k) {
            springs(k);
            s->enqueue_step_fused_into([&]() { bin_into(k + 1); });
            par ^= 1;
        }
        return;
    }
    for (int k = 0; k < steps; ++k) {
        bin_into(k);
        springs(k);
        s->enqueue_step();
        par ^= 1;
    }
}

// The contact adjoints of a recomputed segment's steps, in reverse, on
// stream `st`; they read the slots of `set` and the adjoint buffers only.
void adj_sweep_body(wd_sim* s, int set, int steps, int cur, hipStream_t st) {
    auto* A = s->adj;
    const int blocks = static_cast<int>((s->n + kBlock - 1) / kBlock);
    // four lanes per grain below kAdjLanesUpTo grains, one above (the same bits)
    const bool four = s->n <= kAdjLanesUpTo;
    const int agb = grain_blocks(s->n, four ? kAdjBlock / kAdjLanes : kAdjBlock);
    const int ggb = four ? (s->n + kAdjBlock / kAdjLanes - 1) / (kAdjBlock / kAdjLanes) : (s->n + kAdjBlock - 1) / kAdjBlock;
    // The cotangents arrive in original order (the checkpoint's); the last
    // step's contact adjoint reads them in that step's cell order.
    hipLaunchKernelGGL(k_adj_to_cells, dim3(blocks), dim3(kBlock), 0, st,
                       (const float4*)A->at(set, steps - 1).Sp, (const float4*)A->lx, (const float4*)A->lv,
                       (const float4*)A->lxi[cur], A->lx2, A->lv2, A->lxi[cur ^ 1], s->n);
    cur ^= 1;
    for (int k = steps - 1; k >= 0; --k) {
        const wd_sim::BinSlot& b = A->at(set, k);
        const bool last = k == steps - 1;
        hipLaunchKernelGGL(four ? k_adj_contacts<kAdjLanes> : k_adj_contacts<1>, dim3(agb), dim3(kAdjBlock), 0, st, s->dp, s->ctl,
                           (const float4*)b.Sp, (const float4*)b.Sv, (const uint32_t*)b.start,
                           (const int*)b.hidx, (const float4*)b.H, (const int*)b.Hc,
                           (const int*)b.mbox, (const uint32_t*)b.mcell,
                           (const float4*)(last ? A->lx2 : A->lx), (const float4*)(last ? A->lv2 : A->lv),
                           (const float4*)A->lxi[cur], A->lxi[cur ^ 1], A->own_x, A->own_v, A->cg_d,
                           A->cg_v, A->cg_cnt, A->cg_p, A->gparam, s->n);
        // dL/d(state before step k), at the previous step's positions
        hipLaunchKernelGGL(four ? k_adj_gather<kAdjLanes> : k_adj_gather<1>, dim3(ggb), dim3(kAdjBlock), 0, st, A->own_x,
                           A->own_v, A->cg_d, A->cg_v, A->cg_cnt, A->cg_p, (const int*)b.hidx, A->lx,
                           A->lv, s->n);
        cur ^= 1;
    }
}

void adj_segment_body(wd_sim* s, int steps, int par, int cur) {
    adj_recompute_body(s, 0, steps, par);
    adj_sweep_body(s, 0, steps, cur, s->stream);
}

// Instantiated graph of `body` captured on stream `st`, cached under `key`.
template <class Body>
int adj_graph(std::vector<std::pair<std::array<int, 4>, hipGraphExec_t>>& cache, const std::array<int, 4>& key,
              hipStream_t st, Body body, hipGraphExec_t* out) {
    for (auto& g : cache) {
        if (g.first == key) {
            *out = g.second;
            return WD_OK;
        }
    }
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        body();
        e = hipStreamEndCapture(st, &g);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(WD_ERR_HIP, std::string("adjoint segment capture: ") + hipGetErrorString(e));
    cache.push_back({key, exec});
    *out = exec;
    return WD_OK;
}

int adj_segment(wd_sim* s, const wd_sim::Snap& ck, int steps, int par, int cur) {
    auto* A = s->adj;
    float4* const Sp0 = s->Sp;
    float4* const Sv0 = s->Sv;
    uint32_t* const start0 = s->start;
    int* const hidx0 = s->hidx[0];
    float4* const H0[2] = {s->H[0], s->H[1]};
    int* const Hc0[2] = {s->Hc[0], s->Hc[1]};
    auto restore = [&]() {
        s->bin_box_out = nullptr;
        s->bin_cell_out = nullptr;
        s->Sp = Sp0;
        s->Sv = Sv0;
        s->start = start0;
        s->hidx[0] = hidx0;
        s->H[0] = H0[0];
        s->H[1] = H0[1];
        s->Hc[0] = Hc0[0];
        s->Hc[1] = Hc0[1];
    };
    s->H[par] = A->at(0, 0).H;  // the checkpoint's springs go to slot 0
    s->Hc[par] = A->at(0, 0).Hc;
    snap_load(s, ck);
    int rc = WD_OK;
    if (s->timing) {
        adj_segment_body(s, steps, par, cur);
    } else {
        const std::array<int, 4> key{steps, par, cur, 2 * s->force_var + (s->use_fused() ? 1 : 0)};
        hipGraphExec_t exec = nullptr;
        for (auto& g : A->seg_graphs) {
            if (g.first == key) exec = g.second;
        }
        if (!exec) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                adj_segment_body(s, steps, par, cur);
                e = hipStreamEndCapture(s->stream, &g);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                restore();
                return fail(WD_ERR_HIP, std::string("adjoint segment capture: ") + hipGetErrorString(e));
            }
            A->seg_graphs.push_back({key, exec});
        }
        if (hipGraphLaunch(exec, s->stream) != hipSuccess) rc = fail(WD_ERR_HIP, "adjoint segment launch");
    }
    restore();
    WD_HIP(hipGetLastError());
    return rc;
}

// The engine's buffer pointers and stream, put back when a recompute that
// redirected them is enqueued.
struct EngineBufs {
    wd_sim* s;
    float4 *Sp, *Sv, *H[2];
    uint32_t* start;
    int *hidx0, *Hc[2];
    hipStream_t stream;
    explicit EngineBufs(wd_sim* sim)
        : s(sim), Sp(sim->Sp), Sv(sim->Sv), H{sim->H[0], sim->H[1]}, start(sim->start),
          hidx0(sim->hidx[0]), Hc{sim->Hc[0], sim->Hc[1]}, stream(sim->stream) {}
    ~EngineBufs() {
        s->bin_box_out = nullptr;
        s->bin_cell_out = nullptr;
        s->Sp = Sp;
        s->Sv = Sv;
        s->start = start;
        s->hidx[0] = hidx0;
        s->H[0] = H[0];
        s->H[1] = H[1];
        s->Hc[0] = Hc[0];
        s->Hc[1] = Hc[1];
        s->stream = stream;
    }
};

// Recomputes a segment from checkpoint `ck` into slot set `set` on the
// recompute stream, then records ev_rec[set] there.
int adj_recompute(wd_sim* s, const wd_sim::Snap& ck, int set, int steps, int par) {
    auto* A = s->adj;
    EngineBufs keep(s);
    s->stream = A->rstream;
    s->H[par] = A->at(set, 0).H;  // the checkpoint's springs go to the set's slot 0
    s->Hc[par] = A->at(set, 0).Hc;
    snap_load(s, ck);
    hipGraphExec_t exec = nullptr;
    const int rc = adj_graph(A->rec_graphs, {steps, par, set, 2 * s->force_var + (s->use_fused() ? 1 : 0)}, A->rstream,
                             [&]() { adj_recompute_body(s, set, steps, par); }, &exec);
    if (rc != WD_OK) return rc;
    if (hipGraphLaunch(exec, A->rstream) != hipSuccess) return fail(WD_ERR_HIP, "adjoint recompute launch");
    WD_HIP(hipEventRecord(A->ev_rec[set], A->rstream));
    return WD_OK;
}

// The contact adjoints of the segment in slot set `set`, on the engine's
// stream once its recompute is done, then records ev_swp[set].
int adj_sweep(wd_sim* s, int set, int steps, int cur) {
    auto* A = s->adj;
    WD_HIP(hipStreamWaitEvent(s->stream, A->ev_rec[set], 0));
    hipGraphExec_t exec = nullptr;
    const int rc = adj_graph(A->swp_graphs, {steps, cur, set, -1}, s->stream,
                             [&]() { adj_sweep_body(s, set, steps, cur, s->stream); }, &exec);
    if (rc != WD_OK) return rc;
    if (hipGraphLaunch(exec, s->stream) != hipSuccess) return fail(WD_ERR_HIP, "adjoint sweep launch");
    WD_HIP(hipEventRecord(A->ev_swp[set], s->stream));
    return WD_OK;
}

}  // namespace

extern "C" int wd_adjoint_record(wd_sim* s, int nsteps, int checkpoint_every) {
    WD_CHECK_SIM(s);
    if (nsteps < 1 || checkpoint_every < 1) return fail(WD_ERR_ARG, "wd_adjoint_record: bad window");
    if (s->picard() || !s->history()) {
        return fail(WD_ERR_STATE, "wd_adjoint_record: the adjoint covers semi-implicit Euler with spring_history");
    }
    if (s->n == 0) return fail(WD_ERR_STATE, "wd_adjoint_record: no grains");
    const int every = std::min(checkpoint_every, nsteps);
    int rc = WD_OK;
    if (s->adj && s->adj->nsteps == nsteps && s->adj->every == every) {
        rc = read_overflow(s, &s->adj->overflow0);  // same window shape: keep the buffers
        if (rc != WD_OK) return rc;
        return adj_run_record(s);
    }
    s->drop_adjoint();
    s->adj = new wd_sim::AdjointState();
    auto* A = s->adj;
    A->nsteps = nsteps;
    A->every = every;
    A->nseg = (nsteps + A->every - 1) / A->every;
    const size_t n = static_cast<size_t>(s->n);
    A->ckpt.resize(A->nseg + 1);
    A->sets = A->nseg > 1 ? 2 : 1;
    A->slot.resize(static_cast<size_t>(A->sets) * (A->every + 1));
    if (A->sets == 2) {
        if (hipStreamCreateWithFlags(&A->rstream, hipStreamNonBlocking) != hipSuccess) rc = fail(WD_ERR_HIP, "adjoint stream");
        for (hipEvent_t* e : {&A->ev_start, &A->ev_rec[0], &A->ev_rec[1], &A->ev_swp[0], &A->ev_swp[1]}) {
            if (rc == WD_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
                rc = fail(WD_ERR_HIP, "adjoint events");
            }
        }
    }
    for (auto& sn : A->ckpt) if (rc == WD_OK) rc = adj_snap(s, sn);
    for (auto& b : A->slot) {
        if (rc == WD_OK) rc = adj_alloc(s, &b.Sp, n);
        if (rc == WD_OK) rc = adj_alloc(s, &b.Sv, n);
        if (rc == WD_OK) rc = adj_alloc(s, &b.start, s->scan_cells);
        if (rc == WD_OK) rc = adj_alloc(s, &b.hidx, n);
        if (rc == WD_OK) rc = adj_alloc(s, &b.H, static_cast<size_t>(kHist) * n);
        if (rc == WD_OK) rc = adj_alloc(s, &b.Hc, n);
        if (rc == WD_OK) rc = adj_alloc(s, &b.mbox, 6 * static_cast<size_t>(s->members));
        if (rc == WD_OK) rc = adj_alloc(s, &b.mcell, static_cast<size_t>(s->members) + 1);
    }
    if (rc == WD_OK) rc = adj_alloc(s, &A->cg_p, static_cast<size_t>(kAdjContacts) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lx, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lv, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lx2, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lv2, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->own_x, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->own_v, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lxi[0], static_cast<size_t>(kHist) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->lxi[1], static_cast<size_t>(kHist) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->cg_d, static_cast<size_t>(kAdjContacts) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->cg_v, static_cast<size_t>(kAdjContacts) * n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->cg_cnt, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->gparam, n);
    if (rc == WD_OK) rc = adj_alloc(s, &A->grid_cot, s->cic_count);
    if (rc == WD_OK) rc = adj_alloc(s, &A->state_cot, 6 * n);
    if (rc == WD_OK) rc = read_overflow(s, &A->overflow0);
    if (rc != WD_OK) {
        s->drop_adjoint();
        return rc;
    }
    return adj_run_record(s);
}

namespace {

// wd_adjoint_backward with the cotangents and dL/d(initial state) in host
// memory (dev false) or in device memory of the handle's GPU (dev true).
int adjoint_backward(wd_sim* s, const float* grid_cot, const float* state_cot, float* grad_state0,
                     double* grad_params, bool dev) {
    auto* A = s->adj;
    if (!A) return fail(WD_ERR_STATE, "wd_adjoint_backward: no recorded window (wd_adjoint_record)");
    unsigned long long over = 0;
    int rc = read_overflow(s, &over);
    if (rc != WD_OK) return rc;
    if (over != A->overflow0) {
        return fail(WD_ERR_STATE, "wd_adjoint_backward: cells over particles_per_cell in the window; "
                                  "the neighbor lists are not symmetric");
    }
    const size_t n = static_cast<size_t>(s->n);
    const hipMemcpyKind in_kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (grid_cot) WD_HIP(hipMemcpy(A->grid_cot, grid_cot, sizeof(float) * s->cic_count, in_kind));
    if (state_cot) WD_HIP(hipMemcpy(A->state_cot, state_cot, sizeof(float) * 6 * n, in_kind));
    WD_HIP(hipMemsetAsync(A->lxi[0], 0, sizeof(float4) * kHist * n, s->stream));
    WD_HIP(hipMemsetAsync(A->lxi[1], 0, sizeof(float4) * kHist * n, s->stream));
    WD_HIP(hipMemsetAsync(A->gparam, 0, sizeof(double2) * n, s->stream));
    const int blocks = static_cast<int>((n + kBlock - 1) / kBlock);
    snap_load(s, A->ckpt[A->nseg]);
    hipLaunchKernelGGL(k_adj_seed, dim3(blocks), dim3(kBlock), 0, s->stream, s->dp, s->Tp, s->Tv,
                       grid_cot ? A->grid_cot : nullptr, state_cot ? A->state_cot : nullptr, A->lx,
                       A->lv, s->n);
    const bool energy = s->energy_on;
    s->energy_on = false;  // recomputed steps append nothing to the energy log
    // The device's spring parity, read once: a recomputed step flips it, a
    // checkpoint load leaves it.
    int par = 0;
    WD_HIP(hipStreamSynchronize(s->stream));
    WD_HIP(hipMemcpy(&par, &s->ctl->hist_parity, sizeof(int), hipMemcpyDeviceToHost));
    int cur = 0;
    auto seg_len = [&](int seg) { return std::min(A->nsteps, (seg + 1) * A->every) - seg * A->every; };
    if (A->sets == 2 && !s->timing) {
        // Segment seg - 1 is recomputed on the recompute stream, into the
        // other slot set, while the contact adjoints of segment seg run on
        // the engine's stream.  A set is recomputed into only once the
        // adjoints that read it last are done.
        std::vector<int> seg_par(A->nseg), seg_cur(A->nseg);
        for (int seg = A->nseg - 1; seg >= 0; --seg) {
            seg_par[seg] = par;
            seg_cur[seg] = cur;
            par ^= seg_len(seg) & 1;
            cur ^= (seg_len(seg) + 1) & 1;
        }
        WD_HIP(hipEventRecord(A->ev_start, s->stream));  // after the seed, before any recompute
        WD_HIP(hipStreamWaitEvent(A->rstream, A->ev_start, 0));
        rc = adj_recompute(s, A->ckpt[A->nseg - 1], (A->nseg - 1) & 1, seg_len(A->nseg - 1), seg_par[A->nseg - 1]);
        for (int seg = A->nseg - 1; seg >= 0 && rc == WD_OK; --seg) {
            if (seg >= 1) {
                const int other = (seg - 1) & 1;
                if (seg + 1 < A->nseg && hipStreamWaitEvent(A->rstream, A->ev_swp[other], 0) != hipSuccess) {
                    rc = fail(WD_ERR_HIP, "adjoint stream wait");
                    break;
                }
                rc = adj_recompute(s, A->ckpt[seg - 1], other, seg_len(seg - 1), seg_par[seg - 1]);
                if (rc != WD_OK) break;
            }
            rc = adj_sweep(s, seg & 1, seg_len(seg), seg_cur[seg]);
        }
        // The engine's stream is past every recompute (the sweep of segment
        // 0 waited on the last); on an error, let the other stream drain.
        if (rc != WD_OK) (void)hipStreamSynchronize(A->rstream);
    } else {
        for (int seg = A->nseg - 1; seg >= 0; --seg) {
            const int K = seg_len(seg);
            rc = adj_segment(s, A->ckpt[seg], K, par, cur);
            if (rc != WD_OK) break;
            par ^= K & 1;
            cur ^= (K + 1) & 1;  // K steps and the move into cell order
        }
    }
    s->energy_on = energy;
    if (rc != WD_OK) return rc;
    snap_load(s, A->ckpt[A->nseg]);  // the handle is left at the end of the window
    WD_HIP(hipGetLastError());
    rc = s->sync();
    if (rc != WD_OK) return rc;
    if (grad_state0 && dev) {
        // x and v of each grain's float4s into its six floats
        WD_HIP(hipMemcpy2D(grad_state0, 6 * sizeof(float), A->lx, sizeof(float4), 3 * sizeof(float), n,
                           hipMemcpyDeviceToDevice));
        WD_HIP(hipMemcpy2D(grad_state0 + 3, 6 * sizeof(float), A->lv, sizeof(float4), 3 * sizeof(float), n,
                           hipMemcpyDeviceToDevice));
    } else if (grad_state0) {
        std::vector<float4> x(n), v(n);
        WD_HIP(hipMemcpy(x.data(), A->lx, sizeof(float4) * n, hipMemcpyDeviceToHost));
        WD_HIP(hipMemcpy(v.data(), A->lv, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            float* g = grad_state0 + 6 * i;
            g[0] = x[i].x; g[1] = x[i].y; g[2] = x[i].z;
            g[3] = v[i].x; g[4] = v[i].y; g[5] = v[i].z;
        }
    }
    if (grad_params) {
        std::vector<double2> gp(n);
        WD_HIP(hipMemcpy(gp.data(), A->gparam, sizeof(double2) * n, hipMemcpyDeviceToHost));
        for (int m = 0; m < s->members; ++m) {
            double gmu = 0.0, gk = 0.0;
            for (int i = s->goff[m]; i < s->goff[m + 1]; ++i) {
                gmu += gp[i].x;
                gk += gp[i].y;
            }
            grad_params[2 * m] = gmu;
            grad_params[2 * m + 1] = gk;
        }
    }
    return WD_OK;
}

// WD_OK when p is device memory of the handle's GPU.
int device_buffer(const wd_sim* s, const void* p, const char* what) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice || a.device != s->device) {
        (void)hipGetLastError();
        return fail(WD_ERR_ARG, std::string(what) + ": not device memory of the handle's GPU");
    }
    return WD_OK;
}

}  // namespace

extern "C" int wd_adjoint_backward(wd_sim* s, const float* grid_cot, const float* state_cot,
                                   float* grad_state0, double* grad_params) {
    WD_CHECK_SIM(s);
    return adjoint_backward(s, grid_cot, state_cot, grad_state0, grad_params, false);
}

// ---------------------------------------------------------------------------
// Device-memory variants (extension): the same calls with the caller's
// buffers in device memory of the handle's GPU (torch tensors, say), so that
// a differentiable caller's data stays on the GPU.  Each waits for the
// device first (the caller's producers, on any stream) and returns with its
// results complete.

extern "C" int wd_set_state_device(wd_sim* s, const float* in) {
    WD_CHECK_SIM(s);
    if (!in) return fail(WD_ERR_ARG, "wd_set_state_device: null input");
    if (s->n == 0) return WD_OK;
    int rc = device_buffer(s, in, "wd_set_state_device");
    if (rc != WD_OK) return rc;
    WD_HIP(hipDeviceSynchronize());
    return ingest(s, in, nullptr, true);
}

extern "C" int wd_batch_density_grid_device(wd_sim* s, int member, float* out) {
    WD_CHECK_SIM(s);
    if (!out || member < -1 || member >= s->members) {
        return fail(WD_ERR_ARG, "wd_batch_density_grid_device: bad argument");
    }
    int rc = device_buffer(s, out, "wd_batch_density_grid_device");
    if (rc != WD_OK) return rc;
    WD_HIP(hipDeviceSynchronize());
    hipEvent_t a = s->stage_begin();
    s->enqueue_cic(s->cic_out);
    s->stage_end(WD_STAGE_DENSITY_GRID, a);
    WD_HIP(hipGetLastError());
    const size_t per = 4 * s->nodes;
    const float* src = s->cic_out + (member < 0 ? 0 : per * static_cast<size_t>(member));
    const size_t count = member < 0 ? s->cic_count : per;
    WD_HIP(hipMemcpyAsync(out, src, sizeof(float) * count, hipMemcpyDeviceToDevice, s->stream));
    return s->sync();
}

extern "C" int wd_adjoint_backward_device(wd_sim* s, const float* grid_cot, const float* state_cot,
                                          float* grad_state0, double* grad_params) {
    WD_CHECK_SIM(s);
    for (const void* p : {static_cast<const void*>(grid_cot), static_cast<const void*>(state_cot),
                          static_cast<const void*>(grad_state0)}) {
        if (p) {
            const int rc = device_buffer(s, p, "wd_adjoint_backward_device");
            if (rc != WD_OK) return rc;
        }
    }
    WD_HIP(hipDeviceSynchronize());
    return adjoint_backward(s, grid_cot, state_cot, grad_state0, grad_params, true);
}

// ---------------------------------------------------------------------------
// Launch configuration (extension).

extern "C" int wd_set_force_variant(wd_sim* s, int block, int lanes) {
    WD_CHECK_SIM(s);
    int v = -1;
    for (int k = 0; k < static_cast<int>(sizeof(kForceVariants) / sizeof(kForceVariants[0])); ++k) {
        if (kForceVariants[k].block == block && kForceVariants[k].lanes == lanes) v = k;
    }
    if (!(block == 0 && lanes == 0) && v < 0) {
        return fail(WD_ERR_ARG, "wd_set_force_variant: (block, lanes) must be (0, 0) (autotune), "
                                "(64|128|256, 0) or (64, 1|2|4)");
    }
    WD_HIP(hipStreamSynchronize(s->stream));
    if (s->tuning.pos >= 0) {
        // a tuning cut short: the tile and binning it was timing go back
        if (!s->scan_fixed) s->scan_var = s->tuning.scan_before;
        if (!s->bin_fixed) s->bucket_bin = s->tuning.bin_before;
    }
    s->force_var = v;
    s->force_fixed = v >= 0;
    s->next_tune = s->steps_run;  // (0, 0): tune from the next call on
    s->tuning.pos = -1;
    for (double& t : s->tune_ms) t = 0.0;
    s->drop_graphs();
    return WD_OK;
}

extern "C" int wd_set_scan_tile(wd_sim* s, int cells) {
    WD_CHECK_SIM(s);
    int v = -1;
    for (int k = 0; k < 3; ++k) {
        if (kScanThreads * kScanItemsVariants[k] == cells) v = k;
    }
    if (cells != 0 && v < 0) {
        return fail(WD_ERR_ARG, "wd_set_scan_tile: cells must be 0 (autotune), 1024, 2048 or 4096");
    }
    WD_HIP(hipStreamSynchronize(s->stream));
    s->scan_fixed = v >= 0;
    if (v >= 0 && v != s->scan_var) {
        s->scan_var = v;
        s->drop_graphs();
    }
    return WD_OK;
}

extern "C" int wd_get_scan_tile(const wd_sim* s, int* cells) {
    WD_CHECK_SIM(s);
    if (!cells) return fail(WD_ERR_ARG, "wd_get_scan_tile: null output");
    *cells = kScanThreads * kScanItemsVariants[s->scan_var];
    return WD_OK;
}

extern "C" int wd_set_binning(wd_sim* s, int mode) {
    WD_CHECK_SIM(s);
    if (mode < WD_BINNING_TUNED || mode > WD_BINNING_BUCKETS) {
        return fail(WD_ERR_ARG, "wd_set_binning: mode must be 0 (tuned), 1 (scatter) or 2 (buckets)");
    }
    if (mode == WD_BINNING_BUCKETS && !s->bucket_ok) {
        return fail(WD_ERR_ARG, "wd_set_binning: this handle's buckets would exceed 1 GiB and were not allocated");
    }
    WD_HIP(hipStreamSynchronize(s->stream));
    s->bin_fixed = mode != WD_BINNING_TUNED;
    const bool want = mode == WD_BINNING_BUCKETS;
    if (s->bin_fixed && want != s->bucket_bin) {
        s->bucket_bin = want;
        s->drop_graphs();
    }
    return WD_OK;
}

extern "C" int wd_get_binning(const wd_sim* s, int* mode) {
    WD_CHECK_SIM(s);
    if (!mode) return fail(WD_ERR_ARG, "wd_get_binning: null output");
    *mode = s->bucket_bin ? WD_BINNING_BUCKETS : WD_BINNING_SCATTER;
    return WD_OK;
}

extern "C" int wd_set_graph_replay(wd_sim* s, int on) {
    WD_CHECK_SIM(s);
    if (on != 0 && on != 1) return fail(WD_ERR_ARG, "wd_set_graph_replay: on must be 0 or 1");
    WD_HIP(hipStreamSynchronize(s->stream));
    s->graphs_on = on == 1;
    return WD_OK;
}

extern "C" int wd_set_fused_steps(wd_sim* s, int on) {
    WD_CHECK_SIM(s);
    if (on != 0 && on != 1) return fail(WD_ERR_ARG, "wd_set_fused_steps: on must be 0 or 1");
    WD_HIP(hipStreamSynchronize(s->stream));
    if (s->fused_on != (on == 1)) {
        s->fused_on = on == 1;
        s->fused_ready = false;
        s->drop_graphs();
    }
    return WD_OK;
}

extern "C" int wd_get_fused_steps(const wd_sim* s, int* on) {
    WD_CHECK_SIM(s);
    if (!on) return fail(WD_ERR_ARG, "wd_get_fused_steps: null output");
    *on = s->use_fused() ? 1 : 0;
    return WD_OK;
}

extern "C" int wd_fused_misses(wd_sim* s, unsigned long long* misses) {
    WD_CHECK_SIM(s);
    if (!misses) return fail(WD_ERR_ARG, "wd_fused_misses: null output");
    WD_HIP(hipStreamSynchronize(s->stream));
    WD_HIP(hipMemcpy(misses, &s->ctl->bin_misses, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return WD_OK;
}

// ---------------------------------------------------------------------------
// Collision cell size (extension): coarser cells than the configured grid.

namespace {

int apply_cell_scale(wd_sim* s, float scale) {
    WD_HIP(hipStreamSynchronize(s->stream));
    s->fused_ready = false;
    s->cell_scale = scale;
    int rc = refresh_params(s);
    if (rc != WD_OK || s->n == 0) return rc;
    // the grains' cell keys and boxes under the new grid, for the next bin
    const int gb = grain_blocks(s->n);
    hipLaunchKernelGGL(k_rekey, dim3(gb), dim3(kBlock), 0, s->stream, (const Params*)s->dp, s->parts,
                       (const float4*)s->Tp, (const float4*)s->Tv, s->key, s->n);
    s->finalize(gb, kFinBox);
    WD_HIP(hipGetLastError());
    return s->sync();
}

}  // namespace

extern "C" int wd_set_cell_scale(wd_sim* s, float scale) {
    WD_CHECK_SIM(s);
    if (!(scale >= 1.f && scale <= 4.f)) return fail(WD_ERR_ARG, "wd_set_cell_scale: scale must be in [1, 4]");
    return apply_cell_scale(s, scale);
}

extern "C" int wd_get_cell_scale(const wd_sim* s, float* scale, int* grid) {
    WD_CHECK_SIM(s);
    if (scale) *scale = s->cell_scale;
    if (grid) grid_dims(s->cfg, s->cell_scale, grid);
    return WD_OK;
}

extern "C" int wd_tune_cell_scale(wd_sim* s, int steps, float* chosen) {
    WD_CHECK_SIM(s);
    if (steps < 2) return fail(WD_ERR_ARG, "wd_tune_cell_scale: at least 2 steps per candidate");
    if (s->picard()) return fail(WD_ERR_STATE, "wd_tune_cell_scale: semi-implicit Euler only");
    if (s->n == 0) {
        if (chosen) *chosen = s->cell_scale;
        return WD_OK;
    }
    // Every candidate runs from the same state, which is restored at the end
    // (with the counters the trial steps moved): tuning leaves the
    // simulation where it was, and a trial grid that overflows a cell drops
    // neighbours only in steps that are thrown away.
    const size_t n = static_cast<size_t>(s->n);
    wd_sim::Snap sn;
    std::vector<void*> bufs;
    auto alloc = [&](void** p, size_t bytes) {
        *p = nullptr;
        const bool ok = hipMalloc(p, bytes) == hipSuccess;
        if (ok) bufs.push_back(*p);
        return ok;
    };
    auto release = [&]() {
        for (void* p : bufs) (void)hipFree(p);
    };
    if (!alloc(reinterpret_cast<void**>(&sn.P), sizeof(float4) * n) ||
        !alloc(reinterpret_cast<void**>(&sn.V), sizeof(float4) * n) ||
        !alloc(reinterpret_cast<void**>(&sn.H), sizeof(float4) * kHist * n) ||
        !alloc(reinterpret_cast<void**>(&sn.C), sizeof(int) * n)) {
        release();
        return fail(WD_ERR_HIP, "wd_tune_cell_scale: snapshot allocation failed");
    }
    WD_HIP(hipStreamSynchronize(s->stream));
    Ctl ctl0;
    std::vector<unsigned long long> movf0(static_cast<size_t>(s->members));
    WD_HIP(hipMemcpy(&ctl0, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    WD_HIP(hipMemcpy(movf0.data(), s->d_movf, sizeof(unsigned long long) * movf0.size(), hipMemcpyDeviceToHost));
    snap_save(s, sn);
    const float scale0 = s->cell_scale;
    const bool energy = s->energy_on;
    s->energy_on = false;
    const float cand[3] = {1.f, 1.25f, 1.5f};
    double best_ms = 1e30;
    float best = 1.f;
    int rc = WD_OK;
    for (float sc : cand) {
        s->cell_scale = sc;
        rc = refresh_params(s);
        if (rc != WD_OK) break;
        snap_load(s, sn);  // the snapshot's keys and boxes under this grid
        unsigned long long over0 = 0, over1 = 0;
        rc = read_overflow(s, &over0);
        if (rc != WD_OK) break;
        double total = 0.0;
        for (int k = 0; k < steps && rc == WD_OK; ++k) {
            std::vector<wd_sim::ProfRec> recs;
            s->prof = &recs;
            s->enqueue_step();
            s->prof = nullptr;
            if (hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(WD_ERR_HIP, "wd_tune_cell_scale: step failed");
            double ms = 0.0;
            for (auto& rec : recs) {
                float e = 0.f;
                if (hipEventElapsedTime(&e, rec.a, rec.b) == hipSuccess) ms += e;
                s->spare.push_back(rec.a);
                s->spare.push_back(rec.b);
            }
            if (k > 0) total += ms;  // the first step warms the grid up
        }
        if (rc != WD_OK) break;
        rc = read_overflow(s, &over1);
        if (rc != WD_OK) break;
        const double mean = total / (steps - 1);
        if ((over1 == over0 || sc == 1.f) && mean < best_ms) {
            best_ms = mean;
            best = sc;
        }
    }
    s->energy_on = energy;
    s->cell_scale = rc == WD_OK ? best : scale0;
    int rc2 = refresh_params(s);
    if (rc2 == WD_OK) {
        snap_load(s, sn);
        rc2 = s->sync();
    }
    if (rc2 == WD_OK) {
        // counters as before the trials (the spring generation stays where
        // the restored springs were written)
        Ctl now;
        rc2 = hipMemcpy(&now, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost) == hipSuccess ? WD_OK
                : fail(WD_ERR_HIP, "wd_tune_cell_scale: counters");
        if (rc2 == WD_OK) {
            ctl0.hist_parity = now.hist_parity;
            ctl0.epoch = now.epoch;
            ctl0.tile_counter = now.tile_counter;
            ctl0.bin_par = now.bin_par;  // with the spill counters, as the last scan left them
            ctl0.spill_n[0] = now.spill_n[0];
            ctl0.spill_n[1] = now.spill_n[1];
            if (hipMemcpy(s->ctl, &ctl0, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(s->d_movf, movf0.data(), sizeof(unsigned long long) * movf0.size(),
                          hipMemcpyHostToDevice) != hipSuccess) {
                rc2 = fail(WD_ERR_HIP, "wd_tune_cell_scale: counters");
            }
        }
    }
    release();
    if (rc != WD_OK) return rc;
    if (rc2 != WD_OK) return rc2;
    if (chosen) *chosen = best;
    return s->check_device();
}

extern "C" int wd_get_force_variant(const wd_sim* s, int* block, int* lanes) {
    WD_CHECK_SIM(s);
    if (!block || !lanes) return fail(WD_ERR_ARG, "wd_get_force_variant: null output");
    const ForceVariant fv = s->force_var < 0 ? ForceVariant{0, 0} : kForceVariants[s->force_var];
    *block = fv.block;
    *lanes = fv.lanes;
    return WD_OK;
}

extern "C" int wd_set_force_block(wd_sim* s, int block) {
    WD_CHECK_SIM(s);
    if (block != 0 && block != 64 && block != 128 && block != 256) {
        return fail(WD_ERR_ARG, "wd_set_force_block: block must be 0 (autotune), 64, 128 or 256");
    }
    return wd_set_force_variant(s, block, 0);
}

extern "C" int wd_get_force_block(const wd_sim* s, int* block) {
    WD_CHECK_SIM(s);
    if (!block) return fail(WD_ERR_ARG, "wd_get_force_block: null output");
    int lanes = 0;
    return wd_get_force_variant(s, block, &lanes);
}

extern "C" int wd_contact_stats(wd_sim* s, unsigned long long* passes, unsigned long long* contacts) {
    WD_CHECK_SIM(s);
    if (!passes || !contacts) return fail(WD_ERR_ARG, "wd_contact_stats: null output");
    WD_HIP(hipStreamSynchronize(s->stream));
    WD_HIP(hipMemcpy(passes, &s->ctl->force_passes, sizeof(*passes), hipMemcpyDeviceToHost));
    WD_HIP(hipMemcpy(contacts, &s->ctl->contact_sum, sizeof(*contacts), hipMemcpyDeviceToHost));
    return WD_OK;
}

// ---------------------------------------------------------------------------
// Per-member diagnostics of a batch.

extern "C" int wd_batch_overflow_counts(wd_sim* s, int* counts) {
    WD_CHECK_SIM(s);
    if (!counts) return fail(WD_ERR_ARG, "wd_batch_overflow_counts: null output");
    int rc = s->sync();
    if (rc != WD_OK) return rc;
    std::vector<unsigned long long> v(static_cast<size_t>(s->members));
    WD_HIP(hipMemcpy(v.data(), s->d_movf, sizeof(unsigned long long) * v.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < s->members; ++m) counts[m] = static_cast<int>(std::min<unsigned long long>(v[m], 0x7FFFFFFFull));
    return WD_OK;
}

extern "C" int wd_batch_is_stable(wd_sim* s, int* stable) {
    WD_CHECK_SIM(s);
    if (!stable) return fail(WD_ERR_ARG, "wd_batch_is_stable: null output");
    std::vector<float> st(static_cast<size_t>(6) * s->n);
    const int rc = s->n ? wd_get_state(s, st.data()) : WD_OK;
    if (rc != WD_OK) return rc;
    for (int m = 0; m < s->members; ++m) {
        const wd_config& c = s->mcfg[m];
        const float lo[3] = {c.x_min, c.y_min, c.z_min};
        const float hi[3] = {c.x_max, c.y_max, c.z_max};
        int ok = 1;
        for (int i = s->goff[m]; i < s->goff[m + 1] && ok; ++i) {
            for (int a = 0; a < 3; ++a) {
                const float p = st[6 * static_cast<size_t>(i) + a];
                if (!std::isfinite(p) || p < lo[a] || p > hi[a]) {
                    ok = 0;
                    break;
                }
            }
        }
        stable[m] = ok;
    }
    return WD_OK;
}
