// Kernel declarations and the device control block of one simulation handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_device.h"

namespace wd {

constexpr int kBlock = 256;          // threads per block of the per-grain kernels
constexpr int kScanThreads = 256;
// Cells per thread of the cell scan: tiles of 1,024, 2,048 or 4,096 cells,
// chosen by the online tuner (the start offsets do not depend on it).
constexpr int kScanItemsVariants[3] = {4, 8, 16};
constexpr int kScanTileMin = kScanThreads * 4;
constexpr int kScanTileMax = kScanThreads * 16;
constexpr int kXcds = 8;             // MI355X accelerator complex dies
constexpr int kFinalizeBlock = 1024;  // one block; one wave per ensemble member at a time
constexpr int kContactList = 24;     // touching candidates buffered per grain in LDS
constexpr int kQueueCap = 16;        // contacts listed per lane and pass by the queued force pass
constexpr int kEnergyGroup = 16;     // grains per energy partial (ewave entry)

// Small device-side control block: counters, the occupied-cell bounding box
// and per-pass results.  Only k_count (block 0), k_scan (block 0) and
// k_finalize write it outside of atomics, so no pass needs a grid-wide fence.
struct Ctl {
    unsigned tile_counter;     // scan tile dispenser; zeroed by k_count
    unsigned epoch;            // scan look-back tag; bumped by k_count
    unsigned accel_last;       // float bits, max |a| of the last completed force pass
    unsigned long long overflow_lifetime;  // over-capacity cells, summed over passes
    int error;                 // nonzero: a bounded wait gave up
    int energy_count;          // entries written to the energy log
    int energy_capacity;
    int hist_parity;           // current contact-history generation; flipped by k_finalize
    double energy_last;        // last standalone energy
    unsigned long long force_passes;  // completed force passes (k_finalize)
    unsigned long long contact_sum;   // contacts (walls + pairs) summed over those passes
    unsigned cic_log_next;            // density-log slots taken (k_cic_zero of a logged deposit)
    // Bucket binning: the count pass files grains past a cell's bucket in
    // spill list bin_par (entries spill_n[bin_par]); k_scan empties the other
    // list and flips bin_par, so the reorder reads list bin_par ^ 1.
    unsigned bin_par;
    unsigned spill_n[2];
    // Fused step (a lone simulation): the force pass counts its grains into
    // the pending grid and sets bin_miss when one lies outside it; the next
    // scan then bins the pass's grains itself, on one block, and clears it.
    int bin_miss;
    unsigned long long bin_misses;  // steps binned that way
    // Over-capacity cells of a fused step's bin, which belongs to the next
    // step: the force pass that uses it adds them to the totals (and a
    // classic bin that replaces it clears them), so the totals count the
    // neighbor passes of the steps taken, as in the classic sequence.
    unsigned long long ovf_next;
};

// The count of a force pass that bins for the next step (fused step): the
// grain's slot in its cell of the pending grid goes to slot[p] (mode 1) or its
// id into the cell's bucket (mode 2; past kBucket, the spill list, as
// k_count).  mode 0: the pass counts nothing.
struct BinIO {
    uint32_t* count;
    uint32_t* slot;
    uint32_t* bucket;
    uint2* spill;
    int mode;
};

// Ids a cell's bucket holds (bucket binning): the count pass writes a grain's
// id at bucket[cell * kBucket + slot] when its slot is below kBucket.
constexpr int kBucket = 8;

// Per-block partial results of a grain pass, reduced by k_finalize, in two
// arrays indexed by block.  A block covers consecutive grains, so it holds a
// run of members m0..m1; the boxes of its first and last member are reduced
// here (members strictly between them lie wholly inside the block and are
// reduced by k_finalize from the cell keys).  A lone simulation's pass
// writes and reads only the first array (32 B per block).
struct BlockPart {
    int box0[6];      // min cx, cy, cz / max cx, cy, cz of member m0's grains
    unsigned accel;   // float bits of the block's max |a| (or deposit magnitude)
    unsigned contacts;  // contacts (walls + pairs) of the block's grains
};
struct BlockMembers {
    int m0, m1;       // first and last member of the block's grains (-1: none)
    int box1[6];      // the box of member m1 (when m1 != m0)
};
static_assert(sizeof(BlockPart) == 32 && sizeof(BlockMembers) == 32, "block partials are 32 B");

// The reduction of a lone simulation's block partials runs over kFinSlices
// fixed slices of the partial and energy-group arrays: one wave reduces a
// slice (lane l adds the slice's entries l, l + 64, ... in order, then the
// fixed DPP tree), and one wave combines the slices (lane r holds slice r,
// the same tree).  The bits of the energy sum therefore do not depend on how
// many blocks or threads take part: k_finalize runs them on its one block, a
// fused step on the cell scan's last kFinSliceBlocks blocks, and the
// reorder's block 0 combines them.
constexpr int kFinSlices = 64;
constexpr int kFinSliceBlocks = kFinSlices / (kScanThreads / 64);
struct FinSlice {
    int box[6];
    unsigned accel;
    unsigned pad;
    double contacts;
    double energy;
};

// The cell scan of a fused step (on != 0): it scans the next grid, and its
// last kFinSliceBlocks blocks reduce the slices of the partials of the force
// pass that counted the grains instead of taking tiles (energy groups too
// when with_e); after a miss the last of them also bins the grains (key, Tp)
// itself and the tile blocks return.
struct ScanFuse {
    int on;
    const BlockPart* parts;
    int nparts;
    const double* ewave;
    int ng;
    int with_e;
    FinSlice* slices;
    BinIO bio;
    const uint32_t* key;
    const float4* Tp;
    int n;
};

// The reorder of a fused step (on != 0): its block 0 makes the next grid the
// current one and combines the scan's slices, publishing as k_finalize would
// with `flags`.
struct ReorderFuse {
    int on;
    const FinSlice* slices;
    double* energy_log;
    int flags;
    int* box_out;         // when set: the grid the bin used (box and cell offsets), as k_count
    uint32_t* cell_out;   // keeps it for a recomputed step's slot
};

// What k_finalize publishes.
enum FinalizeFlags {
    kFinBox = 1,        // member boxes and cell offsets <- reduced boxes
    kFinAccel = 2,      // accel_last <- reduced max
    kFinEnergyLog = 4,  // append the energy sum to the log
    kFinEnergyLast = 8, // energy_last <- energy sum
    kFinCicScale = 16,  // per-member CIC magnitudes <- reduced maxima
    kFinFlipHistory = 32,  // the step is complete: its springs become current
    kFinPending = 64    // with kFinBox: the box goes to the pending grid only
};

// Launch geometry helpers: grid rounded up to a multiple of the XCD count,
// logical block b maps to a contiguous slab so neighbouring grains share an
// XCD's L2.
inline int grain_blocks(int n, int bs = kBlock) {
    int b = (n + bs - 1) / bs;
    return ((b + kXcds - 1) / kXcds) * kXcds;
}

// Force-pass variants the autotuner chooses from; the results do not depend
// on the choice.  lanes 0: k_force, one lane per grain, `block` grains per
// block.  lanes L > 0: k_force_q, L lanes per grain, one wave (64 / L grains)
// per block.
struct ForceVariant {
    int block;
    int lanes;
};
constexpr ForceVariant kForceVariants[6] = {{64, 0}, {128, 0}, {256, 0}, {64, 1}, {64, 2}, {64, 4}};
constexpr int kForceBlocks[3] = {64, 128, 256};
// Grains per block of a variant (grains per block partial).
constexpr int variant_grains(ForceVariant v) { return v.lanes == 0 ? v.block : 64 / v.lanes; }

__device__ __forceinline__ int logical_block() {
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

enum ForceMode { kModeSIE = 0, kModeBEP = 1, kModeRHS = 2 };

// Contact-history buffers (slot-major [kHist][n]).  Two generations: the
// force pass reads generation ctl->hist_parity and writes the other one;
// k_finalize flips the parity once a step is complete.  A generation is
// indexed by the T position of its grain (the cell-sorted position at which
// the pass that wrote it found the grain), so that consecutive grains of the
// next pass read and write consecutive entries; hidx[p] (written by
// k_reorder) is the index of cell-sorted grain p in the generation it reads.
struct HistoryIO {
    float4* buf[2];
    int* cnt[2];
    const int* hidx;
};

struct Carry {
    const float4* nin_p;  // Picard base state (NMODE 2)
    const float4* nin_v;
    float4* nout_p;       // Picard base state out (NMODE 1, 2)
    float4* nout_v;
    const int* hidx_in;   // spring index of T grain i (NMODE 2; else i itself)
    int* hidx_out;        // spring index of cell-sorted grain p (null: no springs)
};

}  // namespace wd
