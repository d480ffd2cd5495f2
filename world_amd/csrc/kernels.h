// Kernel declarations and the device control block of one simulation handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dem_device.h"

namespace wd {

constexpr int kBlock = 256;          // threads per block of the per-grain kernels
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // cells per scan tile
constexpr int kXcds = 8;             // MI355X accelerator complex dies
constexpr int kContactList = 16;     // touching candidates buffered per grain in LDS

// Small device-side control block: counters, the occupied-cell bounding box
// and per-pass results.  Only k_count (block 0) and k_finalize write it
// outside of atomics, so no pass needs a grid-wide fence.
struct Ctl {
    unsigned tile_counter;     // scan tile dispenser; zeroed by k_count
    unsigned epoch;            // scan look-back tag; bumped by k_count
    int bbox_cur[6];           // occupied cells (min xyz, max xyz) of the current state
    unsigned accel_last;       // float bits, max |a| of the last completed force pass
    unsigned cic_scale_bits;   // float bits, max per-grain deposit magnitude
    unsigned long long overflow_lifetime;  // over-capacity cells, summed over passes
    int error;                 // nonzero: a bounded wait gave up
    int energy_count;          // entries written to the energy log
    int energy_capacity;
    int hist_parity;           // current contact-history generation; flipped by k_finalize
    double energy_last;        // last standalone energy
};

// Per-block partial results of a grain pass, reduced by k_finalize.
struct BlockPart {
    int box[6];       // min cx, cy, cz / max cx, cy, cz of the block's grains
    unsigned accel;   // float bits of the block's max |a| (or deposit magnitude)
    unsigned pad;
    double energy;    // deterministic block sum of per-grain energies
};

// What k_finalize publishes.
enum FinalizeFlags {
    kFinBox = 1,        // bbox_cur <- reduced box
    kFinAccel = 2,      // accel_last <- reduced max
    kFinEnergyLog = 4,  // append the energy sum to the log
    kFinEnergyLast = 8, // energy_last <- energy sum
    kFinCicScale = 16,  // cic_scale_bits <- reduced max
    kFinFlipHistory = 32  // the step is complete: its springs become current
};

// Launch geometry helpers: grid rounded up to a multiple of the XCD count,
// logical block b maps to a contiguous slab so neighbouring grains share an
// XCD's L2.
inline int grain_blocks(int n) {
    int b = (n + kBlock - 1) / kBlock;
    return ((b + kXcds - 1) / kXcds) * kXcds;
}

__device__ __forceinline__ int logical_block() {
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

enum ForceMode { kModeSIE = 0, kModeBEP = 1, kModeRHS = 2 };

// Contact-history buffers (slot-major [kHist][n], indexed by ORIGINAL grain
// id, so they never move with the cell order).  Two generations: the force
// pass reads generation ctl->hist_parity and writes the other one;
// k_finalize flips the parity once a step is complete.
struct HistoryIO {
    float4* buf[2];
    int* cnt[2];
};

// What k_reorder carries along the permutation besides the grain.
struct Carry {
    const float4* nin_p;  // Picard base state (NMODE 2)
    const float4* nin_v;
    float4* nout_p;       // Picard base state out (NMODE 1, 2)
    float4* nout_v;
};

}  // namespace wd
