// gfx950 kernels of the per-step granular pipeline.
//
// Grain data lives in two float4 streams per grain: position with the
// original grain index in .w, velocity with the material id in .w (32 B per
// grain, one dwordx4 load each).  A step is
//
//   k_count   -> slot of every grain in its cell (atomic per-cell counter);
//                with bucket binning, the grain's id into its cell's bucket
//   k_scan    -> exclusive scan of the counts over the occupied sub-grid,
//                one pass with decoupled look-back; counts re-zeroed here
//   k_scatter -> original ids into cell-sorted slots (not with buckets)
//   k_reorder -> rank inside the cell by original id, move the grain to its
//                cell-sorted position (coalesced float4 streams)
//   k_force   -> walls + 27-cell pair contacts + integration, fused; writes
//                the next state and its cell keys, reduces the next bounding
//                box, max |a| and (optionally) the total energy
//
// The cell order (flat cell, then original index inside a cell) is a
// deterministic instance of the reference's grid (src/find_neighbors_kernel.cu:
// 43-161), whose in-cell order depends on atomic races.  Pair forces are
// summed in stencil order dx -> dy -> dz as the reference's neighbor rows are.
// A cell holding more than particles_per_cell grains contributes only its
// first particles_per_cell grains to every neighbor list, and counts once
// per pass towards the overflow total (src/find_neighbors_kernel.cu:81-112).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "kernels.h"

namespace wd {

namespace {

constexpr int kEmptyMin = 0x7FFFFFFF;
constexpr int kEmptyMax = (-0x7FFFFFFF - 1);

// Deterministic sum over the wave (every lane active): a fixed tree of DPP
// row moves, pairs, quads, half rows, rows, then rows 0-1 into row 1, rows
// 2-3 into row 3 and rows 0-1 into row 3; the result is lane 63's, the same
// for every lane.  Lanes a move does not reach add +0.0 (no effect on lane
// 63's sum).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move_f64<0xB1, 0xF>(v);   // quad_perm 1,0,3,2
    v += dpp_move_f64<0x4E, 0xF>(v);   // quad_perm 2,3,0,1
    v += dpp_move_f64<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_move_f64<0x140, 0xF>(v);  // row_mirror
    v += dpp_move_f64<0x142, 0xA>(v);  // row_bcast:15
    v += dpp_move_f64<0x143, 0xC>(v);  // row_bcast:31
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}
// Deterministic sum of one double per grain over kEnergyGroup consecutive
// grains held L lanes apart: a butterfly over the grain index, so the bits do
// not depend on L or on the block size.
template <int L>
__device__ __forceinline__ double group_energy(double v) {
#pragma unroll
    for (int o = L; o < kEnergyGroup * L; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Wave-wide reduction of an associative, exact integer operation through DPP
// row moves (a few cycles each, against an LDS round trip per step for
// cross-lane shuffles): pairs, quads, half rows and rows, then rows 0-1 into
// row 1 and rows 0-3 into row 3.  Lanes a move does not reach take
// `identity`.  Every lane must be active; the result is lane 63's, the same
// for every lane.
template <class Op>
__device__ __forceinline__ int wave_reduce_dpp(int v, int identity, Op op) {
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(identity, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
// Inclusive prefix sum over the wave through DPP: shifts by 1, 2, 4, 8 within
// rows of 16, then row 0's total into row 1 and row 2's into row 3, then
// rows 0-1's into rows 2-3.  Every lane must be active.
__device__ __forceinline__ int wave_incl_scan_dpp(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}
__device__ __forceinline__ int wave_min_dpp(int v) {
    return wave_reduce_dpp(v, kEmptyMin, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max_dpp(int v) {
    return wave_reduce_dpp(v, kEmptyMax, [](int a, int b) { return max(a, b); });
}
__device__ __forceinline__ unsigned wave_max_u_dpp(unsigned v) {
    return static_cast<unsigned>(wave_reduce_dpp(static_cast<int>(v), 0, [](int a, int b) {
        return static_cast<int>(max(static_cast<unsigned>(a), static_cast<unsigned>(b)));
    }));
}
__device__ __forceinline__ unsigned wave_sum_u_dpp(unsigned v) {
    return static_cast<unsigned>(wave_reduce_dpp(static_cast<int>(v), 0, [](int a, int b) {
        return static_cast<int>(static_cast<unsigned>(a) + static_cast<unsigned>(b));
    }));
}


// Set bits of a wave mask below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u)));
}

// Index of the k-th set bit (k from 0) of a 64-bit lane mask that has more
// than k bits set: a binary search on population counts.
__device__ __forceinline__ int nth_lane(unsigned long long m, int k) {
    int pos = 0;
    int c = __popc(static_cast<uint32_t>(m));
    if (k >= c) {
        k -= c;
        m >>= 32;
        pos = 32;
    }
    uint32_t x = static_cast<uint32_t>(m);
#pragma unroll
    for (int w = 16; w >= 1; w >>= 1) {
        c = __popc(x & ((1u << w) - 1u));
        if (k >= c) {
            k -= c;
            x >>= w;
            pos += w;
        }
    }
    return pos;
}

// True when a column's cell-start quad (starts of cells dz = -1, 0, +1 and of
// the next cell) has a cell holding more than kcap grains.
__device__ __forceinline__ bool column_capped(const uint4& q4, uint32_t kcap) {
    return q4.y - q4.x > kcap || q4.z - q4.y > kcap || q4.w - q4.z > kcap;
}

// Number of candidates of a column: its three cells' grains, at most kcap
// from each.
__device__ __forceinline__ uint32_t column_length(const uint4& q4, uint32_t kcap) {
    return column_capped(q4, kcap) ? min(q4.y - q4.x, kcap) + min(q4.z - q4.y, kcap) + min(q4.w - q4.z, kcap)
                                   : q4.w - q4.x;
}

// Position of cell-sorted slot q: a 32-bit byte offset from the array base
// (grain counts stay below 2^28), so the load takes the scalar-base form and
// no 64-bit address is formed per candidate.
__device__ __forceinline__ const float* slot_pos(const float4* __restrict__ Sp, uint32_t q) {
    return reinterpret_cast<const float*>(reinterpret_cast<const char*>(Sp) + (q << 4));
}

// Candidate positions for the stencil walk, from the cell-sorted state in
// global memory.
struct GlobalPos {
    const float4* __restrict__ Sp;
    __device__ __forceinline__ void operator()(uint32_t q, float& x, float& y, float& z) const {
        const float* s = slot_pos(Sp, q);
        x = s[0];
        y = s[1];
        z = s[2];
    }
};
// Slot of `partner` in an old spring list (slot-major with stride n, grain
// index hs, h_old entries), or -1 when it holds no such partner.  The
// partner ids are loaded eight at a time, all in flight together, rather than
// one dependent load per slot: the batch is compared with selects only, so
// no load can sink behind an early exit.
__device__ __forceinline__ int spring_slot(const float4* __restrict__ h_in, int n, int hs, int h_old,
                                           int partner) {
    const int* __restrict__ ids = reinterpret_cast<const int*>(h_in) + 3;  // .w of each entry
    for (int t0 = 0; t0 < h_old; t0 += 8) {
        int w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = min(t0 + k, h_old - 1);  // slots past the list reload its last entry
            w[k] = ids[(static_cast<size_t>(t) * n + hs) * 4];
        }
        int found = -1;  // the first match of the batch (partners are distinct)
#pragma unroll
        for (int k = 7; k >= 0; --k) found = (t0 + k < h_old && w[k] == partner) ? t0 + k : found;
        if (found >= 0) return found;
    }
    return -1;
}

// Walks one grain's stencil candidates in the reference's neighbor-row order
// (columns dx, then dy; the three cells dz = -1, 0, +1 of a column; at most
// kcap grains of a cell) and calls probe(q, x, y, z) with the cell-sorted slot
// and position of those at positions [s_lo, s_hi) of that sequence.  A run
// (one column, or with GROUPED the three columns of one dx) is walked with
// BATCH position loads in flight, while the next run's cell starts load.
// column(c) returns column c's cell-start quad.  GROUPED walks only the dx
// groups [gx_lo, gx_hi) (columns 3 gx_lo .. 3 gx_hi - 1).  With FLAT, probe is called
// for every slot of a batch, as probe(q, x, y, z, valid), and must ignore the
// slots past the run (valid false) itself: it can then run without branches.
template <int BATCH, bool GROUPED, bool FLAT = false, class Pos, class Column, class Probe>
__device__ __forceinline__ void walk_stencil(Pos pos, uint32_t kcap, uint32_t s_lo, uint32_t s_hi,
                                             Column column, Probe probe, int gx_lo = 0, int gx_hi = 3) {
    uint32_t off = 0;  // sequence position of the run's first candidate
    // Candidates b0 + [0, len): a column with no cell over capacity.
    auto run1 = [&](uint32_t b0, uint32_t len) __attribute__((always_inline)) {
        const uint32_t i0 = min(s_lo > off ? s_lo - off : 0u, len);
        const uint32_t i1 = s_hi > off ? min(s_hi - off, len) : 0u;
        for (uint32_t i = i0; i < i1; i += BATCH) {
            float px[BATCH], py[BATCH], pz[BATCH];
            // Every slot loads (one past the run reloads the run's last
            // candidate), so that all loads are in flight together.
#pragma unroll
            for (int k = 0; k < BATCH; ++k) pos(b0 + min(i + k, i1 - 1), px[k], py[k], pz[k]);
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                if constexpr (FLAT) probe(b0 + i + k, px[k], py[k], pz[k], i + k < i1);
                else if (i + k < i1) probe(b0 + i + k, px[k], py[k], pz[k]);
            }
        }
        off += len;
    };
    // Candidates b0 + [0, l0), b1 + [0, l1), b2 + [0, l2) in that order.
    auto run3 = [&](uint32_t b0, uint32_t l0, uint32_t b1, uint32_t l1, uint32_t b2, uint32_t l2)
            __attribute__((always_inline)) {
        const uint32_t tot = l0 + l1 + l2;
        const uint32_t i0 = min(s_lo > off ? s_lo - off : 0u, tot);
        const uint32_t i1 = s_hi > off ? min(s_hi - off, tot) : 0u;
        for (uint32_t i = i0; i < i1; i += BATCH) {
            uint32_t q[BATCH];
            float px[BATCH], py[BATCH], pz[BATCH];
            // Every slot loads (a slot past the run reloads the first
            // candidate), so that no load sits behind a branch and all are
            // in flight together.
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                const uint32_t t = i + k;
                q[k] = t < l0 ? b0 + t : (t < l0 + l1 ? b1 + (t - l0) : b2 + (t - l0 - l1));
                pos(t < i1 ? q[k] : q[0], px[k], py[k], pz[k]);
            }
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                if constexpr (FLAT) probe(q[k], px[k], py[k], pz[k], i + k < i1);
                else if (i + k < i1) probe(q[k], px[k], py[k], pz[k]);
            }
        }
        off += tot;
    };
    if constexpr (!GROUPED) {
        uint4 next = column(3 * gx_lo);
#pragma unroll 1
        for (int c = 3 * gx_lo; c < 3 * gx_hi; ++c) {
            const uint4 a = next;
            if (c + 1 < 3 * gx_hi) next = column(c + 1);
            if (!column_capped(a, kcap)) {
                run1(a.x, a.w - a.x);
            } else {
                run3(a.x, min(a.y - a.x, kcap), a.y, min(a.z - a.y, kcap), a.z, min(a.w - a.z, kcap));
            }
        }
        return;
    }
    uint4 n0 = column(3 * gx_lo), n1 = column(3 * gx_lo + 1), n2 = column(3 * gx_lo + 2);
#pragma unroll 1
    for (int gx = gx_lo; gx < gx_hi; ++gx) {
        const uint4 a = n0, b = n1, c = n2;
        if (gx + 1 < gx_hi) {
            n0 = column(3 * gx + 3);
            n1 = column(3 * gx + 4);
            n2 = column(3 * gx + 5);
        }
        if (!column_capped(a, kcap) && !column_capped(b, kcap) && !column_capped(c, kcap)) {
            run3(a.x, a.w - a.x, b.x, b.w - b.x, c.x, c.w - c.x);
        } else {
            run3(a.x, min(a.y - a.x, kcap), a.y, min(a.z - a.y, kcap), a.z, min(a.w - a.z, kcap));
            run3(b.x, min(b.y - b.x, kcap), b.y, min(b.z - b.y, kcap), b.z, min(b.w - b.z, kcap));
            run3(c.x, min(c.y - c.x, kcap), c.y, min(c.z - c.y, kcap), c.z, min(c.w - c.z, kcap));
        }
    }
}

// A lane's slice [s_lo, s_hi) of a grain's candidate sequence when no cell
// of its stencil is over capacity: column c is the run cb[c] + [0, len), at
// sequence positions cpre[c] + [0, len).  The whole slice is loaded B
// positions at a time, across column boundaries, so a slice of up to B
// candidates costs one load latency however many columns it spans.  probe
// is called in sequence order, as by walk_stencil.
template <int B, class Probe>
__device__ __forceinline__ void walk_slice(const float4* __restrict__ Sp, const uint32_t cb[9],
                                           const uint32_t cpre[9], uint32_t s_lo, uint32_t s_hi, Probe probe) {
    for (uint32_t t0 = s_lo; t0 < s_hi; t0 += B) {
        uint32_t q[B];
        float px[B], py[B], pz[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const uint32_t t = min(t0 + k, s_hi - 1);  // past the slice: reload its last candidate
            uint32_t s = cb[0] + t;
#pragma unroll
            for (int c = 1; c < 9; ++c) {
                if (t >= cpre[c]) s = cb[c] + (t - cpre[c]);
            }
            q[k] = s;
            const float* r = slot_pos(Sp, s);
            px[k] = r[0];
            py[k] = r[1];
            pz[k] = r[2];
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            if (t0 + k < s_hi) probe(q[k], px[k], py[k], pz[k]);
        }
    }
}

// One dx group of a one-lane-per-grain walk, its three columns c0, c0 + 1,
// c0 + 2 in the reference's neighbor-row order: column k with no cell over
// capacity is the run bk + [0, lk); lk = ~0 marks a capped one, whose
// cell-start quad column(c) is loaded again.  Positions from global memory,
// BATCH per round (a slot past a run reloads its last candidate), every slot
// probed without branches as probe(q, x, y, z, valid).
template <int BATCH, class Column, class Probe>
__device__ __forceinline__ void walk_columns3(const float4* __restrict__ Sp, uint32_t kcap, int c0, uint32_t b0,
                                              uint32_t l0, uint32_t b1, uint32_t l1, uint32_t b2, uint32_t l2,
                                              Column column, Probe probe) {
    const GlobalPos pos{Sp};
    auto run = [&](uint32_t b, uint32_t len) __attribute__((always_inline)) {
        for (uint32_t i = 0; i < len; i += BATCH) {
            float px[BATCH], py[BATCH], pz[BATCH];
#pragma unroll
            for (int k = 0; k < BATCH; ++k) pos(b + min(i + k, len - 1), px[k], py[k], pz[k]);
#pragma unroll
            for (int k = 0; k < BATCH; ++k) probe(b + i + k, px[k], py[k], pz[k], i + k < len);
        }
    };
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        const uint32_t b = c == 0 ? b0 : (c == 1 ? b1 : b2);
        const uint32_t len = c == 0 ? l0 : (c == 1 ? l1 : l2);
        if (len != ~0u) {
            run(b, len);
        } else {
            const uint4 q = column(c0 + c);
            run(q.x, min(q.y - q.x, kcap));
            run(q.y, min(q.z - q.y, kcap));
            run(q.z, min(q.w - q.z, kcap));
        }
    }
}

// The same walk from positions staged in LDS (x, y, z records of slots
// base, base + 1, ..., at most `area` of them), with every lane taking part
// (a lane with nothing to walk has runs of length 0): the wave steps through
// each run to its longest one, so that the control flow stays uniform.  A
// round's BATCH reads start at a clamped index and stay in the area; the
// slots past a run are read but not used.  probe(q, x, y, z, valid, self,
// nan): `self` false tells the probe that q is not its own grain (only
// column c = 4, dx = dy = 0, holds it); `nan` false, that no squared
// distance can be NaN (finite = every staged position and the wave's own
// positions are finite: a squared distance of finite positions is finite or
// +inf).
template <int BATCH, class Column, class Probe>
__device__ __forceinline__ void walk_columns3_staged(const float* __restrict__ sxyz, uint32_t base, uint32_t area,
                                                     bool finite, uint32_t kcap, int c0, uint32_t b0,
                                                     uint32_t l0, uint32_t b1, uint32_t l1, uint32_t b2,
                                                     uint32_t l2, Column column, Probe probe) {
    auto run = [&](uint32_t b, uint32_t len, auto self, auto nan) __attribute__((always_inline)) {
        const uint32_t lmax = wave_max_u_dpp(len);
#pragma unroll 1
        for (uint32_t i = 0; i < lmax; i += BATCH) {
            // a clamped round holds no valid slot
            const int r = static_cast<int>(min(b + i - base, area - BATCH));
            __builtin_assume(r >= 0 && r < 4096);  // the area is a few KB: the reads take immediate offsets
            const float* rec = sxyz + 3 * r;
            float px[BATCH], py[BATCH], pz[BATCH];
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                px[k] = rec[3 * k];
                py[k] = rec[3 * k + 1];
                pz[k] = rec[3 * k + 2];
            }
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                probe(b + i + k, px[k], py[k], pz[k], i + k < len, decltype(self)::value, decltype(nan)::value);
            }
        }
    };
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        const uint32_t b = c == 0 ? b0 : (c == 1 ? b1 : b2);
        const uint32_t len = c == 0 ? l0 : (c == 1 ? l1 : l2);
        const bool capped = len == ~0u;
        if (__ballot(capped) == 0ull) {
            if (!finite) run(b, len, std::true_type{}, std::true_type{});
            else if (c0 + c == 4) run(b, len, std::true_type{}, std::false_type{});
            else run(b, len, std::false_type{}, std::false_type{});
        } else {
            // a capped lane walks its three cells' first kcap grains; the
            // others walk their run as the first of three
            uint4 q = make_uint4(b, b + len, b + len, b + len);
            if (capped) q = column(c0 + c);
            const uint32_t m0 = capped ? min(q.y - q.x, kcap) : len;
            const uint32_t m1 = capped ? min(q.z - q.y, kcap) : 0u;
            const uint32_t m2 = capped ? min(q.w - q.z, kcap) : 0u;
            run(q.x, m0, std::true_type{}, std::true_type{});
            run(q.y, m1, std::true_type{}, std::true_type{});
            run(q.z, m2, std::true_type{}, std::true_type{});
        }
    }
}

// Thread tid's share of the energy groups, ewave[tid], ewave[tid + stride],
// ... added in that order, four loads per round issued together (clamped
// indices, so none is predicated).
__device__ __forceinline__ double ewave_share(const double* __restrict__ ewave, int ng, int tid, int stride) {
    double e = 0.0;
    for (int w = tid; w < ng; w += 4 * stride) {
        double x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = ewave[min(w + k * stride, ng - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (w + k * stride < ng) e += x[k];
        }
    }
    return e;
}

// Deterministic block sum of one double per thread (a fixed tree per wave,
// then the waves in order); the result is valid in thread 0.
__device__ double block_sum(double v) {
    __shared__ double s_part[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    v = wave_sum(v);
    if (lane == 0) s_part[wave] = v;
    __syncthreads();
    double total = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < nw; ++w) total += s_part[w];
    }
    __syncthreads();
    return total;
}

// Reduces, over the block, the boxes of its first and last member (lanes
// with member < 0 contribute nothing), the max of `accel` and the sum of
// `contacts`, and stores them with plain stores at parts[slot] and, when pm
// is not null (an ensemble), pm[slot] (thread 0).
__device__ void store_block_part(BlockPart* parts, BlockMembers* pm, int slot, int member, const int lo[3],
                                 const int hi[3], unsigned accel, unsigned contacts = 0u) {
    __shared__ int s_m[16][2];
    __shared__ int s_box[16][14];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    // block's member range
    int mlo = wave_min_dpp(member < 0 ? kEmptyMin : member);
    int mhi = wave_max_dpp(member);
    if (lane == 0) {
        s_m[wave][0] = mlo;
        s_m[wave][1] = mhi;
    }
    __syncthreads();
    mlo = kEmptyMin;
    mhi = -1;
    for (int w = 0; w < nw; ++w) {
        mlo = min(mlo, s_m[w][0]);
        mhi = max(mhi, s_m[w][1]);
    }
    const bool in0 = member >= 0 && member == mlo;
    const bool in1 = member >= 0 && member == mhi && mhi != mlo;
    int v[13];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = wave_min_dpp(in0 ? lo[k] : kEmptyMin);
        v[3 + k] = wave_max_dpp(in0 ? hi[k] : kEmptyMax);
        v[6 + k] = kEmptyMin;
        v[9 + k] = kEmptyMax;
    }
    if (mhi != mlo && mhi >= 0) {  // block-uniform: a second member only where one starts
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[6 + k] = wave_min_dpp(in1 ? lo[k] : kEmptyMin);
            v[9 + k] = wave_max_dpp(in1 ? hi[k] : kEmptyMax);
        }
    }
    v[12] = static_cast<int>(wave_max_u_dpp(accel));
    const unsigned cw = wave_sum_u_dpp(contacts);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 13; ++k) s_box[wave][k] = v[k];
        s_box[wave][13] = static_cast<int>(cw);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned am = static_cast<unsigned>(s_box[0][12]);
        unsigned ct = static_cast<unsigned>(s_box[0][13]);
        for (int w = 1; w < nw; ++w) {
            ct += static_cast<unsigned>(s_box[w][13]);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = min(v[k], s_box[w][k]);
                v[3 + k] = max(v[3 + k], s_box[w][3 + k]);
                v[6 + k] = min(v[6 + k], s_box[w][6 + k]);
                v[9 + k] = max(v[9 + k], s_box[w][9 + k]);
            }
            am = max(am, static_cast<unsigned>(s_box[w][12]));
        }
        BlockPart bp;
#pragma unroll
        for (int k = 0; k < 6; ++k) bp.box0[k] = v[k];
        bp.accel = am;
        bp.contacts = ct;
        parts[slot] = bp;
        if (pm) {
            BlockMembers bm;
            bm.m0 = mhi < 0 ? -1 : mlo;
            bm.m1 = mhi;
#pragma unroll
            for (int k = 0; k < 6; ++k) bm.box1[k] = v[6 + k];
            pm[slot] = bm;
        }
    }
}

// store_block_part for a block of one wave: the same partial, reduced with
// wave operations only (no LDS, so the pass's LDS bounds its occupancy less).
__device__ void store_wave_part(BlockPart* parts, BlockMembers* pm, int slot, int member, const int lo[3],
                                const int hi[3], unsigned accel, unsigned contacts) {
    const int lane = threadIdx.x & 63;
    // A lone simulation (no member partials) has member 0 or none.
    const int mlo = pm ? wave_min_dpp(member < 0 ? kEmptyMin : member) : 0;
    const int mhi = pm ? wave_max_dpp(member) : 0;
    const bool in0 = member >= 0 && member == mlo;
    const bool in1 = member >= 0 && member == mhi && mhi != mlo;
    int v[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = wave_min_dpp(in0 ? lo[k] : kEmptyMin);
        v[3 + k] = wave_max_dpp(in0 ? hi[k] : kEmptyMax);
        v[6 + k] = kEmptyMin;
        v[9 + k] = kEmptyMax;
    }
    if (mhi != mlo && mhi >= 0) {  // wave-uniform: a second member only where one starts
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[6 + k] = wave_min_dpp(in1 ? lo[k] : kEmptyMin);
            v[9 + k] = wave_max_dpp(in1 ? hi[k] : kEmptyMax);
        }
    }
    const unsigned am = wave_max_u_dpp(accel);
    const unsigned ct = wave_sum_u_dpp(contacts);
    if (lane == 0) {
        BlockPart bp;
#pragma unroll
        for (int k = 0; k < 6; ++k) bp.box0[k] = v[k];
        bp.accel = am;
        bp.contacts = ct;
        parts[slot] = bp;
        if (pm) {
            BlockMembers bm;
            bm.m0 = mhi < 0 ? -1 : mlo;
            bm.m1 = mhi;
#pragma unroll
            for (int k = 0; k < 6; ++k) bm.box1[k] = v[6 + k];
            pm[slot] = bm;
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Ingest an original-order AoS state: grain i gets id i.
__global__ __launch_bounds__(kBlock) void k_ingest(const Params* __restrict__ Pp,
                                                   BlockPart* __restrict__ parts,
                                                   const float* __restrict__ st6,
                                                   const int* __restrict__ mat,
                                                   float4* __restrict__ Tp, float4* __restrict__ Tv,
                                                   uint32_t* __restrict__ key,
                                                   int* __restrict__ hist_cnt0,
                                                   int* __restrict__ hist_cnt1, int n) {
    const Params& P = *Pp;
    const int i = logical_block() * blockDim.x + threadIdx.x;
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    int member = -1;
    if (i < n) {
        member = mat[i] >> 4;
        const float* s = st6 + 6 * static_cast<size_t>(i);
        const float x = s[0], y = s[1], z = s[2];
        Tp[i] = make_float4(x, y, z, __int_as_float(i));
        Tv[i] = make_float4(s[3], s[4], s[5], __int_as_float(mat[i]));  // member << 4 | material
        if (hist_cnt0) {  // an ingested state starts with no contact history
            hist_cnt0[i] = 0;
            hist_cnt1[i] = 0;
        }
        int cx, cy, cz;
        grain_cell(P, x, y, z, cx, cy, cz);
        key[i] = pack_cell(cx, cy, cz);
        lo[0] = hi[0] = cx;
        lo[1] = hi[1] = cy;
        lo[2] = hi[2] = cz;
    }
    store_block_part(parts, P.members > 1 ? P.pmem : nullptr, logical_block(), member, lo, hi, 0u);
}

// Original-order material ids pushed into the working layout.
__global__ __launch_bounds__(kBlock) void k_set_material(const float4* __restrict__ Tp,
                                                         float4* __restrict__ Tv,
                                                         const int* __restrict__ mat, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Tv[i].w = __int_as_float(mat[__float_as_int(Tp[i].w)]);
}

// Working layout -> original-order AoS state.
__global__ __launch_bounds__(kBlock) void k_export(const float4* __restrict__ Tp,
                                                   const float4* __restrict__ Tv,
                                                   float* __restrict__ st6, int* __restrict__ mat,
                                                   int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 p = Tp[i], v = Tv[i];
        const int id = __float_as_int(p.w);
        float* s = st6 + 6 * static_cast<size_t>(id);
        s[0] = p.x; s[1] = p.y; s[2] = p.z;
        s[3] = v.x; s[4] = v.y; s[5] = v.z;
        mat[id] = __float_as_int(v.w);
    }
}

// Cell keys and block boxes of the current state under the current grid
// (after the collision grid changes; the next k_finalize publishes the boxes).
__global__ __launch_bounds__(kBlock) void k_rekey(const Params* __restrict__ Pp,
                                                  BlockPart* __restrict__ parts,
                                                  const float4* __restrict__ Tp,
                                                  const float4* __restrict__ Tv,
                                                  uint32_t* __restrict__ key, int n) {
    const Params& P = *Pp;
    const int i = logical_block() * blockDim.x + threadIdx.x;
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    int member = -1;
    if (i < n) {
        const float4 p = Tp[i];
        member = member_of(P, Tv[i].w);
        int cx, cy, cz;
        grain_cell(P, p.x, p.y, p.z, cx, cy, cz);
        key[i] = pack_cell(cx, cy, cz);
        lo[0] = hi[0] = cx;
        lo[1] = hi[1] = cy;
        lo[2] = hi[2] = cz;
    }
    store_block_part(parts, P.members > 1 ? P.pmem : nullptr, logical_block(), member, lo, hi, 0u);
}

// Cells of the padded sub-grid over occupied box bx.
__device__ __forceinline__ uint32_t grid_cells(const int bx[6]) {
    return static_cast<uint32_t>((bx[3] - bx[0] + 3) * (bx[4] - bx[1] + 3) * (bx[5] - bx[2] + 3));
}

// The pending grid of a lone simulation: occupied box bx widened by one cell
// on every side (within the collision grid), so that the grains of the next
// state can be counted into it before its box is known, unless one moves
// more than a cell.
__device__ void store_pending(const Params& P, const int bx[6]) {
    const int g[3] = {P.gx, P.gy, P.gz};
    int w[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] = max(0, bx[k] - 1);
        w[3 + k] = min(g[k] - 1, bx[3 + k] + 1);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) P.pbox[k] = w[k];
    P.pcell[0] = 0u;
    P.pcell[1] = grid_cells(w);
}

// Slice r of a lone simulation's block partials (and, with with_e, of its ng
// energy groups) reduced by one wave, every lane active: lane l folds the
// slice's entries l, l + 64, ... in order (loads issued four at a time, the
// index clamped into the slice so that none is predicated), then the fixed
// DPP trees.  Lane 0 stores the slice.
__device__ void slice_reduce(const BlockPart* __restrict__ parts, int nparts, const double* __restrict__ ewave,
                             int ng, bool with_e, int r, FinSlice* out) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int p0 = static_cast<int>(static_cast<long long>(nparts) * r / kFinSlices);
    const int p1 = static_cast<int>(static_cast<long long>(nparts) * (r + 1) / kFinSlices);
    int bx[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
    unsigned am = 0u;
    unsigned long long cs = 0ull;
    for (int b = p0 + lane; b < p1; b += 4 * 64) {
        BlockPart q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = parts[min(b + 64 * k, p1 - 1)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b + 64 * k < p1) {
                am = max(am, q[k].accel);
                cs += q[k].contacts;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    bx[a] = min(bx[a], q[k].box0[a]);
                    bx[3 + a] = max(bx[3 + a], q[k].box0[3 + a]);
                }
            }
        }
    }
    double e = 0.0;
    if (with_e) {
        const int g0 = static_cast<int>(static_cast<long long>(ng) * r / kFinSlices);
        const int g1 = static_cast<int>(static_cast<long long>(ng) * (r + 1) / kFinSlices);
        for (int g = g0 + lane; g < g1; g += 4 * 64) {
            double x[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) x[k] = ewave[min(g + 64 * k, g1 - 1)];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (g + 64 * k < g1) e += x[k];
            }
        }
    }
    e = wave_sum(e);
    const double cw = wave_sum(static_cast<double>(cs));  // exact below 2^53
    am = wave_max_u_dpp(am);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bx[a] = wave_min_dpp(bx[a]);
        bx[3 + a] = wave_max_dpp(bx[3 + a]);
    }
    if (lane == 0) {
        FinSlice o;
#pragma unroll
        for (int a = 0; a < 6; ++a) o.box[a] = bx[a];
        o.accel = am;
        o.pad = 0u;
        o.contacts = cw;
        o.energy = e;
        *out = o;
    }
}

// The kFinSlices slices combined by one wave (lane r: slice r, every lane
// active), published by lane 0 as k_finalize's flags ask for a lone
// simulation; box_out (null or shared memory): lane 0 leaves the box there.
__device__ void slice_combine(const Params& P, Ctl* ctl, const FinSlice* sl, double* __restrict__ energy_log,
                              int flags, int* box_out) {
    static_assert(kFinSlices == 64, "one slice per lane");
    const int lane = static_cast<int>(threadIdx.x & 63);
    const FinSlice s = sl[lane];
    int bx[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bx[a] = wave_min_dpp(s.box[a]);
        bx[3 + a] = wave_max_dpp(s.box[3 + a]);
    }
    const unsigned am = wave_max_u_dpp(s.accel);
    const double contacts = wave_sum(s.contacts);
    const double total = wave_sum(s.energy);
    if (lane != 0) return;
    if (flags & kFinBox) {
        if (!(flags & kFinPending)) {
#pragma unroll
            for (int k = 0; k < 6; ++k) P.mbox[k] = bx[k];
            P.mcell[0] = 0u;
            P.mcell[1] = grid_cells(bx);
        }
        store_pending(P, bx);
    }
    if (box_out) {
#pragma unroll
        for (int k = 0; k < 6; ++k) box_out[k] = bx[k];
    }
    if (flags & kFinAccel) {
        ctl->accel_last = am;
        ctl->force_passes = ctl->force_passes + 1;
        ctl->contact_sum = ctl->contact_sum + static_cast<unsigned long long>(contacts);
    }
    if (flags & kFinEnergyLast) ctl->energy_last = total;
    if (flags & kFinFlipHistory) ctl->hist_parity = ctl->hist_parity ^ 1;
    if ((flags & kFinEnergyLog) && ctl->energy_count < ctl->energy_capacity) {
        energy_log[ctl->energy_count] = total;
        ctl->energy_count = ctl->energy_count + 1;
    }
}

// The reduction of a lone simulation's block partials on one block: its waves
// take the slices in turn, then wave 0 combines them.
__device__ void slices_on_block(const Params& P, Ctl* ctl, const BlockPart* __restrict__ parts, int nparts,
                                const double* __restrict__ ewave, double* __restrict__ energy_log, int flags,
                                int* box_out) {
    __shared__ FinSlice s_sl[kFinSlices];
    const int wave = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
    const bool with_e = (flags & (kFinEnergyLog | kFinEnergyLast)) != 0;
    const int ng = with_e ? (P.n + kEnergyGroup - 1) / kEnergyGroup : 0;
    for (int r = wave; r < kFinSlices; r += nw) slice_reduce(parts, nparts, ewave, ng, with_e, r, &s_sl[r]);
    __syncthreads();
    if (wave == 0) slice_combine(P, ctl, s_sl, energy_log, flags, box_out);
}

// ---------------------------------------------------------------------------
// The slot of a grain (lane) in cell c: count[c] is its cell's counter, every
// lane of the wave calls this (valid: the lane carries a grain, c is then its
// cell; grain i, original id `id`).  Grains arrive in the last step's cell
// order, so neighbouring lanes often share a cell: a run of equal cells in the
// wave takes its slots with one atomic (the order of slots inside a cell is
// free: the reorder ranks a cell's grains by id).  BUCKET: the id goes into
// the cell's bucket (or, past kBucket, spill list ctl->bin_par), otherwise
// the slot into slot[i].
template <bool BUCKET>
__device__ __forceinline__ void bin_grain(Ctl* ctl, uint32_t* __restrict__ count, uint32_t* __restrict__ slot,
                                          uint32_t* __restrict__ bucket, uint2* __restrict__ spill, bool valid,
                                          uint32_t c, int id, int i, int n) {
    const int lane = threadIdx.x & 63;
    if (!valid) c = 0xFFFFFFFFu;  // no cell
    // The previous lane's cell (wave_shr:1; lane 0 gets none).
    const uint32_t cprev = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(-2, static_cast<int>(c), 0x138, 0xF, 0xF, false));
    const unsigned long long heads = __ballot(!valid || c != cprev || lane == 0);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h = 63 - __clzll(static_cast<long long>(heads & upto));  // this run's first lane
    const unsigned long long after = heads & ~upto;
    const int e = after ? __ffsll(static_cast<long long>(after)) - 1 : 64;  // the next run's
    uint32_t base = 0u;
    if (valid && lane == h) base = atomicAdd(&count[c], static_cast<uint32_t>(e - h));
    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), h, 64));
    if (valid) {
        const uint32_t s = base + static_cast<uint32_t>(lane - h);
        if constexpr (BUCKET) {
            if (s < static_cast<uint32_t>(kBucket)) {
                bucket[static_cast<size_t>(c) * kBucket + s] = static_cast<uint32_t>(id);
            } else {
                const unsigned bp = ctl->bin_par;  // only k_scan writes it
                const unsigned e = atomicAdd(&ctl->spill_n[bp], 1u);  // below n: one entry per grain
                spill[static_cast<size_t>(bp) * n + e] = make_uint2(c, static_cast<uint32_t>(id));
            }
        } else {
            slot[i] = s;
        }
    }
}

// Fused step: the count of grain p (new cell nx, ny, nz; `has`: the lane
// carries one) into the pending grid's box pb, from the force pass that
// produced it; every lane of the wave calls both halves, in turn.
// fused_count_issue takes the slots (one atomic per run of equal cells, as
// bin_grain); fused_count_finish files the grain (bucket, spill list or
// slot).  A grain outside the pending grid is not counted; it sets bin_miss,
// and the next scan bins every grain itself.  (Issuing the atomic earlier, to
// overlap the pass's last reductions, measured no faster and held registers
// across them that cost the one-lane pass a wave per SIMD.)
struct FusedCount {
    uint32_t c;      // the grain's cell, 0xFFFFFFFF: none
    int h;           // its run's first lane
    uint32_t base;   // that lane's atomic result
    bool valid;
};
__device__ __forceinline__ FusedCount fused_count_issue(Ctl* ctl, const BinIO& bio, const int* pb, bool has,
                                                        int nx, int ny, int nz) {
    FusedCount fc{0xFFFFFFFFu, 0, 0u, false};
    if (bio.mode == 0) return fc;
    const bool inside = has && nx >= pb[0] && nx <= pb[3] && ny >= pb[1] && ny <= pb[4] && nz >= pb[2] &&
                        nz <= pb[5];
    if (has && !inside) ctl->bin_miss = 1;
    const int lane = static_cast<int>(threadIdx.x & 63);
    fc.valid = inside;
    if (inside) {
        SubGrid g;
        g.ox = pb[0] - 1;
        g.oy = pb[1] - 1;
        g.oz = pb[2] - 1;
        g.bx = pb[3] - pb[0] + 3;
        g.by = pb[4] - pb[1] + 3;
        g.bz = pb[5] - pb[2] + 3;
        fc.c = static_cast<uint32_t>(g.index(nx, ny, nz));
    }
    const uint32_t cprev = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(-2, static_cast<int>(fc.c), 0x138, 0xF, 0xF, false));
    const unsigned long long heads = __ballot(!inside || fc.c != cprev || lane == 0);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    fc.h = 63 - __clzll(static_cast<long long>(heads & upto));
    const unsigned long long after = heads & ~upto;
    const int e = after ? __ffsll(static_cast<long long>(after)) - 1 : 64;
    if (inside && lane == fc.h) fc.base = atomicAdd(&bio.count[fc.c], static_cast<uint32_t>(e - fc.h));
    return fc;
}
__device__ __forceinline__ void fused_count_finish(Ctl* ctl, const BinIO& bio, const FusedCount& fc, int id, int p,
                                                   int n) {
    if (bio.mode == 0) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint32_t base = static_cast<uint32_t>(__shfl(static_cast<int>(fc.base), fc.h, 64));
    if (!fc.valid) return;
    const uint32_t s = base + static_cast<uint32_t>(lane - fc.h);
    if (bio.mode == 2) {
        if (s < static_cast<uint32_t>(kBucket)) {
            bio.bucket[static_cast<size_t>(fc.c) * kBucket + s] = static_cast<uint32_t>(id);
        } else {
            const unsigned bp = ctl->bin_par;  // only k_scan writes it
            const unsigned e = atomicAdd(&ctl->spill_n[bp], 1u);
            bio.spill[static_cast<size_t>(bp) * n + e] = make_uint2(fc.c, static_cast<uint32_t>(id));
        }
    } else {
        bio.slot[p] = s;
    }
}

// Fused step, block 0 of the force pass: re-arms the scan (as k_count) and
// hands the pending grid to the scan and reorder that bin this pass.
__device__ __forceinline__ void fused_arm(const Params& P, Ctl* ctl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->tile_counter = 0;
        ctl->epoch = ctl->epoch + 1;
        const unsigned long long o = ctl->ovf_next;  // this pass uses the bin they were found in
        if (o) {
            ctl->overflow_lifetime = ctl->overflow_lifetime + o;
            P.movf[0] = P.movf[0] + o;
            ctl->ovf_next = 0ull;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) P.nbox[k] = P.pbox[k];
        P.ncell[0] = P.pcell[0];
        P.ncell[1] = P.pcell[1];
    }
}

// ---------------------------------------------------------------------------
// Pass 1: slot of each grain in its cell.  Block 0 also re-arms the scan.
// box_out / cell_out (null unless a caller keeps the bin of this pass): a
// copy of the member boxes and cell offsets the pass bins with.  BUCKET:
// instead of the slot, the grain's id goes into its cell's bucket (slots
// below kBucket) or, past it, into spill list ctl->bin_par as (cell, id); the
// reorder then ranks a cell's ids without a scatter pass.
template <bool BUCKET>
__global__ __launch_bounds__(kBlock) void k_count(const Params* __restrict__ Pp, Ctl* ctl,
                                                  const float4* __restrict__ Tp,
                                                  const float4* __restrict__ Tv,
                                                  const uint32_t* __restrict__ key,
                                                  uint32_t* __restrict__ slot,
                                                  uint32_t* __restrict__ count, int n,
                                                  int* __restrict__ box_out,
                                                  uint32_t* __restrict__ cell_out,
                                                  uint32_t* __restrict__ bucket,
                                                  uint2* __restrict__ spill, int clear_next) {
    const Params& P = *Pp;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->tile_counter = 0;
        ctl->epoch = ctl->epoch + 1;
        if (clear_next) ctl->ovf_next = 0ull;  // a fused bin this one replaces
    }
    if (blockIdx.x == 0 && box_out) {
        for (int k = threadIdx.x; k < 6 * P.members; k += blockDim.x) box_out[k] = P.mbox[k];
        for (int k = threadIdx.x; k <= P.members; k += blockDim.x) cell_out[k] = P.mcell[k];
    }
    const int i = logical_block() * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t c = 0xFFFFFFFFu;  // no cell
    int id = 0;
    if (valid) {
        if (BUCKET) id = __float_as_int(Tp[i].w);
        const MemberCells mc = member_cells(P, P.members == 1 ? 0 : grain_member(Tv[i].w));
        int cx, cy, cz;
        unpack_cell(key[i], cx, cy, cz);
        c = mc.index(cx, cy, cz);
    }
    bin_grain<BUCKET>(ctl, count, slot, bucket, spill, valid, c, id, i, n);
}

// Pass 2: start[c] = sum of count[0..c) over the concatenated member
// sub-grids, written up to and including index `cells`; counts are zeroed
// for the next pass and over-capacity cells are added to the lifetime
// overflow total.  A tile only waits on lower tiles: block b's first tile is
// b, and a lower-indexed block is dispatched first on its XCD; a tile past
// the grid is claimed by a running block, which reaches it once its current
// tile (waiting only on lower ones) is done.  So every wait ends, without
// co-residency limits.
// Fused step, after the force pass left a grain outside the pending grid:
// the scan's reduction block bins the pass's grains on its own, as k_count
// and k_scan would with the grid of their exact box bx: the counts the force
// pass made (over the old next grid's old_cells + 1 entries) are cleared,
// the grains counted again (spill list ctl->bin_par, then flipped, as
// k_scan does), and the cells scanned 16 per thread.  The next grid becomes
// the exact one; the reorder then proceeds as usual.
__device__ void fallback_bin(const Params& P, Ctl* ctl, const ScanFuse& fz, uint32_t* __restrict__ count,
                             uint32_t* __restrict__ start) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    __shared__ int bx[6];
    const int tid = static_cast<int>(threadIdx.x), nt = static_cast<int>(blockDim.x);
    const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    const uint32_t old_cells = P.ncell[1];  // the grid the force pass counted into
    // the exact box of the pass's grains (its partials; nothing published)
    slices_on_block(P, ctl, fz.parts, fz.nparts, nullptr, nullptr, 0, bx);
    __syncthreads();
    const uint32_t cells = grid_cells(bx);
    for (uint32_t k = static_cast<uint32_t>(tid); k <= old_cells; k += static_cast<uint32_t>(nt)) count[k] = 0u;
    const unsigned bp = ctl->bin_par;
    __syncthreads();  // counts cleared, bin_par read by every thread
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) P.nbox[k] = bx[k];
        P.ncell[0] = 0u;
        P.ncell[1] = cells;
        ctl->spill_n[0] = 0u;
        ctl->spill_n[1] = 0u;
    }
    __syncthreads();
    const SubGrid g = make_subgrid(bx);
    for (int i = tid; i < fz.n; i += nt) {
        int cx, cy, cz;
        unpack_cell(fz.key[i], cx, cy, cz);
        const uint32_t c = static_cast<uint32_t>(g.index(cx, cy, cz));
        const uint32_t sl = atomicAdd(&count[c], 1u);
        if (fz.bio.mode == 2) {
            const uint32_t id = __float_as_uint(fz.Tp[i].w);
            if (sl < static_cast<uint32_t>(kBucket)) {
                fz.bio.bucket[static_cast<size_t>(c) * kBucket + sl] = id;
            } else {
                const unsigned e = atomicAdd(&ctl->spill_n[bp], 1u);
                fz.bio.spill[static_cast<size_t>(bp) * fz.n + e] = make_uint2(c, id);
            }
        } else {
            fz.bio.slot[i] = sl;
        }
    }
    __syncthreads();  // every count is in (device atomics, read back below at device scope)
    if (tid == 0) ctl->bin_par = bp ^ 1u;
    const uint32_t kcap = static_cast<uint32_t>(P.kcap);
    const uint32_t total = cells + 1u;
    uint32_t carry = 0u, over = 0u;
    for (uint32_t base = 0; base < total; base += static_cast<uint32_t>(nt) * 16u) {
        const uint32_t i0 = base + static_cast<uint32_t>(tid) * 16u;
        uint32_t v[16], sum = 0u;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t idx = i0 + static_cast<uint32_t>(k);
            const uint32_t c = idx < total ? __hip_atomic_load(&count[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                           : 0u;
            over += c > kcap ? 1u : 0u;
            v[k] = sum;
            sum += c;
        }
        const uint32_t incl = static_cast<uint32_t>(wave_incl_scan_dpp(static_cast<int>(sum)));
        if (lane == 63) s_w[wave] = incl;
        __syncthreads();
        uint32_t woff = 0u, btot = 0u;
        for (int w = 0; w < nw; ++w) {
            if (w < wave) woff += s_w[w];
            btot += s_w[w];
        }
        const uint32_t off = carry + woff + (incl - sum);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t idx = i0 + static_cast<uint32_t>(k);
            if (idx < total) {
                start[idx] = off + v[k];
                const uint32_t c = (k < 15 ? v[k + 1] : sum) - v[k];
                if (c) count[idx] = 0u;
            }
        }
        carry += btot;
        __syncthreads();  // s_w is rewritten by the next round
    }
    over = wave_sum_u_dpp(over);
    if (lane == 0 && over) atomicAdd(&ctl->ovf_next, static_cast<unsigned long long>(over));
    if (tid == 0) ctl->bin_misses = ctl->bin_misses + 1ull;
}

// A reduction block of a fused step's scan (rb of kFinSliceBlocks): its waves
// reduce one slice each of the force pass's partials; the last one also
// takes the spill-list turn k_scan's block 0 makes otherwise or, after a
// miss, the whole bin.
__device__ void scan_reduction_block(const Params& P, Ctl* ctl, uint32_t* __restrict__ count,
                                     uint32_t* __restrict__ start, const ScanFuse& fz, int rb) {
    const int r = rb * (kScanThreads / 64) + static_cast<int>(threadIdx.x >> 6);
    slice_reduce(fz.parts, fz.nparts, fz.ewave, fz.ng, fz.with_e != 0, r, fz.slices + r);
    if (rb != kFinSliceBlocks - 1) return;
    if (!ctl->bin_miss) {
        if (threadIdx.x == 0) {
            const unsigned bp = ctl->bin_par;
            ctl->spill_n[bp ^ 1u] = 0u;
            ctl->bin_par = bp ^ 1u;
        }
        return;
    }
    fallback_bin(P, ctl, fz, count, start);
}

template <int ITEMS>
__global__ __launch_bounds__(kScanThreads) void k_scan(const Params* __restrict__ Pp, Ctl* ctl,
                                                       uint32_t* __restrict__ count,
                                                       uint32_t* __restrict__ start,
                                                       unsigned long long* __restrict__ status, ScanFuse fz) {
    static_assert(ITEMS % 4 == 0, "tiles are loaded as uint4");
    constexpr int kScanItems = ITEMS;
    constexpr int kScanTile = kScanThreads * ITEMS;
    const bool fused = fz.on != 0;
    const int nb = static_cast<int>(gridDim.x) - (fused ? kFinSliceBlocks : 0);  // blocks that take tiles
    if (fused && static_cast<int>(blockIdx.x) >= nb) {
        scan_reduction_block(*Pp, ctl, count, start, fz, static_cast<int>(blockIdx.x) - nb);
        return;
    }
    if (fused && ctl->bin_miss) return;  // the reduction block bins alone
    const int cells = fused ? static_cast<int>(Pp->ncell[1]) : static_cast<int>(Pp->mcell[Pp->members]);
    const int ntiles = (cells + 1 + kScanTile - 1) / kScanTile;
    const unsigned long long tag = static_cast<unsigned long long>(ctl->epoch & 0x3FFFFFFFu) << 34;
    const uint32_t kcap = static_cast<uint32_t>(Pp->kcap);
    __shared__ int s_tile;
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_over[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (!fused && blockIdx.x == 0 && threadIdx.x == 0) {
        // bucket binning: the next count pass spills into the other list,
        // emptied here (its last reader, the previous reorder, is done)
        const unsigned bp = ctl->bin_par;
        ctl->spill_n[bp ^ 1u] = 0u;
        ctl->bin_par = bp ^ 1u;
    }
    if (static_cast<int>(blockIdx.x) >= ntiles) return;  // never takes a tile
    // A block scans tile blockIdx.x first, then (when the grid has fewer
    // blocks than tiles) tiles handed out by the counter past the grid, each
    // claimed once the block is free.
    const bool more = ntiles > nb;
    for (int tile = static_cast<int>(blockIdx.x); tile < ntiles;) {
        const size_t base = static_cast<size_t>(tile) * kScanTile + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        const uint4* src = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = src[k];
            v[4 * k + 0] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
        }
        uint32_t sum = 0, over = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (v[k] > kcap) {
                ++over;
                // The member owning the cell (rare: over-capacity cells only);
                // a fused step's bin counts when the next force pass uses it.
                if (!fused) {
                    const uint32_t cell = static_cast<uint32_t>(base) + k;
                    int lo = 0, hi = Pp->members - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (Pp->mcell[mid] <= cell) lo = mid;
                        else hi = mid - 1;
                    }
                    atomicAdd(&Pp->movf[lo], 1ull);
                }
            }
            const uint32_t c = v[k];
            v[k] = sum;  // exclusive within the thread
            sum += c;
        }
        // Wave-level scan of the thread sums.
        const uint32_t incl = static_cast<uint32_t>(wave_incl_scan_dpp(static_cast<int>(sum)));
        const uint32_t wave_total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
        over = wave_sum_u_dpp(over);
        if (lane == 0) {
            s_wave[wave] = wave_total;
            s_over[wave] = over;
        }
        __syncthreads();
        uint32_t wave_off = 0, agg = 0, over_total = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) {
            if (w < wave) wave_off += s_wave[w];
            agg += s_wave[w];
            over_total += s_over[w];
        }
        if (threadIdx.x == 0 && over_total) {
            atomicAdd(fused ? &ctl->ovf_next : &ctl->overflow_lifetime, static_cast<unsigned long long>(over_total));
        }
        // Look-back by the first wave: lane k reads the status of tile
        // j - k, so one round covers 64 predecessors; the nearest one with
        // its inclusive prefix ends the walk, and the aggregates between are
        // added.  A round repeats until every status it needs is published.
        if (wave == 0) {
            uint32_t prefix = 0;
            if (tile == 0) {
                if (lane == 0) {
                    __hip_atomic_store(&status[0], tag | (2ull << 32) | agg, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                if (lane == 0) {
                    __hip_atomic_store(&status[tile], tag | (1ull << 32) | agg, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
                int j = tile - 1;
                unsigned spins = 0;
                for (;;) {
                    const int idx = j - lane;
                    unsigned long long w = 0ull;
                    if (idx >= 0) {
                        w = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    const unsigned flag = static_cast<unsigned>((w >> 32) & 3u);
                    const bool ok = idx >= 0 && (w & ~((1ull << 34) - 1)) == tag && flag != 0u;
                    const unsigned long long inc = __ballot(ok && flag == 2u);
                    const int stop = inc ? __builtin_ctzll(inc) : 64;  // nearest inclusive
                    const unsigned long long need = stop >= 63 ? ~0ull : ((2ull << stop) - 1ull);
                    if ((__ballot(ok) & need) != need) {
                        if (++spins > (1u << 22)) {
                            if (lane == 0) ctl->error = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    prefix += wave_sum_u_dpp(lane <= stop && ok ? static_cast<uint32_t>(w) : 0u);
                    if (stop < 64) break;
                    j -= 64;
                }
                if (lane == 0) {
                    __hip_atomic_store(&status[tile], tag | (2ull << 32) | (prefix + agg),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            if (lane == 0) s_prefix = prefix;
        }
        __syncthreads();
        const uint32_t off = s_prefix + wave_off + (incl - sum);
        uint4* dst = reinterpret_cast<uint4*>(start + base);
        uint4* cz = reinterpret_cast<uint4*>(count + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            dst[k] = make_uint4(off + v[4 * k + 0], off + v[4 * k + 1], off + v[4 * k + 2],
                                off + v[4 * k + 3]);
            const uint4 q = src[k];
            if (q.x | q.y | q.z | q.w) cz[k] = make_uint4(0u, 0u, 0u, 0u);
        }
        if (!more) break;
        if (threadIdx.x == 0) s_tile = nb + static_cast<int>(atomicAdd(&ctl->tile_counter, 1u));
        __syncthreads();
        tile = s_tile;
        __syncthreads();  // every thread has read the claim before the next one is written
    }
}

// Pass 3: original ids into cell-sorted slots (order inside a cell is still
// arbitrary here).
__global__ __launch_bounds__(kBlock) void k_scatter(const Params* __restrict__ Pp,
                                                    const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ slot,
                                                    const float4* __restrict__ Tp,
                                                    const float4* __restrict__ Tv,
                                                    const uint32_t* __restrict__ start,
                                                    int* __restrict__ tmp_id, int n, int fused) {
    const Params& P = *Pp;
    const int i = logical_block() * blockDim.x + threadIdx.x;
    if (i < n) {
        const MemberCells mc = fused ? MemberCells{make_subgrid(P.nbox), 0u}
                                     : member_cells(P, P.members == 1 ? 0 : grain_member(Tv[i].w));
        int cx, cy, cz;
        unpack_cell(key[i], cx, cy, cz);
        tmp_id[start[mc.index(cx, cy, cz)] + slot[i]] = __float_as_int(Tp[i].w);
    }
}

// Pass 4: final position = cell start + rank of the original id among the
// cell's ids (from the scattered ids, or with BUCKET from the cell's bucket
// and, for a cell of more than kBucket grains, the spill list the count pass
// filled).  NMODE 1 also copies the grain into the Picard base buffers,
// NMODE 2 carries the Picard base along the same permutation.
// rf.on (a fused step): the cells are the next grid's, which block 0 then
// makes the current one (no block of this pass reads the current grid); it
// also clears the force pass's miss flag and combines the scan's slices of
// the pass's partials.
template <int NMODE, bool BUCKET>
__global__ __launch_bounds__(kBlock) void k_reorder(
        const Params* __restrict__ Pp, Ctl* ctl, const uint32_t* __restrict__ key,
        const uint32_t* __restrict__ start, const int* __restrict__ tmp_id,
        const uint32_t* __restrict__ bucket, const uint2* __restrict__ spill,
        const float4* __restrict__ Tp, const float4* __restrict__ Tv, float4* __restrict__ Sp,
        float4* __restrict__ Sv, Carry carry, int n, ReorderFuse rf) {
    const Params& P = *Pp;
    const bool fused = rf.on != 0;
    if (fused && blockIdx.x == 0) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) P.mbox[k] = P.nbox[k];
            P.mcell[0] = P.ncell[0];
            P.mcell[1] = P.ncell[1];
            if (rf.box_out) {
#pragma unroll
                for (int k = 0; k < 6; ++k) rf.box_out[k] = P.nbox[k];
                rf.cell_out[0] = P.ncell[0];
                rf.cell_out[1] = P.ncell[1];
            }
            ctl->bin_miss = 0;
        }
        slice_combine(P, ctl, rf.slices, rf.energy_log, rf.flags, nullptr);
    }
    const int i = logical_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = Tp[i];
    const float4 v = Tv[i];
    const MemberCells mc = fused ? MemberCells{make_subgrid(P.nbox), 0u} : member_cells(P, member_of(P, v.w));
    int cx, cy, cz;
    unpack_cell(key[i], cx, cy, cz);
    const uint32_t c = mc.index(cx, cy, cz);
    const uint32_t b = start[c], e = start[c + 1];
    const int id = __float_as_int(p.w);
    uint32_t rank = 0;
    if (e - b > 1) {
        if constexpr (BUCKET) {
            const uint32_t* bk = bucket + static_cast<size_t>(c) * kBucket;
            const uint32_t m = min(e - b, static_cast<uint32_t>(kBucket));
            for (uint32_t q = 0; q < m; ++q) rank += (static_cast<int>(bk[q]) < id) ? 1u : 0u;
            if (e - b > static_cast<uint32_t>(kBucket)) {
                const unsigned bp = ctl->bin_par ^ 1u;  // the list the count pass filled
                const unsigned ns = ctl->spill_n[bp];
                const uint2* sp = spill + static_cast<size_t>(bp) * n;
                for (unsigned q = 0; q < ns; ++q) {
                    const uint2 v = sp[q];
                    rank += (v.x == c && static_cast<int>(v.y) < id) ? 1u : 0u;
                }
            }
        } else {
            for (uint32_t q = b; q < e; ++q) rank += (tmp_id[q] < id) ? 1u : 0u;
        }
    }
    const uint32_t dst = b + rank;
    Sp[dst] = p;
    Sv[dst] = v;
    if (carry.hidx_out) carry.hidx_out[dst] = NMODE == 2 ? carry.hidx_in[i] : i;
    if (NMODE == 1) {
        carry.nout_p[dst] = p;
        carry.nout_v[dst] = v;
    } else if (NMODE == 2) {
        carry.nout_p[dst] = carry.nin_p[i];
        carry.nout_v[dst] = carry.nin_v[i];
    }
}

// ---------------------------------------------------------------------------
// Pass 5: forces at the cell-sorted state S, fused with the integrator.
// MODE kModeSIE / kModeBEP write the next state to T (same slot), its cell
// key and the per-block box / max |a|; kModeRHS only writes
// rhs = [v, F/m] in original order.  ENERGY adds the per-grain energy of S
// (pair terms counted once, from the lower original index) to the block
// partials.  HIST selects the tangential spring with history: the stored
// springs of S are read from generation ctl->hist_parity (at hist.hidx[p])
// and the updated springs of this step are written to the other generation
// at p, in contact order (walls, then pairs in stencil order).
//
// Phase 1 walks the stencil with batched position loads and lists the
// touching candidates (squared distance below touch2, exactly the contact
// law's dist < 0), up to kContactList of them; phase 2 evaluates them in that
// order with the next contact's loads in flight.  A grain with more contacts
// walks again and lists the ones after those evaluated.  A candidate beyond
// contact adds exact zeros in the reference's arithmetic and is skipped; a
// NaN candidate (a diverged state) makes the force NaN whatever the order,
// which is applied at the end.
template <int MODE, bool ENERGY, bool HIST, int BS>
__global__ __launch_bounds__(BS, 5) void k_force(
        const Params* __restrict__ Pp, Ctl* ctl, const float4* __restrict__ Sp,
        const float4* __restrict__ Sv, const float4* __restrict__ Np,
        const float4* __restrict__ Nv, const uint32_t* __restrict__ start,
        float4* __restrict__ Tp, float4* __restrict__ Tv, uint32_t* __restrict__ key_out,
        float* __restrict__ rhs_out, BlockPart* __restrict__ parts, double* __restrict__ ewave,
        HistoryIO hist, int n, BinIO bio) {
    const Params& P = *Pp;
    if (MODE != kModeRHS && bio.mode) fused_arm(P, ctl);
    __shared__ uint32_t s_list[kContactList][BS];
    __shared__ float s_cpair[kMaxMat * kMaxMat];
    // A lone simulation's pair damping table, in LDS: one lookup per contact.
    const bool lds_tab = P.mtab == nullptr;
    if (lds_tab) {
        for (int k = threadIdx.x; k < kMaxMat * kMaxMat; k += BS) s_cpair[k] = P.c_pair[k];
        __syncthreads();
    }
    const int lb = logical_block();
    const int p = lb * blockDim.x + threadIdx.x;
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    unsigned accel_bits = 0;
    unsigned ncont = 0;
    float e_grain = 0.f;
    int lane_member = -1;
    int fc[3] = {0, 0, 0}, fid = 0;  // the grain's next cell and id (fused count)
    if (p < n) {
        const float4 pi = Sp[p];
        const float4 vi4 = Sv[p];
        const float x[3] = {pi.x, pi.y, pi.z};
        const float v[3] = {vi4.x, vi4.y, vi4.z};
        const int id = __float_as_int(pi.w);
        const int mat = grain_mat(vi4.w);
        const int member = member_of(P, vi4.w);
        const float mu = P.members == 1 ? P.mu : P.mmu[member];
        const float m = mat_mass(P, member, mat);
        auto cpair = [&](int mj) __attribute__((always_inline)) {
            return lds_tab ? s_cpair[mat * kMaxMat + mj] : mat_cpair(P, member, mat, mj);
        };
        float f[3] = {0.f, 0.f, 0.f};
        if (ENERGY) e_grain = grain_energy_self(P, x, v, m);

        // Contact history: the old list was written in the same contact order,
        // so its next entry is prefetched one contact ahead and normally
        // matches; a miss falls back to a search of the whole list.  Slot 0 is
        // loaded with the list's length (it exists whether or not in use).
        const int par = HIST ? ctl->hist_parity : 0;
        const float4* __restrict__ h_in = hist.buf[par];
        float4* __restrict__ h_out = hist.buf[par ^ 1];
        const int hs = HIST ? hist.hidx[p] : 0;  // this grain's index in generation par
        const int h_old = HIST ? hist.cnt[par][hs] : 0;
        int h_cur = 0, h_new = 0;
        float4 h_next = make_float4(0.f, 0.f, 0.f, 0.f);
        if (HIST) h_next = h_in[hs];
        auto h_take = [&](int partner, float xi[3]) __attribute__((always_inline)) {
            if (h_cur < h_old && __float_as_int(h_next.w) == partner) {
                xi[0] = h_next.x; xi[1] = h_next.y; xi[2] = h_next.z;
            } else {
                xi[0] = xi[1] = xi[2] = 0.f;
                const int found = spring_slot(h_in, n, hs, h_old, partner);
                if (found < 0) return;  // a new contact: the spring starts at zero
                const float4 h = h_in[static_cast<size_t>(found) * n + hs];
                xi[0] = h.x; xi[1] = h.y; xi[2] = h.z;
                h_cur = found;
            }
            ++h_cur;
            // the next slot, loaded unconditionally (slots below kHist exist)
            h_next = h_in[static_cast<size_t>(min(h_cur, kHist - 1)) * n + hs];
        };
        auto h_put = [&](int partner, const float xi[3]) __attribute__((always_inline)) {
            if (MODE != kModeRHS && h_new < kHist) {
                h_out[static_cast<size_t>(h_new) * n + p] =
                        make_float4(xi[0], xi[1], xi[2], __int_as_float(partner));
            }
            ++h_new;
        };

        // Gravity and the six walls, lower then upper per axis
        // (src/compute_rhs_kernel.cu:105-141).
        const float c_wall = mat_cwall(P, member, mat);
        f[2] -= m * P.gravity;
        #pragma unroll
        for (int a = 0; a < 3; ++a) {
            #pragma unroll
            for (int side = 0; side < 2; ++side) {
                const float dist = side == 0 ? x[a] - P.wlo[a] - P.radius : P.whi[a] - x[a] - P.radius;
                const float fw = fmaxf(0.f, fminf(P.max_force, P.kneg * dist));
                if (side == 0) f[a] += fw;
                else f[a] -= fw;
                if (dist < 0.f) {
                    ++ncont;
                    f[a] -= c_wall * v[a];
                    if (HIST) {
                        float nrm[3] = {0.f, 0.f, 0.f};
                        nrm[a] = side == 0 ? 1.f : -1.f;
                        float vn = 0.f;
                        vn += v[0] * nrm[0];
                        vn += v[1] * nrm[1];
                        vn += v[2] * nrm[2];
                        float xi[3];
                        h_take(wall_partner(a, side), xi);
                        if (ENERGY) {
                            float s2 = 0.f;
                            s2 += xi[0] * xi[0];
                            s2 += xi[1] * xi[1];
                            s2 += xi[2] * xi[2];
                            e_grain += 0.5f * P.kt * s2;
                        }
                        spring_friction(P, mu, f, nrm, v, vn, fw, c_wall, xi);
                        h_put(wall_partner(a, side), xi);
                    } else {
                        float vt[3];
                        #pragma unroll
                        for (int b = 0; b < 3; ++b) vt[b] = (b == a) ? 0.f : v[b];
                        viscous_friction(f, vt, mu * fw, c_wall);
                    }
                }
            }
        }

        const MemberCells mc = member_cells(P, member);
        int cx, cy, cz;
        grain_cell(P, x[0], x[1], x[2], cx, cy, cz);
        const uint32_t kcap = static_cast<uint32_t>(P.kcap);
        // Cell starts of stencil column c (cells cz-1 .. cz+1 and the next
        // cell's start), one 16-byte load.
        auto column = [&](int c) __attribute__((always_inline)) {
            uint4 q4;
            __builtin_memcpy(&q4, start + mc.index(cx + c / 3 - 1, cy + c % 3 - 1, cz - 1), sizeof(uint4));
            return q4;
        };
        // Phase 2: the full contact law for the listed contacts, in the
        // order phase 1 found them (the canonical stencil order).
        int listed = 0;
        auto flush = [&]() __attribute__((always_inline)) {
            uint32_t qn = 0;
            float4 pjn = make_float4(0.f, 0.f, 0.f, 0.f), vjn = pjn, pjn2 = pjn, vjn2 = pjn;
            if (listed > 0) {
                qn = s_list[0][threadIdx.x];
                pjn = Sp[qn];
                vjn = Sv[qn];
                qn = s_list[min(1, listed - 1)][threadIdx.x];
                pjn2 = Sp[qn];
                vjn2 = Sv[qn];
            }
            for (int k = 0; k < listed; ++k) {
                const float4 pj = pjn, vj = vjn;
                pjn = pjn2;
                vjn = vjn2;
                // the loads of the contact after the next (the last one's
                // again past the end: unconditional, so that two contacts'
                // loads are in flight during this one)
                qn = s_list[min(k + 2, listed - 1)][threadIdx.x];
                pjn2 = Sp[qn];
                vjn2 = Sv[qn];
                float d0, d1, d2;
                const float norm = pair_norm(x, pj.x, pj.y, pj.z, d0, d1, d2);
                const int jd = __float_as_int(pj.w);
                if (ENERGY && jd > id) e_grain += contact_energy(P, 2.f * P.radius - norm);
                const float dist = norm - P.radius - P.radius;
                const float nrm[3] = {d0 / norm, d1 / norm, d2 / norm};
                const float fp = fmaxf(0.f, fminf(P.max_force, P.kneg * dist));
                f[0] += fp * nrm[0];
                f[1] += fp * nrm[1];
                f[2] += fp * nrm[2];
                if (dist < 0.f) {  // every listed candidate
                    ++ncont;
                    const float c = cpair(grain_mat(vj.w));
                    const float vr[3] = {v[0] - vj.x, v[1] - vj.y, v[2] - vj.z};
                    float vn = 0.f;
                    vn += vr[0] * nrm[0];
                    vn += vr[1] * nrm[1];
                    vn += vr[2] * nrm[2];
#pragma unroll
                    for (int a = 0; a < 3; ++a) f[a] -= c * vn * nrm[a];
                    if (HIST) {
                        float xi[3];
                        h_take(jd, xi);
                        if (ENERGY && jd > id) {
                            float s2 = 0.f;
                            s2 += xi[0] * xi[0];
                            s2 += xi[1] * xi[1];
                            s2 += xi[2] * xi[2];
                            e_grain += 0.5f * P.kt * s2;
                        }
                        spring_friction(P, mu, f, nrm, vr, vn, fp, c, xi);
                        h_put(jd, xi);
                    } else {
                        float vt[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a) vt[a] = vr[a] - vn * nrm[a];
                        viscous_friction(f, vt, mu * fp, c);
                    }
                }
            }
            listed = 0;
        };
        // Phase 1: squared distances only.
        int consumed = 0, seen = 0;
        bool nan_f = false, nan_e = false;
        auto probe = [&](uint32_t q, float px, float py, float pz) __attribute__((always_inline)) {
            const float d0 = x[0] - px, d1 = x[1] - py, d2 = x[2] - pz;
            float s2 = 0.f;
            s2 += d0 * d0;
            s2 += d1 * d1;
            s2 += d2 * d2;
            if (!(s2 >= P.touch2) && q != static_cast<uint32_t>(p)) {
                if (s2 < P.touch2) {
                    if (seen >= consumed && listed < kContactList) {
                        s_list[listed][threadIdx.x] = q;
                        ++listed;
                    }
                    ++seen;
                } else {
                    nan_f = true;
                    if (ENERGY && __float_as_int(Sp[q].w) > id) nan_e = true;
                }
            }
        };
        for (;;) {
            seen = 0;
            walk_stencil<4, false>(GlobalPos{Sp}, kcap, 0u, 0xFFFFFFFFu, column, probe);
            consumed += listed;
            flush();
            if (seen <= consumed) break;
        }
        if (nan_f) {
            const float qnan = __builtin_nanf("");
            f[0] += qnan;
            f[1] += qnan;
            f[2] += qnan;
        }
        if (ENERGY && nan_e) e_grain += __builtin_nanf("");
        if (HIST && MODE != kModeRHS) hist.cnt[par ^ 1][p] = min(h_new, kHist);

        const float a[3] = {f[0] / m, f[1] / m, f[2] / m};
        if (MODE == kModeRHS) {
            if (rhs_out) {
                float* r = rhs_out + 6 * static_cast<size_t>(id);
                r[0] = v[0]; r[1] = v[1]; r[2] = v[2];
                r[3] = a[0]; r[4] = a[1]; r[5] = a[2];
            }
        } else {
            float vn[3], xn[3];
            if (MODE == kModeSIE) {
                #pragma unroll
                for (int k = 0; k < 3; ++k) vn[k] = v[k] + P.dt * a[k];
                #pragma unroll
                for (int k = 0; k < 3; ++k) xn[k] = x[k] + P.dt * vn[k];
            } else {
                const float4 bp = Np[p], bv = Nv[p];
                const float x0[3] = {bp.x, bp.y, bp.z}, v0[3] = {bv.x, bv.y, bv.z};
                #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    vn[k] = v0[k] + P.dt * a[k];
                    xn[k] = x0[k] + P.dt * vn[k];
                }
            }
            Tp[p] = make_float4(xn[0], xn[1], xn[2], pi.w);
            Tv[p] = make_float4(vn[0], vn[1], vn[2], vi4.w);
            lane_member = member;
            int nx, ny, nz;
            grain_cell(P, xn[0], xn[1], xn[2], nx, ny, nz);
            key_out[p] = pack_cell(nx, ny, nz);
            lo[0] = hi[0] = nx;
            lo[1] = hi[1] = ny;
            lo[2] = hi[2] = nz;
            fc[0] = nx;
            fc[1] = ny;
            fc[2] = nz;
            fid = id;
            float s2 = 0.f;
            s2 += a[0] * a[0];
            s2 += a[1] * a[1];
            s2 += a[2] * a[2];
            const float acc = sqrtf(s2);
            if (acc == acc) accel_bits = __float_as_uint(acc);
        }
    }
    if (MODE != kModeRHS) {
        const FusedCount fcount = fused_count_issue(ctl, bio, bio.mode ? P.pbox : nullptr, p < n, fc[0], fc[1], fc[2]);
        fused_count_finish(ctl, bio, fcount, fid, p, n);
    }
    if (ENERGY) {
        const double ew = group_energy<1>(static_cast<double>(e_grain));
        if ((threadIdx.x & (kEnergyGroup - 1)) == 0 && p < n) ewave[p / kEnergyGroup] = ew;
    }
    if (MODE != kModeRHS) {
        store_block_part(parts, P.members > 1 ? P.pmem : nullptr, lb, lane_member, lo, hi, accel_bits, ncont);
    }
}

// ---------------------------------------------------------------------------
// Pass 5, queued: the forces, springs and integration of k_force with the
// work of one wave (64 / L grains, L adjacent lanes per grain) spread over
// its lanes by contact rather than by grain.
//   Phase 1: the L lanes of a grain scan equal slices of the grain's stencil
//            candidates (the nine columns concatenated) and list the touching
//            ones: squared distance below touch2, which is exactly the
//            contact law's `dist < 0`.  A grain's lists, lane after lane, are
//            its contacts in stencil order, so every contact's spring slot
//            (walls first) is known before any is evaluated.
//   Phase 2: the wave's listed contacts form one queue, grain after grain,
//            evaluated 64 at a time, one per lane.  Each lane leaves its
//            force and energy addends in LDS, and the grain's lanes add them
//            in queue order, which is the order in which k_force adds them:
//            the sums are the same bits.
// A lane with more than QCAP contacts in its slice drains them over
// further passes; a grain's later lanes queue nothing until its earlier ones
// are drained, so the order holds.  A NaN candidate (a diverged state) is
// not queued: it makes the force (and, from a higher id, the energy) NaN
// whatever the order, which is applied at the end.  Blocks are one wave.
// Waves per SIMD each queued variant is compiled for: four lanes per grain
// need ~166 VGPRs, and three waves (168 at most) rather than two.
template <int L>
constexpr int kForceQWaves = L == 4 ? 3 : 1;

template <int MODE, bool ENERGY, bool HIST, int L>
__global__ __launch_bounds__(64, kForceQWaves<L>) void k_force_q(
        const Params* __restrict__ Pp, Ctl* ctl, const float4* __restrict__ Sp,
        const float4* __restrict__ Sv, const float4* __restrict__ Np,
        const float4* __restrict__ Nv, const uint32_t* __restrict__ start,
        float4* __restrict__ Tp, float4* __restrict__ Tv, uint32_t* __restrict__ key_out,
        BlockPart* __restrict__ parts, double* __restrict__ ewave, HistoryIO hist, int n, BinIO bio) {
    static_assert(MODE != kModeRHS && (L == 1 || L == 2 || L == 4), "queued force pass variants");
    constexpr int G = 64 / L;  // grains per wave
    constexpr int QCAP = kQueueCap;
    const Params& P = *Pp;
    if (bio.mode) fused_arm(P, ctl);
    // Per lane: cell-sorted slots of listed contacts; with one lane per
    // grain, row QCAP takes the stores that no list row takes.
    // One lane per grain keeps its LDS small (about 7 KB: five waves per
    // SIMD at the register count it takes without the next round's
    // prefetch): the owner's spring slot and old spring list come by
    // shuffles instead of LDS, an entry carries its owner lane (or, past 2^22
    // grains, it is found from its row's ballot), a round's records are
    // packed floats, the record area holds the staged candidate positions
    // while the stencil is walked, and the damping table holds the handle's
    // materials only (dynamic LDS, nmat x nmat).
    constexpr int NR = ENERGY ? 11 : 9;         // floats per record (L == 1)
    constexpr int RB = NR * 64 / 3;             // L == 1: staged slots (x, y, z each)
    __shared__ uint32_t s_list[L == 1 ? QCAP + 1 : QCAP][64];  // L == 1: row QCAP is the spare row
    __shared__ int s_q[L == 1 ? 1 : 64];        // per lane: queue index of its first listed contact
    __shared__ uint8_t s_src[L == 1 ? 1 : QCAP * 64];  // owner lane of each queue entry (L > 1)
    __shared__ unsigned long long s_rowmask[L == 1 ? QCAP : 1];  // L == 1: lanes listing row j
    __shared__ int s_rowstart[L == 1 ? QCAP : 1];  // L == 1: queue index of row j's first entry
    __shared__ int s_slot[L == 1 ? 1 : 64];     // per lane: spring slot of its first listed contact
    __shared__ float4 s_rec[L == 1 ? 1 : 3][L == 1 ? 1 : 64];  // per queue entry: force, energy addends
    __shared__ __align__(16) float s_r[L == 1 ? NR : 1][64];  // the same, packed (L == 1)
    __shared__ int2 s_gh[L == 1 ? 1 : G];       // per grain: old spring list (index, count)
    __shared__ float s_cpair[L == 1 ? 1 : kMaxMat * kMaxMat];
    extern __shared__ float s_cdyn[];           // L == 1: c_pair of the handle's nmat materials
    float* const s_sxyz = &s_r[0][0];  // staged x, y, z records (L == 1)
    // A lone simulation's pair damping table, in LDS: one lookup per contact.
    const bool lds_tab = P.mtab == nullptr;
    const int nmat = P.nmat;
    if (lds_tab) {
        if constexpr (L == 1) {
            for (int k = threadIdx.x; k < nmat * nmat; k += 64) s_cdyn[k] = P.c_pair[(k / nmat) * kMaxMat + k % nmat];
        } else {
            for (int k = threadIdx.x; k < kMaxMat * kMaxMat; k += 64) s_cpair[k] = P.c_pair[k];
        }
        __syncthreads();
    }
    const int lane = static_cast<int>(threadIdx.x);
    const int g = lane / L, sub = lane % L;
    const int lb = logical_block();
    const int p = lb * G + g;
    const bool owner_bits = n <= (1 << 22);  // L == 1: queue entries carry their owner lane
    const bool live = p < n;

    const int par = HIST ? ctl->hist_parity : 0;
    const float4* __restrict__ h_in = hist.buf[par];
    float4* __restrict__ h_out = hist.buf[par ^ 1];
    // Spring of `partner` in the old list (index hs, h_old entries), zero for
    // a new contact.  The old list was written in contact order, so slot
    // `guess` (the contact's new slot) normally holds it.
    // hg: slot `guess` of the old list, loaded by the caller (any slot below
    // kHist exists).
    auto h_find_at = [&](int hs, int h_old, int partner, int guess, const float4& hg,
                         float xi[3]) __attribute__((always_inline)) {
        xi[0] = xi[1] = xi[2] = 0.f;
        if (guess < h_old && __float_as_int(hg.w) == partner) {
            xi[0] = hg.x; xi[1] = hg.y; xi[2] = hg.z;
            return;
        }
        const int t = spring_slot(h_in, n, hs, h_old, partner);
        if (t >= 0) {
            const float4 h = h_in[static_cast<size_t>(t) * n + hs];
            xi[0] = h.x; xi[1] = h.y; xi[2] = h.z;
        }
    };
    auto h_find = [&](int hs, int h_old, int partner, int guess, float xi[3]) __attribute__((always_inline)) {
        h_find_at(hs, h_old, partner, guess, h_in[static_cast<size_t>(min(guess, kHist - 1)) * n + hs], xi);
    };

    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi4 = pi;
    float x[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f};
    int id = 0, mat = 0, member = 0;
    float mu = 0.f, m = 0.f;
    float f[3] = {0.f, 0.f, 0.f};
    float e_grain = 0.f;
    int hs = 0, h_old = 0, wc = 0;  // old spring list; wall contacts
    unsigned ncont = 0;
    if (live) {
        pi = Sp[p];
        vi4 = Sv[p];
        x[0] = pi.x; x[1] = pi.y; x[2] = pi.z;
        v[0] = vi4.x; v[1] = vi4.y; v[2] = vi4.z;
        id = __float_as_int(pi.w);
        mat = grain_mat(vi4.w);
        member = member_of(P, vi4.w);
        mu = P.members == 1 ? P.mu : P.mmu[member];
        m = mat_mass(P, member, mat);
        if (HIST) {
            hs = hist.hidx[p];
            h_old = hist.cnt[par][hs];
        }
        if (ENERGY) e_grain = grain_energy_self(P, x, v, m);
        // Gravity and the six walls, as k_force (every lane of the grain).
        const float c_wall = mat_cwall(P, member, mat);
        f[2] -= m * P.gravity;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const float dist = side == 0 ? x[a] - P.wlo[a] - P.radius : P.whi[a] - x[a] - P.radius;
                const float fw = fmaxf(0.f, fminf(P.max_force, P.kneg * dist));
                if (side == 0) f[a] += fw;
                else f[a] -= fw;
                if (dist < 0.f) {
                    f[a] -= c_wall * v[a];
                    if (HIST) {
                        float nrm[3] = {0.f, 0.f, 0.f};
                        nrm[a] = side == 0 ? 1.f : -1.f;
                        float vn = 0.f;
                        vn += v[0] * nrm[0];
                        vn += v[1] * nrm[1];
                        vn += v[2] * nrm[2];
                        float xi[3];
                        h_find(hs, h_old, wall_partner(a, side), wc, xi);
                        if (ENERGY) {
                            float s2 = 0.f;
                            s2 += xi[0] * xi[0];
                            s2 += xi[1] * xi[1];
                            s2 += xi[2] * xi[2];
                            e_grain += 0.5f * P.kt * s2;
                        }
                        spring_friction(P, mu, f, nrm, v, vn, fw, c_wall, xi);
                        if (sub == 0 && wc < kHist) {
                            h_out[static_cast<size_t>(wc) * n + p] =
                                    make_float4(xi[0], xi[1], xi[2], __int_as_float(wall_partner(a, side)));
                        }
                    } else {
                        float vt[3];
#pragma unroll
                        for (int b = 0; b < 3; ++b) vt[b] = (b == a) ? 0.f : v[b];
                        viscous_friction(f, vt, mu * fw, c_wall);
                    }
                    ++wc;
                }
            }
        }
    }
    if (L > 1 && sub == 0) s_gh[g] = make_int2(hs, h_old);

    const uint32_t kcap = static_cast<uint32_t>(P.kcap);
    // The grain's cell in its member's padded sub-grid and the index steps
    // of one cell in x and y: the stencil's cells lie at fixed offsets.
    uint32_t c_self = 0u, step_x = 0u, step_y = 0u;
    if (live) {
        const MemberCells mc = member_cells(P, member);
        int cx, cy, cz;
        grain_cell(P, x[0], x[1], x[2], cx, cy, cz);
        c_self = mc.index(cx, cy, cz);
        step_y = static_cast<uint32_t>(mc.g.bz);
        step_x = static_cast<uint32_t>(mc.g.by) * step_y;
    }
    auto column = [&](int c) __attribute__((always_inline)) {
        const uint32_t cell = c_self + static_cast<uint32_t>(c / 3 - 1) * step_x + static_cast<uint32_t>(c % 3 - 1) * step_y - 1u;
        uint4 q4;
        __builtin_memcpy(&q4, start + cell, sizeof(uint4));
        return q4;
    };
    // This lane's slice [s_lo, s_hi) of the grain's candidate sequence (set
    // by the first walk when L > 1).
    uint32_t s_lo = 0u, s_hi = 0xFFFFFFFFu;
    int cons = 0;     // this lane's contacts already queued
    int seen = 0;     // touching candidates met by the current scan
    int rowc = 0;     // one lane per grain: seen - cons (the list row of the next listed one)
    bool nan_f = false, nan_e = false;
    int nan_n = 0;    // one lane per grain: NaN candidates met
    const float touch2 = P.touch2;
    // One lane per grain: every slot of a batch is probed without branches
    // (an unlisted candidate's store goes to a row no list entry reads), and
    // a NaN candidate only marks the lane; with more lanes per grain, only
    // the slots in the run, with branches.
    auto probe = [&](uint32_t q, float px, float py, float pz, bool valid = true, bool self = true,
                     bool nan = true) __attribute__((always_inline)) {
        const float d0 = x[0] - px, d1 = x[1] - py, d2 = x[2] - pz;
        float s2 = 0.f;
        s2 += d0 * d0;
        s2 += d1 * d1;
        s2 += d2 * d2;
        if constexpr (L == 1) {
            const bool other = valid & (!self | (q != static_cast<uint32_t>(p)));
            const bool hit = other & (s2 < touch2);
            // Every slot stores: into the row this pass's next listed
            // candidate takes (a later one overwrites it; the rows past the
            // list are not read) or, for the candidates listed by an earlier
            // pass or past this pass's rows, into the spare row.
            const uint32_t li = min(static_cast<uint32_t>(rowc), static_cast<uint32_t>(QCAP));
            s_list[li][lane] = q;
            rowc += hit ? 1 : 0;
            if (nan) nan_n += (other & (s2 != s2)) ? 1 : 0;
        } else if (!(s2 >= touch2) && q != static_cast<uint32_t>(p)) {
            if (s2 < touch2) {
                if (seen >= cons && seen < cons + QCAP) s_list[seen - cons][lane] = q;
                ++seen;
            } else {
                nan_f = true;
                if (__float_as_int(Sp[q].w) > id) nan_e = true;
            }
        }
    };

    int total = 0;      // touching candidates in this lane's slice
    int lane_slot = 0;  // spring slot of this lane's first contact
    int grain_slots = 0;
    // One lane per grain, lone simulation: the cells a wave's grains reach
    // with one dx of the stencil are the contiguous range from its first
    // grain's cell - (dy, dz) = (1, 1) to its last grain's + (1, 1) (the
    // grains are in cell order and the sub-grid is padded), so their
    // candidates are one contiguous range of slots.  Each dx group's range
    // is staged in LDS with coalesced loads and walked from there; a range
    // over RB slots is walked from global memory.
    // Blocks past the last grain (the grid is rounded up to whole XCD
    // slabs) stage nothing; a range outside the cell-start array (which the
    // padded sub-grid rules out) is walked from global memory.
    uint32_t c_first = 0u, c_last = 0u, Sx = 0u, Sy = 0u;
    const bool stage_ok = L == 1 && P.members == 1 && lb * G < n;
    // every grain of the wave at a finite position
    const bool own_finite =
            __ballot(live && !(__builtin_isfinite(x[0]) && __builtin_isfinite(x[1]) && __builtin_isfinite(x[2]))) ==
            0ull;
    if constexpr (L == 1) {
        if (stage_ok) {
            const MemberCells m0 = member_cells(P, 0);
            // lane 0 holds the block's first grain, lane n - 1 - lb G (or
            // 63) its last
            c_first = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(c_self)));
            c_last = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c_self), min(63, n - 1 - lb * G)));
            Sx = static_cast<uint32_t>(m0.g.by * m0.g.bz);
            Sy = static_cast<uint32_t>(m0.g.bz);
        }
    }
    // One lane per grain: the slot ranges [g_lo, g_lo + len) of the three dx
    // groups (uniform), loaded together before the walk; a group whose cell
    // range leaves the cell-start array (ruled out by the padded sub-grid)
    // keeps length ~0 and is walked from global memory.
    uint32_t g_lo0 = 0u, g_lo1 = 0u, g_lo2 = 0u, g_len0 = ~0u, g_len1 = ~0u, g_len2 = ~0u;
    if constexpr (L == 1) {
        if (stage_ok) {
            const uint32_t n_cells = P.mcell[1];  // lone simulation: start[] holds n_cells + 1 entries
            auto range = [&](int gx, uint32_t& g_lo, uint32_t& len) __attribute__((always_inline)) {
                const uint32_t lo = c_first + static_cast<uint32_t>(gx - 1) * Sx - Sy - 1u;
                const uint32_t hi = c_last + static_cast<uint32_t>(gx - 1) * Sx + Sy + 2u;
                if (lo <= hi && hi <= n_cells) {
                    g_lo = start[lo];
                    len = start[hi] - g_lo;
                }
            };
            range(0, g_lo0, g_len0);
            range(1, g_lo1, g_len1);
            range(2, g_lo2, g_len2);
        }
    }
    for (int pass = 0;; ++pass) {
        seen = 0;
        rowc = -cons;
        const bool walking = live && (pass == 0 || cons < total);
        if constexpr (L == 1) {
#pragma unroll 1
            for (int gx = 0; gx < 3; ++gx) {
                // The group's column quads are in flight while its positions
                // are staged: a group waits on one load after the first.
                uint4 qa0{}, qa1{}, qa2{};
                if (walking) {
                    qa0 = column(3 * gx);
                    qa1 = column(3 * gx + 1);
                    qa2 = column(3 * gx + 2);
                }
                // a column as its run (begin, length), ~0: a cell over
                // capacity; (0, 0) for a lane with nothing to walk
                auto run_of = [&](const uint4& q, uint32_t& b, uint32_t& l) __attribute__((always_inline)) {
                    b = walking ? q.x : 0u;
                    l = !walking ? 0u : (column_capped(q, kcap) ? ~0u : q.w - q.x);
                };
                const uint32_t g_lo = gx == 0 ? g_lo0 : (gx == 1 ? g_lo1 : g_lo2);
                const uint32_t len = gx == 0 ? g_len0 : (gx == 1 ? g_len1 : g_len2);
                // (at most RB - 4 slots: a round of staged reads clamped into
                // the area then holds no slot of a run)
                const bool staged = len <= static_cast<uint32_t>(RB - 4);
                bool bad = false;  // a staged position that is not finite
                if (staged) {
                    for (uint32_t k = static_cast<uint32_t>(lane); k < len; k += 64u) {
                        const float4 a = Sp[g_lo + k];
                        s_sxyz[3 * k] = a.x;
                        s_sxyz[3 * k + 1] = a.y;
                        s_sxyz[3 * k + 2] = a.z;
                        bad |= !(__builtin_isfinite(a.x) && __builtin_isfinite(a.y) && __builtin_isfinite(a.z));
                    }
                }
                const bool finite = own_finite && __ballot(bad) == 0ull;
                __syncthreads();
                uint32_t b0, l0, b1, l1, b2, l2;
                run_of(qa0, b0, l0);
                run_of(qa1, b1, l1);
                run_of(qa2, b2, l2);
                if (staged) {
                    walk_columns3_staged<3>(s_sxyz, g_lo, static_cast<uint32_t>(RB), finite, kcap, 3 * gx, b0, l0,
                                            b1, l1, b2, l2, column, probe);
                } else if (walking) {
                    walk_columns3<4>(Sp, kcap, 3 * gx, b0, l0, b1, l1, b2, l2, column, probe);
                }
                __syncthreads();  // the staged range is read before the next one is written
            }
        } else if (walking) {
            if constexpr (L == 4) {
                // Lanes 0, 1, 2 of the grain walk dx groups 0, 1, 2 (a
                // contiguous part of the sequence each, in lane order);
                // lane 3 lists nothing.
                if (sub < 3) walk_stencil<12, true>(GlobalPos{Sp}, kcap, 0u, 0xFFFFFFFFu, column, probe, sub, sub + 1);
            } else {
                // The grain's nine columns: its sequence length (for the
                // slices) and, when no cell is over capacity, the runs.
                uint32_t cb[9], cpre[9], tot = 0;
                bool capped = false;
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    const uint4 a = column(c);
                    capped |= column_capped(a, kcap);
                    cb[c] = a.x;
                    cpre[c] = tot;
                    tot += column_length(a, kcap);
                }
                if (pass == 0) {
                    s_lo = tot * static_cast<uint32_t>(sub) / L;
                    s_hi = tot * static_cast<uint32_t>(sub + 1) / L;
                }
                if (capped) walk_stencil<8, true>(GlobalPos{Sp}, kcap, s_lo, s_hi, column, probe);
                else walk_slice<8>(Sp, cb, cpre, s_lo, s_hi, probe);
            }
        }
        if (pass == 0) {
            total = L == 1 ? rowc : seen;  // cons is 0
            int b = wc;
#pragma unroll
            for (int k = 1; k < L; ++k) {
                const int o = __shfl(total, lane - k, 64);
                if (sub >= k) b += o;
            }
            lane_slot = b;
            grain_slots = __shfl(lane_slot + total, g * L + L - 1, 64);
        }
        const int listed = min(total - cons, QCAP);
        const unsigned long long drained = __ballot(cons + listed == total);
        const unsigned long long before = ((1ull << sub) - 1ull) << (g * L);
        const int cnt = (drained & before) == before ? listed : 0;
        int T = 0, gq0 = 0, gq1 = 0;
        uint32_t* const s_flat = &s_list[0][0];
        if constexpr (L == 1) {
            // Column-major queue: every grain's first listed contact, then
            // every grain's second, and so on.  A grain's contacts keep their
            // order, and a round of 64 holds few of any one grain's, so the
            // per-grain sums after a round are short.  The list is compacted
            // in place: row j's entries land below index 64 (j + 1), where
            // no unread row lies.  An entry is the partner slot with its row
            // in the top four bits; below 2^22 grains the owner lane sits in
            // bits 22-27, otherwise (slots below 2^28) the owner is found
            // from the row's ballot.
            for (int j = 0; j < QCAP; ++j) {
                const unsigned long long mask = __ballot(j < cnt);
                if (mask == 0ull) break;
                const uint32_t q = s_list[j][lane];
                __syncthreads();  // the row is read before any of it is overwritten
                if (j < cnt) {
                    const int pos = T + lanes_below(mask);
                    s_flat[pos] = q | (static_cast<uint32_t>(j) << 28) |
                                  (owner_bits ? static_cast<uint32_t>(lane) << 22 : 0u);
                }
                if (!owner_bits && lane == 0) {
                    s_rowmask[j] = mask;
                    s_rowstart[j] = T;
                }
                T += __popcll(mask);
            }
        } else {
            const int incl = wave_incl_scan_dpp(cnt);
            const int q0 = incl - cnt;
            T = __builtin_amdgcn_readlane(incl, 63);
            gq0 = __shfl(q0, g * L, 64);
            const int gq1n = __shfl(q0, ((g + 1) * L) & 63, 64);
            gq1 = (g + 1) * L < 64 ? gq1n : T;
            s_q[lane] = q0;
#pragma unroll
            for (int j = 0; j < QCAP; ++j) {
                if (j < cnt) s_src[q0 + j] = static_cast<uint8_t>(lane);
            }
        }
        if (L > 1) s_slot[lane] = lane_slot + cons;
        const int slot_base = lane_slot + cons;  // L == 1: read by shuffles
        __syncthreads();
        // One queue entry's indices and loads; the owner grain's own state
        // comes from its lanes' registers by shuffles.  The next round's
        // entries are fetched while the current round is evaluated.  An
        // index past the queue fetches its last entry: every load is
        // unconditional.
        struct Entry {
            int slot, og, hs, hold;  // hs, hold: the owner's old spring list (L == 1)
            uint32_t q;
            float4 pj, vj, hg;
        };
        auto fetch = [&](int e, Entry& t) __attribute__((always_inline)) {
            if constexpr (L == 1) {
                const uint32_t w = s_flat[e];
                const int j = static_cast<int>(w >> 28);
                // the owner: the (e - start)-th lane of the row's ballot
                int src;
                if (owner_bits) {
                    src = static_cast<int>((w >> 22) & 63u);
                    t.q = w & 0x003FFFFFu;
                } else {
                    src = nth_lane(s_rowmask[j], e - s_rowstart[j]);
                    t.q = w & 0x0FFFFFFFu;
                }
                t.slot = __shfl(slot_base, src, 64) + j;
                t.og = src;
                t.hs = HIST ? __shfl(hs, src, 64) : 0;
                t.hold = HIST ? __shfl(h_old, src, 64) : 0;
            } else {
                const int src = s_src[e];
                const int j = e - s_q[src];
                t.q = s_list[j][src];
                t.slot = s_slot[src] + j;
                t.og = src / L;
            }
            t.pj = Sp[t.q];
            t.vj = Sv[t.q];
            if (HIST) t.hg = h_in[static_cast<size_t>(min(t.slot, kHist - 1)) * n + (L == 1 ? t.hs : s_gh[t.og].x)];
        };
        // With one lane per grain the next round is not prefetched: its
        // loads would hold 17 more registers and cost a wave per SIMD.
        constexpr bool kAhead = L != 1;
        Entry cur, nxt;
        int jc = 0, cjc = 0;  // one lane per grain: first row in the round, its first queue index
        if (kAhead && T > 0) fetch(min(lane, T - 1), cur);
        for (int base = 0; base < T; base += 64) {
            if (kAhead) {
                if (base + 64 < T) fetch(min(base + 64 + lane, T - 1), nxt);
            } else {
                fetch(min(base + lane, T - 1), cur);
            }
            const int e = base + lane;
            // the owner's state, from a lane of its grain (every lane takes part)
            const int ol = cur.og * L;
            const float ox[3] = {__shfl(x[0], ol, 64), __shfl(x[1], ol, 64), __shfl(x[2], ol, 64)};
            const float ov[3] = {__shfl(v[0], ol, 64), __shfl(v[1], ol, 64), __shfl(v[2], ol, 64)};
            const int oid = __shfl(id, ol, 64);
            const float ovw = __shfl(vi4.w, ol, 64);
            const float omu = __shfl(mu, ol, 64);
            float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0;
            if (e < T) {
                const int slot = cur.slot;
                const int og = cur.og;
                const int po = lb * G + og;
                const int omat = grain_mat(ovw);
                const int omember = member_of(P, ovw);
                const float4 pj = cur.pj;
                float d0, d1, d2;
                const float norm = pair_norm(ox, pj.x, pj.y, pj.z, d0, d1, d2);
                const int jd = __float_as_int(pj.w);
                if (ENERGY && jd > oid) r0.w = contact_energy(P, 2.f * P.radius - norm);
                const float dist = norm - P.radius - P.radius;
                const float nrm[3] = {d0 / norm, d1 / norm, d2 / norm};
                const float fp = fmaxf(0.f, fminf(P.max_force, P.kneg * dist));
                r0.x = fp * nrm[0];
                r0.y = fp * nrm[1];
                r0.z = fp * nrm[2];
                ++ncont;
                const float4 vj = cur.vj;
                const float c = !lds_tab ? mat_cpair(P, omember, omat, grain_mat(vj.w))
                                : L == 1 ? s_cdyn[omat * nmat + grain_mat(vj.w)]
                                         : s_cpair[omat * kMaxMat + grain_mat(vj.w)];
                const float vr[3] = {ov[0] - vj.x, ov[1] - vj.y, ov[2] - vj.z};
                float vn = 0.f;
                vn += vr[0] * nrm[0];
                vn += vr[1] * nrm[1];
                vn += vr[2] * nrm[2];
                r1.x = -(c * vn * nrm[0]);
                r1.y = -(c * vn * nrm[1]);
                r1.z = -(c * vn * nrm[2]);
                if (HIST) {
                    float xi[3];
                    if constexpr (L == 1) {
                        h_find_at(cur.hs, cur.hold, jd, slot, cur.hg, xi);
                    } else {
                        const int2 gh = s_gh[og];
                        h_find_at(gh.x, gh.y, jd, slot, cur.hg, xi);
                    }
                    if (ENERGY && jd > oid) {
                        float s2 = 0.f;
                        s2 += xi[0] * xi[0];
                        s2 += xi[1] * xi[1];
                        s2 += xi[2] * xi[2];
                        r1.w = 0.5f * P.kt * s2;
                    }
                    float ft[3] = {0.f, 0.f, 0.f};
                    spring_friction(P, omu, ft, nrm, vr, vn, fp, c, xi);
                    r2.x = ft[0];
                    r2.y = ft[1];
                    r2.z = ft[2];
                    if (slot < kHist) {
                        h_out[static_cast<size_t>(slot) * n + po] =
                                make_float4(xi[0], xi[1], xi[2], __int_as_float(jd));
                    }
                } else {
                    float ft[3] = {0.f, 0.f, 0.f};
                    float vt[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) vt[a] = vr[a] - vn * nrm[a];
                    viscous_friction(ft, vt, omu * fp, c);
                    r2.x = ft[0];
                    r2.y = ft[1];
                    r2.z = ft[2];
                }
            }
            if constexpr (L == 1) {
                s_r[0][lane] = r0.x;
                s_r[1][lane] = r0.y;
                s_r[2][lane] = r0.z;
                s_r[3][lane] = r1.x;
                s_r[4][lane] = r1.y;
                s_r[5][lane] = r1.z;
                s_r[6][lane] = r2.x;
                s_r[7][lane] = r2.y;
                s_r[8][lane] = r2.z;
                if constexpr (ENERGY) {
                    s_r[NR - 2][lane] = r0.w;
                    s_r[NR - 1][lane] = r1.w;
                }
            } else {
                s_rec[0][lane] = r0;
                s_rec[1][lane] = r1;
                s_rec[2][lane] = r2;
            }
            __syncthreads();
            auto add_rec = [&](const float4& a0, const float4& a1, const float4& a2) __attribute__((always_inline)) {
                f[0] += a0.x;
                f[1] += a0.y;
                f[2] += a0.z;
                f[0] += a1.x;
                f[1] += a1.y;
                f[2] += a1.z;
                f[0] += a2.x;
                f[1] += a2.y;
                f[2] += a2.z;
                if (ENERGY) {
                    e_grain += a0.w;
                    e_grain += a1.w;
                }
            };
            auto add = [&](int r) __attribute__((always_inline)) {
                if constexpr (L == 1) {
                    const float4 a0 = make_float4(s_r[0][r], s_r[1][r], s_r[2][r], ENERGY ? s_r[NR - 2][r] : 0.f);
                    const float4 a1 = make_float4(s_r[3][r], s_r[4][r], s_r[5][r], ENERGY ? s_r[NR - 1][r] : 0.f);
                    const float4 a2 = make_float4(s_r[6][r], s_r[7][r], s_r[8][r], 0.f);
                    add_rec(a0, a1, a2);
                } else {
                    add_rec(s_rec[0][r], s_rec[1][r], s_rec[2][r]);
                }
            };
            if constexpr (L == 1) {
                // Rows jc, jc + 1, ... overlap the round; a grain adds its
                // entry of each, in row order.
                for (int j = jc, cj = cjc; j < QCAP && cj < base + 64; ++j) {
                    const unsigned long long mask = __ballot(j < cnt);
                    if (mask == 0ull) break;
                    const int r = cj + lanes_below(mask) - base;
                    if (j < cnt && r >= 0 && r < 64) add(r);
                    const int cn = cj + __popcll(mask);
                    if (cn <= base + 64) {
                        jc = j + 1;
                        cjc = cn;
                    }
                    cj = cn;
                }
            } else {
                // A grain's entries of the round, two at a time with both
                // records' reads in flight (the second clamped to the range).
                const int r_lo = max(gq0, base), r_hi = min(gq1, base + 64);
                for (int r = r_lo; r < r_hi; r += 2) {
                    const int r2 = min(r + 1, r_hi - 1) - base;
                    const float4 b0 = s_rec[0][r2], b1 = s_rec[1][r2], b2 = s_rec[2][r2];
                    add(r - base);
                    if (r + 1 < r_hi) add_rec(b0, b1, b2);
                }
            }
            __syncthreads();
            if (kAhead) cur = nxt;
        }
        cons += cnt;
        if (__ballot(cons < total) == 0ull) break;
    }
    if constexpr (L == 1) {
        // The walk only counted the NaN candidates a lane met: its force is
        // NaN, and its energy is when such a candidate has the higher id (a
        // pair's energy is counted from its lower id), which a second walk
        // of the stencil finds (a diverged state only).
        nan_f = nan_n != 0;
        if (ENERGY && nan_f) {
            walk_stencil<4, false, true>(GlobalPos{Sp}, kcap, 0u, 0xFFFFFFFFu, column,
                                         [&](uint32_t q, float px, float py, float pz, bool valid)
                                                 __attribute__((always_inline)) {
                const float d0 = x[0] - px, d1 = x[1] - py, d2 = x[2] - pz;
                float s2 = 0.f;
                s2 += d0 * d0;
                s2 += d1 * d1;
                s2 += d2 * d2;
                if (valid && q != static_cast<uint32_t>(p) && s2 != s2 && __float_as_int(Sp[q].w) > id) nan_e = true;
            });
        }
    }
    const unsigned long long gmask = ((L == 1) ? 1ull : ((1ull << L) - 1ull)) << (g * L);
    if (__ballot(nan_f) & gmask) {
        const float qnan = __builtin_nanf("");
        f[0] += qnan;
        f[1] += qnan;
        f[2] += qnan;
    }
    if (ENERGY && (__ballot(nan_e) & gmask)) e_grain += __builtin_nanf("");

    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    unsigned accel_bits = 0;
    int lane_member = -1;
    int fc[3] = {0, 0, 0};  // the grain's next cell (fused count)
    if (live && sub == 0) {
        if (HIST) hist.cnt[par ^ 1][p] = min(grain_slots, kHist);
        const float a[3] = {f[0] / m, f[1] / m, f[2] / m};
        float vn[3], xn[3];
        if (MODE == kModeSIE) {
#pragma unroll
            for (int k = 0; k < 3; ++k) vn[k] = v[k] + P.dt * a[k];
#pragma unroll
            for (int k = 0; k < 3; ++k) xn[k] = x[k] + P.dt * vn[k];
        } else {
            const float4 bp = Np[p], bv = Nv[p];
            const float x0[3] = {bp.x, bp.y, bp.z}, v0[3] = {bv.x, bv.y, bv.z};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vn[k] = v0[k] + P.dt * a[k];
                xn[k] = x0[k] + P.dt * vn[k];
            }
        }
        Tp[p] = make_float4(xn[0], xn[1], xn[2], pi.w);
        Tv[p] = make_float4(vn[0], vn[1], vn[2], vi4.w);
        lane_member = member;
        int nx, ny, nz;
        grain_cell(P, xn[0], xn[1], xn[2], nx, ny, nz);
        key_out[p] = pack_cell(nx, ny, nz);
        lo[0] = hi[0] = nx;
        lo[1] = hi[1] = ny;
        lo[2] = hi[2] = nz;
        fc[0] = nx;
        fc[1] = ny;
        fc[2] = nz;
        float s2 = 0.f;
        s2 += a[0] * a[0];
        s2 += a[1] * a[1];
        s2 += a[2] * a[2];
        const float acc = sqrtf(s2);
        if (acc == acc) accel_bits = __float_as_uint(acc);
        ncont += static_cast<unsigned>(wc);
    }
    // (issued and filed here: holding the count's registers across the
    // reductions below costs the one-lane pass a wave per SIMD)
    const FusedCount fcount =
            fused_count_issue(ctl, bio, bio.mode ? P.pbox : nullptr, live && sub == 0, fc[0], fc[1], fc[2]);
    fused_count_finish(ctl, bio, fcount, id, p, n);
    if (ENERGY) {
        const double ew = group_energy<L>(live ? static_cast<double>(e_grain) : 0.0);
        if (sub == 0 && (g & (kEnergyGroup - 1)) == 0 && live) ewave[p / kEnergyGroup] = ew;
    }
    store_wave_part(parts, P.members > 1 ? P.pmem : nullptr, lb, lane_member, lo, hi, accel_bits, ncont);
}

// Largest CIC deposit magnitude of one grain as float bits (0 for a
// massless grain, +inf bits for a non-finite one).
__device__ __forceinline__ unsigned cic_grain_bits(const Params& P, float4 p, float4 v) {
    const float m = mat_mass(P, member_of(P, v.w), grain_mat(v.w));
    float mag = fabsf(m);
    mag = fmaxf(mag, fabsf(m * v.x));
    mag = fmaxf(mag, fabsf(m * v.y));
    mag = fmaxf(mag, fabsf(m * v.z));
    const bool finite = isfinite(m * v.x) && isfinite(m * v.y) && isfinite(m * v.z) &&
                        isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(m);
    return (m == 0.f) ? 0u : (finite ? __float_as_uint(mag) : 0x7F800000u);
}

// Reduces, with one wave, the boxes that the block partials hold for member
// m over its grain range [mgoff[m], mgoff[m+1]) into v (valid in every lane:
// min lo[3], max hi[3]).  With first / stride the wave takes partials
// first, first + stride, ... (its share of a block-wide reduction).  A partial block whose first and last members are
// both other than m (it spans three or more members) is folded grain by
// grain with `grain(i, lo, hi)`.
template <class Grain>
__device__ void wave_member_reduce(const Params& P, const BlockPart* __restrict__ parts, int gpp,
                                   int m, int v[6], Grain grain, int first = -1, int stride = 64) {
    const int lane = first < 0 ? static_cast<int>(threadIdx.x & 63) : first;
    const int bs = gpp;  // grains per partial
    const int g0 = P.mgoff[m], g1 = P.mgoff[m + 1];
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    if (g1 > g0) {
        for (int b = g0 / bs + lane; b <= (g1 - 1) / bs; b += stride) {
            const BlockPart& bp = parts[b];
            const BlockMembers& bm = P.pmem[b];
            if (bm.m0 == m || bm.m1 == m) {
                const int* bx = bm.m0 == m ? bp.box0 : bm.box1;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    lo[k] = min(lo[k], bx[k]);
                    hi[k] = max(hi[k], bx[3 + k]);
                }
            } else {
                for (int i = max(g0, b * bs); i < min(g1, (b + 1) * bs); ++i) grain(i, lo, hi);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[k] = wave_min_dpp(lo[k]);
        v[3 + k] = wave_max_dpp(hi[k]);
    }
}

// k_finalize for a lone simulation's force pass: one pass over the partials
// and the energy groups, one block-wide combine.  The sums run in the order
// of the general path (per thread strided, then waves in order), so the bits
// are the same.
// The reduction of a lone simulation's block partials (the slices, on this
// block).
__device__ void finalize_single(const Params& P, Ctl* ctl, const BlockPart* __restrict__ parts, int nparts,
                                const double* __restrict__ ewave, double* __restrict__ energy_log,
                                int flags) {
    slices_on_block(P, ctl, parts, nparts, ewave, energy_log, flags, nullptr);
}

// Members k_finalize reduces through LDS (more take the general path below).
constexpr int kFinMembers = 512;

// k_finalize for up to kFinMembers members: one pass over the partials, each
// adding its first and last member's box (force pass: occupied cells; CIC
// scale pass: the largest deposit magnitude, in [3]) to that member's entry in
// LDS with integer min / max atomics (order-free, so the result is the general
// path's), and the grains of a member lying wholly inside one block folded
// one by one.  The energy and contact sums run in the order of the general
// path, so their bits are the same.
__device__ void finalize_members(const Params& P, Ctl* ctl, const BlockPart* __restrict__ parts, int nparts,
                                 int gpp, const uint32_t* __restrict__ key, const float4* __restrict__ Tp,
                                 const float4* __restrict__ Tv, const double* __restrict__ ewave,
                                 double* __restrict__ energy_log, int flags) {
    __shared__ int s_mb[kFinMembers][6];
    __shared__ int s_iv[16];
    __shared__ double s_dv[16][2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    const int E = P.members;
    const bool scale = (flags & kFinCicScale) != 0;
    for (int k = tid; k < 6 * E; k += blockDim.x) s_mb[k / 6][k % 6] = (k % 6) < 3 ? kEmptyMin : kEmptyMax;
    __syncthreads();
    auto merge = [&](int m, const int* b) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (b[k] != kEmptyMin) atomicMin(&s_mb[m][k], b[k]);
            if (b[3 + k] != kEmptyMax) atomicMax(&s_mb[m][3 + k], b[3 + k]);
        }
    };
    unsigned am = 0u;
    unsigned long long csum = 0;
    const int rounds = (nparts + static_cast<int>(blockDim.x) - 1) / static_cast<int>(blockDim.x);
    for (int r = 0; r < rounds; ++r) {  // the same trip count in every lane
        const int b = r * static_cast<int>(blockDim.x) + tid;
        const bool valid = b < nparts;
        const BlockPart* bp = parts + (valid ? b : 0);
        // A lone simulation writes no member part: every partial is member 0's
        // (an empty block's box is the identity, which merges as a no-op).
        const BlockMembers* bm = P.members > 1 ? P.pmem + (valid ? b : 0) : nullptr;
        const int m0 = valid ? (bm ? bm->m0 : 0) : -1, m1 = valid ? (bm ? bm->m1 : 0) : -1;
        if (valid) {
            am = max(am, bp->accel);
            csum += bp->contacts;
        }
        if (!(flags & (kFinBox | kFinCicScale))) continue;
        // Consecutive partials almost always belong to one member: a wave
        // whose partials all hold one member (and no second one) merges
        // its reduced box with one set of atomics.
        const int mw = __shfl(m0, 0, 64);
        const bool one = __all(!valid || (m0 == mw && m1 == m0)) && mw >= 0;
        if (one) {
            int w6[6];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                w6[k] = wave_min_dpp(valid ? bp->box0[k] : kEmptyMin);
                w6[3 + k] = wave_max_dpp(valid ? bp->box0[3 + k] : kEmptyMax);
            }
            if (lane == 0) merge(mw, w6);
            continue;
        }
        if (m0 < 0) continue;
        merge(m0, bp->box0);
        if (m1 > m0) merge(m1, bm->box1);
        for (int mm = m0 + 1; mm < m1; ++mm) {  // members wholly inside this block
            int g[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
            for (int i = P.mgoff[mm]; i < P.mgoff[mm + 1]; ++i) {
                if (scale) {
                    g[3] = max(g[3], static_cast<int>(cic_grain_bits(P, Tp[i], Tv[i])));
                } else {
                    int c3[3];
                    unpack_cell(key[i], c3[0], c3[1], c3[2]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        g[k] = min(g[k], c3[k]);
                        g[3 + k] = max(g[3 + k], c3[k]);
                    }
                }
            }
            merge(mm, g);
        }
    }
    double e = 0.0;
    if (flags & (kFinEnergyLog | kFinEnergyLast)) {
        e = ewave_share(ewave, (P.n + kEnergyGroup - 1) / kEnergyGroup, tid, static_cast<int>(blockDim.x));
    }
    e = wave_sum(e);
    const double cw = wave_sum(static_cast<double>(csum));  // exact below 2^53
    am = wave_max_u_dpp(am);
    if (lane == 0) {
        s_dv[wave][0] = e;
        s_dv[wave][1] = cw;
        s_iv[wave] = static_cast<int>(am);
    }
    __syncthreads();  // also completes the LDS atomics
    if (flags & kFinBox) {
        for (int k = tid; k < 6 * E; k += blockDim.x) P.mbox[k] = s_mb[k / 6][k % 6];
    }
    if (scale) {
        for (int m = tid; m < E; m += blockDim.x) P.mscale[m] = s_mb[m][3] < 0 ? 0u : static_cast<unsigned>(s_mb[m][3]);
    }
    if (tid != 0) return;
    if (flags & kFinBox) {
        uint32_t off = 0;
        for (int m = 0; m < E; ++m) {
            P.mcell[m] = off;
            const int* bx = s_mb[m];
            off += static_cast<uint32_t>((bx[3] - bx[0] + 3) * (bx[4] - bx[1] + 3) * (bx[5] - bx[2] + 3));
        }
        P.mcell[E] = off;
    }
    double total = 0.0, contacts = 0.0;
    for (int w = 0; w < nw; ++w) {
        total += s_dv[w][0];
        contacts += s_dv[w][1];
        am = max(am, static_cast<unsigned>(s_iv[w]));
    }
    if (flags & kFinAccel) {
        ctl->accel_last = am;
        ctl->force_passes = ctl->force_passes + 1;
        ctl->contact_sum = ctl->contact_sum + static_cast<unsigned long long>(contacts);
    }
    if (flags & kFinEnergyLast) ctl->energy_last = total;
    if (flags & kFinFlipHistory) ctl->hist_parity = ctl->hist_parity ^ 1;
    if ((flags & kFinEnergyLog) && ctl->energy_count < ctl->energy_capacity) {
        energy_log[ctl->energy_count] = total;
        ctl->energy_count = ctl->energy_count + 1;
    }
}

__global__ __launch_bounds__(kFinalizeBlock) void k_finalize(const Params* __restrict__ Pp, Ctl* ctl,
                                                     const BlockPart* __restrict__ parts,
                                                     int nparts, int gpp,
                                                     const uint32_t* __restrict__ key,
                                                     const float4* __restrict__ Tp,
                                                     const float4* __restrict__ Tv,
                                                     const double* __restrict__ ewave,
                                                     double* __restrict__ energy_log, int flags) {
    const Params& P = *Pp;
    if (P.members == 1 && !(flags & kFinCicScale)) {
        finalize_single(P, ctl, parts, nparts, ewave, energy_log, flags);
        return;
    }
    if (P.members <= kFinMembers) {
        finalize_members(P, ctl, parts, nparts, gpp, key, Tp, Tv, ewave, energy_log, flags);
        return;
    }
    __shared__ int s_box[16][1];
    unsigned am = 0u;
    double e = 0.0;
    unsigned long long csum = 0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        am = max(am, parts[b].accel);
        csum += parts[b].contacts;
    }
    if (flags & (kFinEnergyLog | kFinEnergyLast)) {
        e = ewave_share(ewave, (P.n + kEnergyGroup - 1) / kEnergyGroup, static_cast<int>(threadIdx.x),
                        static_cast<int>(blockDim.x));
    }
    const double total = block_sum(e);
    const double contacts = block_sum(static_cast<double>(csum));  // exact below 2^53
    const unsigned amax_w = wave_max_u_dpp(am);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    if (lane == 0) s_box[wave][0] = static_cast<int>(amax_w);
    __syncthreads();
    unsigned amax = 0u;
    for (int w = 0; w < nw; ++w) amax = max(amax, static_cast<unsigned>(s_box[w][0]));
    __syncthreads();
    // A member with many partials is reduced by the whole block (members in
    // order), the others by one wave each.
    __shared__ int s_big[16][6];
    const int tid = threadIdx.x;
    auto big = [&](int m) { return (P.mgoff[m + 1] - P.mgoff[m]) / gpp > 256; };
    auto block_combine = [&](int v[6]) {
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) s_big[wave][k] = v[k];
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < nw; ++w) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    v[k] = min(v[k], s_big[w][k]);
                    v[3 + k] = max(v[3 + k], s_big[w][3 + k]);
                }
            }
        }
        __syncthreads();
    };
    if (flags & (kFinBox | kFinCicScale)) {
        for (int m = 0; m < P.members; ++m) {
            if (!big(m)) continue;
            int v[6];
            if (flags & kFinBox) {
                wave_member_reduce(P, parts, gpp, m, v, [&](int i, int lo[3], int hi[3]) {
                    int c3[3];
                    unpack_cell(key[i], c3[0], c3[1], c3[2]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = min(lo[k], c3[k]);
                        hi[k] = max(hi[k], c3[k]);
                    }
                }, tid, static_cast<int>(blockDim.x));
                block_combine(v);
                if (tid == 0) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) P.mbox[6 * m + k] = v[k];
                }
            }
            if (flags & kFinCicScale) {
                wave_member_reduce(P, parts, gpp, m, v, [&](int i, int lo[3], int hi[3]) {
                    hi[0] = max(hi[0], static_cast<int>(cic_grain_bits(P, Tp[i], Tv[i])));
                }, tid, static_cast<int>(blockDim.x));
                block_combine(v);
                if (tid == 0) P.mscale[m] = v[3] < 0 ? 0u : static_cast<unsigned>(v[3]);
            }
        }
        for (int m = wave; m < P.members; m += nw) {
            if (big(m)) continue;
            int v[6];
            if (flags & kFinBox) {
                wave_member_reduce(P, parts, gpp, m, v, [&](int i, int lo[3], int hi[3]) {
                    int c3[3];
                    unpack_cell(key[i], c3[0], c3[1], c3[2]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = min(lo[k], c3[k]);
                        hi[k] = max(hi[k], c3[k]);
                    }
                });
                if (lane < 6) P.mbox[6 * m + lane] = v[lane];
            }
            if (flags & kFinCicScale) {
                wave_member_reduce(P, parts, gpp, m, v, [&](int i, int lo[3], int hi[3]) {
                    hi[0] = max(hi[0], static_cast<int>(cic_grain_bits(P, Tp[i], Tv[i])));
                });
                if (lane == 0) P.mscale[m] = v[3] < 0 ? 0u : static_cast<unsigned>(v[3]);
            }
        }
    }
    if (flags & kFinBox) {
        __syncthreads();  // every member's box is stored
        if (threadIdx.x == 0) {
            uint32_t off = 0;
            for (int m = 0; m < P.members; ++m) {
                P.mcell[m] = off;
                const int* bx = P.mbox + 6 * m;
                off += static_cast<uint32_t>((bx[3] - bx[0] + 3) * (bx[4] - bx[1] + 3) * (bx[5] - bx[2] + 3));
            }
            P.mcell[P.members] = off;
        }
    }
    if (threadIdx.x != 0) return;
    if (flags & kFinAccel) {
        ctl->accel_last = amax;
        ctl->force_passes = ctl->force_passes + 1;
        ctl->contact_sum = ctl->contact_sum + static_cast<unsigned long long>(contacts);
    }
    if (flags & kFinEnergyLast) ctl->energy_last = total;
    if (flags & kFinFlipHistory) ctl->hist_parity = ctl->hist_parity ^ 1;
    if ((flags & kFinEnergyLog) && ctl->energy_count < ctl->energy_capacity) {
        energy_log[ctl->energy_count] = total;
        ctl->energy_count = ctl->energy_count + 1;
    }
}

#define WD_INST_REORDER(M, B)                                                                \
    template __global__ void k_reorder<M, B>(const Params*, Ctl*, const uint32_t*,                \
                                             const uint32_t*, const int*, const uint32_t*,        \
                                             const uint2*, const float4*, const float4*, float4*, \
                                             float4*, Carry, int, ReorderFuse);
template __global__ void k_count<false>(const Params*, Ctl*, const float4*, const float4*, const uint32_t*,
                                        uint32_t*, uint32_t*, int, int*, uint32_t*, uint32_t*, uint2*, int);
template __global__ void k_count<true>(const Params*, Ctl*, const float4*, const float4*, const uint32_t*,
                                       uint32_t*, uint32_t*, int, int*, uint32_t*, uint32_t*, uint2*, int);
template __global__ void k_scan<4>(const Params*, Ctl*, uint32_t*, uint32_t*, unsigned long long*, ScanFuse);
template __global__ void k_scan<8>(const Params*, Ctl*, uint32_t*, uint32_t*, unsigned long long*, ScanFuse);
template __global__ void k_scan<16>(const Params*, Ctl*, uint32_t*, uint32_t*, unsigned long long*, ScanFuse);
WD_INST_REORDER(0, false)
WD_INST_REORDER(1, false)
WD_INST_REORDER(2, false)
WD_INST_REORDER(0, true)
WD_INST_REORDER(1, true)
WD_INST_REORDER(2, true)
#undef WD_INST_REORDER

#define WD_INST_FORCE(M, E, H, B)                                                               \
    template __global__ void k_force<M, E, H, B>(const Params*, Ctl*, const float4*, const float4*, \
                                                 const float4*, const float4*, const uint32_t*,     \
                                                 float4*, float4*, uint32_t*, float*, BlockPart*,   \
                                                 double*, HistoryIO, int, BinIO);
#define WD_INST_FORCE_B(M, E, H) WD_INST_FORCE(M, E, H, 64) WD_INST_FORCE(M, E, H, 128) WD_INST_FORCE(M, E, H, 256)
#define WD_INST_FORCE_H(M, E) WD_INST_FORCE_B(M, E, false) WD_INST_FORCE_B(M, E, true)
WD_INST_FORCE_H(kModeSIE, false)
WD_INST_FORCE_H(kModeSIE, true)
WD_INST_FORCE_H(kModeBEP, false)
WD_INST_FORCE_H(kModeBEP, true)
WD_INST_FORCE(kModeRHS, false, false, 256)
WD_INST_FORCE(kModeRHS, false, true, 256)
WD_INST_FORCE(kModeRHS, true, false, 256)
WD_INST_FORCE(kModeRHS, true, true, 256)
#define WD_INST_FORCE_Q(M, E, H, L)                                                              \
    template __global__ void k_force_q<M, E, H, L>(const Params*, Ctl*, const float4*, const float4*, \
                                                   const float4*, const float4*, const uint32_t*,     \
                                                   float4*, float4*, uint32_t*, BlockPart*, double*,  \
                                                   HistoryIO, int, BinIO);
#define WD_INST_FORCE_QL(M, E, H) WD_INST_FORCE_Q(M, E, H, 1) WD_INST_FORCE_Q(M, E, H, 2) WD_INST_FORCE_Q(M, E, H, 4)
WD_INST_FORCE_QL(kModeSIE, false, false)
WD_INST_FORCE_QL(kModeSIE, false, true)
WD_INST_FORCE_QL(kModeSIE, true, false)
WD_INST_FORCE_QL(kModeSIE, true, true)
WD_INST_FORCE_QL(kModeBEP, false, false)
WD_INST_FORCE_QL(kModeBEP, false, true)
WD_INST_FORCE_QL(kModeBEP, true, false)
WD_INST_FORCE_QL(kModeBEP, true, true)
#undef WD_INST_FORCE_QL
#undef WD_INST_FORCE_Q
#undef WD_INST_FORCE_H
#undef WD_INST_FORCE_B
#undef WD_INST_FORCE

// ---------------------------------------------------------------------------
// Inspection: cell order and reference-format neighbor rows from S.
__global__ __launch_bounds__(kBlock) void k_export_order(const Params* __restrict__ Pp,
                                                         const float4* __restrict__ Sp,
                                                         int* __restrict__ ids,
                                                         uint32_t* __restrict__ keys, int n) {
    const Params& P = *Pp;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        const float4 q = Sp[p];
        int cx, cy, cz;
        grain_cell(P, q.x, q.y, q.z, cx, cy, cz);
        ids[p] = __float_as_int(q.w);
        keys[p] = (static_cast<uint32_t>(cx) * static_cast<uint32_t>(P.gy) + static_cast<uint32_t>(cy)) *
                          static_cast<uint32_t>(P.gz) + static_cast<uint32_t>(cz);
    }
}

__global__ __launch_bounds__(kBlock) void k_neighbor_rows(const Params* __restrict__ Pp,
                                                          const float4* __restrict__ Sv,
                                                          const float4* __restrict__ Sp,
                                                          const uint32_t* __restrict__ start,
                                                          int* __restrict__ out, int n) {
    const Params& P = *Pp;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 pi = Sp[p];
    const size_t row_len = static_cast<size_t>(27) * P.kcap;
    int* row = out + static_cast<size_t>(__float_as_int(pi.w)) * row_len;
    const MemberCells g = member_cells(P, P.members == 1 ? 0 : grain_member(Sv[p].w));
    int cx, cy, cz;
    grain_cell(P, pi.x, pi.y, pi.z, cx, cy, cz);
    const uint32_t kcap = static_cast<uint32_t>(P.kcap);
    size_t w = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                const uint32_t c = g.index(cx + dx, cy + dy, cz + dz);
                const uint32_t b = start[c];
                const uint32_t e = min(start[c + 1], b + kcap);
                for (uint32_t q = b; q < e; ++q) {
                    // A grain dropped from its own full cell can see 27 * k
                    // candidates; the row keeps the first 27 * k - 1 so the
                    // terminator stays inside it.
                    if (q != static_cast<uint32_t>(p) && w + 1 < row_len) row[w++] = __float_as_int(Sp[q].w);
                }
            }
        }
    }
    row[w] = -1;
}

// Contact-history lists of one generation (indexed by T position),
// transposed to per-grain rows in original order: partner[id * kHist + k],
// xi[(id * kHist + k) * 3 + a], count[id].
__global__ __launch_bounds__(kBlock) void k_export_history(const float4* __restrict__ Tp,
                                                           const float4* __restrict__ hist,
                                                           const int* __restrict__ cnt,
                                                           int* __restrict__ partner,
                                                           float* __restrict__ xi,
                                                           int* __restrict__ count, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // T position
    if (i >= n) return;
    const int id = __float_as_int(Tp[i].w);
    const int h = cnt[i];
    count[id] = h;
    for (int k = 0; k < kHist; ++k) {
        const size_t o = static_cast<size_t>(id) * kHist + k;
        if (k < h) {
            const float4 e = hist[static_cast<size_t>(k) * n + i];
            partner[o] = __float_as_int(e.w);
            xi[3 * o] = e.x; xi[3 * o + 1] = e.y; xi[3 * o + 2] = e.z;
        } else {
            partner[o] = 0;
            xi[3 * o] = xi[3 * o + 1] = xi[3 * o + 2] = 0.f;
        }
    }
}

// ---------------------------------------------------------------------------
// Cloud-in-cell deposit (src/density_grid_kernel.cu:25-83), made
// order-independent: every FP32 contribution (m*w, m*v_a*w, formed as in the
// reference) is rounded to a 64-bit fixed-point integer whose scale 2^S is
// chosen from the largest per-grain magnitude so no node sum can overflow;
// integer sums do not depend on the order of the atomics, so the grid is
// bitwise reproducible.
// log_next (null unless the deposit goes to the density log): the number of
// log slots taken, which this deposit's k_cic_convert reads back.
__global__ __launch_bounds__(kBlock) void k_cic_zero(unsigned long long* __restrict__ acc,
                                                     size_t count, unsigned* __restrict__ log_next) {
    if (log_next && blockIdx.x == 0 && threadIdx.x == 0) *log_next = *log_next + 1u;
    for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
         k += static_cast<size_t>(gridDim.x) * blockDim.x) {
        acc[k] = 0ull;
    }
}

__global__ __launch_bounds__(kBlock) void k_cic_scale(const Params* __restrict__ Pp,
                                                      BlockPart* __restrict__ parts,
                                                      const float4* __restrict__ Tp,
                                                      const float4* __restrict__ Tv, int n) {
    const Params& P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    int member = -1;
    if (i < n) {
        const float4 v = Tv[i];
        member = member_of(P, v.w);
        hi[0] = static_cast<int>(cic_grain_bits(P, Tp[i], v));
    }
    store_block_part(parts, P.members > 1 ? P.pmem : nullptr, blockIdx.x, member, lo, hi, 0u);
}

// Fixed-point exponent S for n grains of largest magnitude M: n*M*2^S < 2^62.
__device__ __forceinline__ int cic_shift(float M, int n) {
    const double nm = static_cast<double>(n) * static_cast<double>(M);
    if (!(nm > 0.0)) return 0;
    return 61 - ilogb(nm);
}

__device__ __forceinline__ void axis_weights(float p, float lo, float hi, int size, int& i0,
                                             int& i1, float& w0, float& w1) {
    if (size <= 1) {
        i0 = 0; i1 = 0; w0 = 1.f; w1 = 0.f;
        return;
    }
    const float h = (hi - lo) / static_cast<float>(size - 1);
    const float u = (p - lo) / h;
    const float base = floorf(u);
    const float frac = u - base;
    const int b = sat_f2i(base);
    i0 = min(size - 1, max(0, b));
    i1 = (b == 2147483647) ? (size - 1) : min(size - 1, max(0, b + 1));
    w0 = 1.f - frac;
    w1 = frac;
}

// u64 LDS accumulators of one k_cic_deposit block (nodes of its box x 4 channels).
constexpr int kCicWindow = 2048;

// Adds each grain's mass and momentum, CIC-weighted, to the 8 surrounding
// nodes as fixed-point integers (so the sums do not depend on order).  Grains
// are in cell order, so a block's grains cover a small box of nodes: the block
// accumulates there in LDS and adds each touched node to the grid once.  A
// block whose box does not fit, or that spans two ensemble members, adds to
// the grid directly.
__global__ __launch_bounds__(kBlock) void k_cic_deposit(const Params* __restrict__ Pp,
                                                        const float4* __restrict__ Tp,
                                                        const float4* __restrict__ Tv,
                                                        unsigned long long* __restrict__ acc,
                                                        int n) {
    const Params& P = *Pp;
    __shared__ unsigned long long s_acc[kCicWindow];
    __shared__ int s_red[kBlock / 64][8];
    const int i = logical_block() * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t nodes = static_cast<size_t>(P.dgx) * P.dgy * P.dgz;
    bool active = false;
    int member = 0;
    double scale = 0.0;
    float m = 0.f, mom[3] = {0.f, 0.f, 0.f};
    int ix[2] = {0, 0}, iy[2] = {0, 0}, iz[2] = {0, 0};
    float wx[2] = {0.f, 0.f}, wy[2] = {0.f, 0.f}, wz[2] = {0.f, 0.f};
    if (i < n) {
        const float4 p = Tp[i], v = Tv[i];
        member = member_of(P, v.w);
        const unsigned sb = P.mscale[member];
        m = mat_mass(P, member, grain_mat(v.w));
        // a non-finite member grid is reported as NaN; massless grains add nothing
        if (sb < 0x7F800000u && m != 0.f) {
            active = true;
            scale = ldexp(1.0, cic_shift(__uint_as_float(sb), P.mgoff[member + 1] - P.mgoff[member]));
            axis_weights(p.x, P.lo[0], P.hi[0], P.dgx, ix[0], ix[1], wx[0], wx[1]);
            axis_weights(p.y, P.lo[1], P.hi[1], P.dgy, iy[0], iy[1], wy[0], wy[1]);
            axis_weights(p.z, P.lo[2], P.hi[2], P.dgz, iz[0], iz[1], wz[0], wz[1]);
            mom[0] = m * v.x;
            mom[1] = m * v.y;
            mom[2] = m * v.z;
        }
    }
    // The block's node box and member range (active grains only).
    int r[8] = {active ? ix[0] : kEmptyMin, active ? iy[0] : kEmptyMin, active ? iz[0] : kEmptyMin,
                active ? ix[1] : kEmptyMax, active ? iy[1] : kEmptyMax, active ? iz[1] : kEmptyMax,
                active ? member : kEmptyMin, active ? member : kEmptyMax};
#pragma unroll
    for (int k = 0; k < 8; ++k) r[k] = (k < 3 || k == 6) ? wave_min_dpp(r[k]) : wave_max_dpp(r[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) s_red[wave][k] = r[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            r[k] = (k < 3 || k == 6) ? min(r[k], s_red[w][k]) : max(r[k], s_red[w][k]);
        }
    }
    if (r[6] > r[7]) return;  // no active grain in the block
    const int bx = r[3] - r[0] + 1, by = r[4] - r[1] + 1, bz = r[5] - r[2] + 1;
    const long long vol = static_cast<long long>(bx) * by * bz;
    auto value = [&](float q, float w) __attribute__((always_inline)) {
        return static_cast<unsigned long long>(llrint(static_cast<double>(q * w) * scale));
    };
    if (r[6] != r[7] || 4 * vol > kCicWindow) {
        if (!active) return;
        unsigned long long* g = acc + static_cast<size_t>(member) * 4 * nodes;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const float w = wx[a] * wy[b] * wz[c];
                    if (w == 0.f) continue;
                    const size_t node = (static_cast<size_t>(ix[a]) * P.dgy + iy[b]) * P.dgz + iz[c];
                    atomicAdd(&g[node], value(m, w));
#pragma unroll
                    for (int k = 0; k < 3; ++k) atomicAdd(&g[(k + 1) * nodes + node], value(mom[k], w));
                }
            }
        }
        return;
    }
    const int nv = static_cast<int>(vol);
    for (int k = threadIdx.x; k < 4 * nv; k += blockDim.x) s_acc[k] = 0ull;
    __syncthreads();
    if (active) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const float w = wx[a] * wy[b] * wz[c];
                    if (w == 0.f) continue;
                    const int ln = ((ix[a] - r[0]) * by + (iy[b] - r[1])) * bz + (iz[c] - r[2]);
                    atomicAdd(&s_acc[ln], value(m, w));
#pragma unroll
                    for (int k = 0; k < 3; ++k) atomicAdd(&s_acc[(k + 1) * nv + ln], value(mom[k], w));
                }
            }
        }
    }
    __syncthreads();
    unsigned long long* g = acc + static_cast<size_t>(r[6]) * 4 * nodes;
    for (int k = threadIdx.x; k < 4 * nv; k += blockDim.x) {
        const unsigned long long v = s_acc[k];
        if (v == 0ull) continue;
        const int ch = k / nv, ln = k - ch * nv;
        const int lx = ln / (by * bz), rem = ln - lx * by * bz, ly = rem / bz, lz = rem - ly * bz;
        const size_t node = (static_cast<size_t>(r[0] + lx) * P.dgy + (r[1] + ly)) * P.dgz + (r[2] + lz);
        atomicAdd(&g[static_cast<size_t>(ch) * nodes + node], v);
    }
}

// log_next (null for a plain deposit into `out`): the deposit goes to log
// slot *log_next - 1, `count` floats apart.
__global__ __launch_bounds__(kBlock) void k_cic_convert(const Params* __restrict__ Pp,
                                                        const unsigned long long* __restrict__ acc,
                                                        float* __restrict__ out, size_t count,
                                                        const unsigned* __restrict__ log_next) {
    const Params& P = *Pp;
    if (log_next) out += static_cast<size_t>(*log_next - 1u) * count;
    const size_t per = static_cast<size_t>(4) * P.dgx * P.dgy * P.dgz;
    for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
         k += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int member = static_cast<int>(k / per);
        const unsigned sb = P.mscale[member];
        const bool bad = sb >= 0x7F800000u;
        const int nm = P.mgoff[member + 1] - P.mgoff[member];
        const double inv = bad ? 0.0 : ldexp(1.0, -cic_shift(__uint_as_float(sb), max(nm, 1)));
        out[k] = bad ? __int_as_float(0x7FC00000)
                     : static_cast<float>(static_cast<double>(static_cast<long long>(acc[k])) * inv);
    }
}

}  // namespace wd
