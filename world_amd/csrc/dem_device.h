// Per-grain physics shared by every gfx950 kernel of the engine.
//
// Each function restates one piece of the reference's per-thread arithmetic
// in the same FP32 operation order (the build compiles with
// -ffp-contract=off and correctly rounded division and square root, so the
// results are reproducible bit for bit by a CPU evaluation of the same order):
//   cell_axis          src/find_neighbors_kernel.cu:11-22
//   pair_contact       src/compute_rhs_kernel.cu:145-197
//   viscous_friction   src/compute_rhs_kernel.cu:66-78
//   contact_energy     src/energy_kernel.cu:22-28
// The restitution-to-damping constants (src/compute_rhs_kernel.cu:41-47)
// depend only on material pairs, so they are evaluated once on the host and
// read here from a table.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/world_amd.h"

namespace wd {

constexpr int kMaxMat = WD_MAX_MATERIALS;
constexpr int kHist = 24;  // contact-history slots per grain (walls + pairs)
constexpr int kMatTab = 2 * kMaxMat + kMaxMat * kMaxMat;  // one member's material tables

// Device-resident solver constants (one copy per handle, read through a
// pointer so that dt can change without re-capturing the step graph).
struct BlockMembers;

struct Params {
    int n;
    int gx, gy, gz;        // collision grid
    int kcap;              // particles_per_cell
    int nmat;
    float lo[3], hi[3];    // domain (cell mapping, CIC)
    float wlo[3], whi[3];  // walls: x_min, y_min, floor / x_max, y_max, ceiling
    float gravity, radius, max_force;
    float kneg;            // -max_force / radius (or -normal_stiffness)
    float spring_k;        // k of the energy's quadratic branch
    float clamp_delta;     // overlap where the force clamp engages
    float cut2;            // squared centre distance above which a pair cannot touch
    float touch2;          // smallest FP32 squared distance with (sqrt - r) - r >= 0: a pair
                           // touches exactly when its squared distance is below it
    float mu;
    float dt;
    float kt;              // tangential spring stiffness (history model)
    float inv_kt;          // 1 / kt in FP32: a slip resets the spring by multiplying with it
    int history;           // 1: tangential spring with history, 0: viscous (reference)
    int dgx, dgy, dgz;     // density grid
    // Ensemble members (1 for a single simulation).  Member m owns grains
    // with ids [mgoff[m], mgoff[m+1]) and the cell range
    // [mcell[m], mcell[m+1]) of the concatenated member sub-grids; mbox and
    // mcell are rewritten by k_finalize for every new state.
    int members;
    const int* mgoff;      // [members + 1]
    int* mbox;             // [members][6] occupied cells of the current state
    uint32_t* mcell;       // [members + 1]
    const float* mmu;      // [members] friction coefficient
    unsigned* mscale;      // [members] float bits of the largest CIC deposit magnitude
    const float* mtab;     // [members][kMatTab] mass, c_wall, c_pair per member (null: the
                           // tables below serve every grain)
    unsigned long long* movf;  // [members] over-capacity cells, summed over passes
    BlockMembers* pmem;    // member part of the block partials (written when members > 1)
    // Fused step of a lone simulation: the grid the force pass counts into
    // (pending: the last reduced box widened by one cell per side, written by
    // the reduction of the partials) and the grid the scan and reorder bin
    // with (next: the pending grid as the force pass found it; the reorder
    // makes it current).  Layouts as mbox / mcell.
    int* pbox;
    uint32_t* pcell;
    int* nbox;
    uint32_t* ncell;
    float mass[kMaxMat];
    float c_wall[kMaxMat];
    float c_pair[kMaxMat * kMaxMat];
};

// Saturating float -> int truncation (NaN -> 0), the conversion the
// reference's static_cast<int> performs on the GPU.
__device__ __forceinline__ int sat_f2i(float t) {
    if (!(t == t)) return 0;
    if (t >= 2147483648.0f) return 2147483647;
    if (t <= -2147483648.0f) return (-2147483647 - 1);
    return static_cast<int>(t);
}

__device__ __forceinline__ int cell_axis(float x, float lo, float hi, int g) {
    const float t = (x - lo) / (hi - lo) * static_cast<float>(g);
    const int c = sat_f2i(t);
    return min(g - 1, max(0, c));
}

// Packed cell key: cx in bits 21..31, cy in 10..20, cz in 0..9.
__device__ __forceinline__ uint32_t pack_cell(int cx, int cy, int cz) {
    return (static_cast<uint32_t>(cx) << 21) | (static_cast<uint32_t>(cy) << 10) |
           static_cast<uint32_t>(cz);
}
__device__ __forceinline__ void unpack_cell(uint32_t k, int& cx, int& cy, int& cz) {
    cx = static_cast<int>(k >> 21);
    cy = static_cast<int>((k >> 10) & 0x7FFu);
    cz = static_cast<int>(k & 0x3FFu);
}

__device__ __forceinline__ void grain_cell(const Params& P, float x, float y, float z,
                                           int& cx, int& cy, int& cz) {
    cx = cell_axis(x, P.lo[0], P.hi[0], P.gx);
    cy = cell_axis(y, P.lo[1], P.hi[1], P.gy);
    cz = cell_axis(z, P.lo[2], P.hi[2], P.gz);
}

// Material constants of a grain of `member` (members of a batch may differ
// in grain mass and restitution).
__device__ __forceinline__ float mat_mass(const Params& P, int member, int mat) {
    return P.mtab ? P.mtab[member * kMatTab + mat] : P.mass[mat];
}
__device__ __forceinline__ float mat_cwall(const Params& P, int member, int mat) {
    return P.mtab ? P.mtab[member * kMatTab + kMaxMat + mat] : P.c_wall[mat];
}
__device__ __forceinline__ float mat_cpair(const Params& P, int member, int a, int b) {
    return P.mtab ? P.mtab[member * kMatTab + 2 * kMaxMat + a * kMaxMat + b]
                  : P.c_pair[a * kMaxMat + b];
}

// Member index and material id of a grain, packed in its velocity's .w.
__device__ __forceinline__ int grain_member(float w) { return __float_as_int(w) >> 4; }
__device__ __forceinline__ int grain_mat(float w) { return __float_as_int(w) & (kMaxMat - 1); }

// Sub-grid covering the occupied cells' bounding box with one empty cell of
// padding on every side, so every 27-cell stencil of an occupied cell is
// addressable without bounds checks.
struct SubGrid {
    int ox, oy, oz;
    int bx, by, bz;
    __device__ __forceinline__ int index(int cx, int cy, int cz) const {
        return ((cx - ox) * by + (cy - oy)) * bz + (cz - oz);
    }
    __device__ __forceinline__ int cells() const { return bx * by * bz; }
};

__device__ __forceinline__ SubGrid make_subgrid(const int* bbox) {
    SubGrid s;
    s.ox = bbox[0] - 1;
    s.oy = bbox[1] - 1;
    s.oz = bbox[2] - 1;
    s.bx = bbox[3] - bbox[0] + 3;
    s.by = bbox[4] - bbox[1] + 3;
    s.bz = bbox[5] - bbox[2] + 3;
    return s;
}

// The cells of member m: its padded sub-grid, placed at `base` in the
// concatenated cell arrays.
struct MemberCells {
    SubGrid g;
    uint32_t base;
    __device__ __forceinline__ uint32_t index(int cx, int cy, int cz) const {
        return base + static_cast<uint32_t>(g.index(cx, cy, cz));
    }
};

__device__ __forceinline__ MemberCells member_cells(const Params& P, int m) {
    MemberCells c;
    c.g = make_subgrid(P.mbox + 6 * m);
    c.base = P.mcell[m];
    return c;
}

// Member of a grain given its velocity .w (uniformly 0 for one member).
__device__ __forceinline__ int member_of(const Params& P, float w) {
    return P.members == 1 ? 0 : grain_member(w);
}

// Reference viscous Coulomb friction: oppose the tangential velocity,
// magnitude min(limit, c * |v_t|); no force below |v_t| = 1e-8.
__device__ __forceinline__ void viscous_friction(float f[3], const float vt[3], float limit,
                                                 float c) {
    float s = 0.f;
    s += vt[0] * vt[0];
    s += vt[1] * vt[1];
    s += vt[2] * vt[2];
    const float n = sqrtf(s);
    if (n < 1e-8f) return;
    const float mag = fminf(limit, c * n);
    f[0] -= mag * vt[0] / n;
    f[1] -= mag * vt[1] / n;
    f[2] -= mag * vt[2] / n;
}

// Centre distance |x_i - x_j| with the reference's summation order; also
// the separation vector.
__device__ __forceinline__ float pair_norm(const float xi[3], float xj0, float xj1, float xj2,
                                           float& d0, float& d1, float& d2) {
    d0 = xi[0] - xj0;
    d1 = xi[1] - xj1;
    d2 = xi[2] - xj2;
    float s = 0.f;
    s += d0 * d0;
    s += d1 * d1;
    s += d2 * d2;
    return sqrtf(s);
}

// One neighbor j (cell-sorted slot q) acting on grain i, given the
// separation d and its norm.  A pair beyond contact adds a zero spring force,
// as in the reference.
__device__ __forceinline__ void pair_contact(const Params& P, float mu, int member, const float vi[3],
                                             int mat_i, float d0, float d1, float d2, float norm,
                                             const float4* __restrict__ vel, int q, float f[3]) {
    const float dist = norm - P.radius - P.radius;
    const float n0 = d0 / norm, n1 = d1 / norm, n2 = d2 / norm;
    const float fp = fmaxf(0.f, fminf(P.max_force, P.kneg * dist));
    f[0] += fp * n0;
    f[1] += fp * n1;
    f[2] += fp * n2;
    if (dist < 0.f) {
        const float4 vj = vel[q];
        const int mat_j = grain_mat(vj.w);
        const float c = mat_cpair(P, member, mat_i, mat_j);
        const float r0 = vi[0] - vj.x, r1 = vi[1] - vj.y, r2 = vi[2] - vj.z;
        float vn = 0.f;
        vn += r0 * n0;
        vn += r1 * n1;
        vn += r2 * n2;
        float vt[3];
        f[0] -= c * vn * n0;
        vt[0] = r0 - vn * n0;
        f[1] -= c * vn * n1;
        vt[1] = r1 - vn * n1;
        f[2] -= c * vn * n2;
        vt[2] = r2 - vn * n2;
        viscous_friction(f, vt, mu * fp, c);
    }
}

// Tangential spring with history (extension; the reference's friction is
// stateless, src/compute_rhs_kernel.cu:50-65).  The stored displacement xi is
// rotated into the current tangent plane, advanced by the tangential relative
// displacement v_t dt, and turned into F_t = -k_t xi - c v_t; beyond the
// Coulomb limit mu * fp the force is scaled back to the limit and the spring
// reset to match it (slip).  Returns 1 on slip.  n: unit normal into grain i;
// vr: relative velocity of i; vn = vr . n.
__device__ __forceinline__ int spring_friction(const Params& P, float mu, float f[3],
                                               const float n[3], const float vr[3], float vn,
                                               float fp, float c, float xi[3]) {
    float vt[3];
    #pragma unroll
    for (int a = 0; a < 3; ++a) vt[a] = vr[a] - vn * n[a];
    float xn = 0.f;
    xn += xi[0] * n[0];
    xn += xi[1] * n[1];
    xn += xi[2] * n[2];
    #pragma unroll
    for (int a = 0; a < 3; ++a) xi[a] = xi[a] - xn * n[a];
    #pragma unroll
    for (int a = 0; a < 3; ++a) xi[a] = xi[a] + vt[a] * P.dt;
    float ft[3];
    #pragma unroll
    for (int a = 0; a < 3; ++a) ft[a] = -(P.kt * xi[a]) - c * vt[a];
    float s = 0.f;
    s += ft[0] * ft[0];
    s += ft[1] * ft[1];
    s += ft[2] * ft[2];
    const float fm = sqrtf(s);
    const float lim = mu * fp;
    int slip = 0;
    if (fm > lim) {
        const float scale = lim / fm;
        #pragma unroll
        for (int a = 0; a < 3; ++a) {
            ft[a] = ft[a] * scale;
            xi[a] = -(ft[a] + c * vt[a]) * P.inv_kt;
        }
        slip = 1;
    }
    #pragma unroll
    for (int a = 0; a < 3; ++a) f[a] += ft[a];
    return slip;
}

// Partner code of wall (axis a, side 0 lower / 1 upper) in a history list.
__device__ __forceinline__ int wall_partner(int a, int side) { return -(1 + 2 * a + side); }

// Elastic energy of one contact of overlap delta: quadratic up to the clamp,
// linear beyond it.
__device__ __forceinline__ float contact_energy(const Params& P, float delta) {
    if (delta <= 0.f) return 0.f;
    if (delta <= P.clamp_delta) return 0.5f * P.spring_k * delta * delta;
    return 0.5f * P.max_force * P.clamp_delta + P.max_force * (delta - P.clamp_delta);
}

// Kinetic + gravitational + wall elastic energy of one grain (pair terms are
// added by the caller in neighbor order).
__device__ __forceinline__ float grain_energy_self(const Params& P, const float x[3],
                                                   const float v[3], float m) {
    float total = 0.f;
    float s = 0.f;
    s += v[0] * v[0];
    s += v[1] * v[1];
    s += v[2] * v[2];
    total += 0.5f * m * s;
    total += m * P.gravity * (x[2] - P.wlo[2]);
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
        total += contact_energy(P, P.radius - (x[a] - P.wlo[a]));
        total += contact_energy(P, P.radius - (P.whi[a] - x[a]));
    }
    return total;
}

}  // namespace wd
