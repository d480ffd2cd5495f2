// gfx950 kernels of the discrete adjoint (extension; the reference has no
// differentiation).
//
// The adjoint differentiates the semi-implicit Euler step with the
// history-carrying tangential spring (dem_device.h: pair_contact /
// spring_friction arithmetic) along a recorded forward trajectory.  The
// contact set of every step is the one the forward used (the pass re-derives
// it bit for bit from the restored state), and so are the clamp and
// stick/slip branches: the adjoint is the exact derivative of that frozen
// piecewise-smooth step.
//
// Per step s (state s -> s+1), with lambda = dL/d(state s+1) given:
//   k_adj_contacts  grains in cell order (one or four lanes each): replays
//                   every contact of the forward pass, applies the contact's vector-Jacobian
//                   product to (dL/dF_i, dL/dxi_new), keeps the grain's own
//                   share, writes dL/dxi_old into the previous spring
//                   generation, and stores each pair contact's (d, v_r)
//                   cotangents for the partner;
//   k_adj_gather    per grain: adds the partners' shares (the pair
//                   law depends on x_i - x_j and v_i - v_j only, so the
//                   partner's share is the negative of the stored one), in a
//                   fixed order of the grain's contacts: no atomics, deterministic.
// CIC cotangents enter through k_adj_seed.  Everything is accumulated in FP32
// except the parameter gradients (per-grain double sums over the window).
//
// Layout: within a step the adjoint arrays follow that step's cell order (a
// lane's own entries are contiguous with its neighbours'): dL/d(state s+1) at
// the step's cell-sorted positions, the cotangents of the springs the step
// wrote at those positions and of the springs it read at the recorded spring
// index, pair contacts by partner position.  k_adj_gather hands the result
// to the previous step's order through the recorded spring index (nearly the
// identity between steps); a checkpoint is in original order, and
// k_adj_to_cells maps into a segment's last cell order.  The parameter
// gradients stay in original order.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace wd {

constexpr int kAdjContacts = 32;  // pair contacts per grain kept for the partner gather

enum AdjError { kAdjErrContacts = 2 };

namespace {

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    float s = 0.f;
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

// Reverse of one contact of the history model: the normal spring-dashpot
// (f += fp n - c vn n) followed by spring_friction, recomputed in the same
// FP32 order so that the slip decision is the forward's.  Inputs: unit
// normal n, relative velocity vr, vn = vr . n, fp and whether fp is on its
// linear branch (fp = kneg * dist), the damping c, the stored spring xi, and
// the cotangents gF (of the grain's force) and gX (of the updated spring).
// Outputs: cotangents of n, dist, vr, the stored spring, and the parameters
// k_n and mu (added).
__device__ __forceinline__ void contact_vjp(const Params& P, float mu, const float n[3],
                                            const float vr[3], float vn, float fp, bool fp_linear,
                                            float dist, float c, const float xi[3],
                                            const float gF[3], const float gX[3], float g_n[3],
                                            float& g_dist, float g_vr[3], float g_xi[3],
                                            float& g_k, float& g_mu) {
    // forward (spring_friction)
    float vt[3], xi1[3], xi2[3], ft0[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) vt[a] = vr[a] - vn * n[a];
    float xn = 0.f;
    xn += xi[0] * n[0];
    xn += xi[1] * n[1];
    xn += xi[2] * n[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) xi1[a] = xi[a] - xn * n[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) xi2[a] = xi1[a] + vt[a] * P.dt;
#pragma unroll
    for (int a = 0; a < 3; ++a) ft0[a] = -(P.kt * xi2[a]) - c * vt[a];
    float s = 0.f;
    s += ft0[0] * ft0[0];
    s += ft0[1] * ft0[1];
    s += ft0[2] * ft0[2];
    const float fm = sqrtf(s);
    const float lim = mu * fp;
    const bool slip = fm > lim;

    // reverse
    float g_ft0[3], g_xi2[3], g_vt[3] = {0.f, 0.f, 0.f};
    float g_c = 0.f, g_kt = 0.f, g_fp = 0.f;
    if (slip) {
        const float scale = lim / fm;
        float ft[3], g_ft[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            ft[a] = ft0[a] * scale;
            // xi_new = -(ft + c vt) / kt
            g_ft[a] = gF[a] - gX[a] / P.kt;
            g_vt[a] = -c * gX[a] / P.kt;
        }
        float w[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) w[a] = ft[a] + c * vt[a];
        g_c -= dot3(gX, vt) / P.kt;
        g_kt += dot3(gX, w) / (P.kt * P.kt);
        // ft = ft0 * lim / |ft0|
        const float g_scale = dot3(g_ft, ft0);
        const float g_lim = g_scale / fm;
        const float g_fm = -g_scale * lim / (fm * fm);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            g_ft0[a] = scale * g_ft[a] + g_fm * ft0[a] / fm;
            g_xi2[a] = 0.f;
        }
        g_mu += g_lim * fp;
        g_fp += g_lim * mu;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            g_ft0[a] = gF[a];
            g_xi2[a] = gX[a];
        }
    }
    // ft0 = -(kt xi2) - c vt
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g_xi2[a] -= P.kt * g_ft0[a];
        g_vt[a] -= c * g_ft0[a];
    }
    g_kt -= dot3(xi2, g_ft0);
    g_c -= dot3(vt, g_ft0);
    // xi2 = xi1 + vt dt
#pragma unroll
    for (int a = 0; a < 3; ++a) g_vt[a] += P.dt * g_xi2[a];
    // xi1 = xi - (xi . n) n
    const float g_xn = -dot3(g_xi2, n);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g_xi[a] = g_xi2[a] + g_xn * n[a];
        g_n[a] = -xn * g_xi2[a] + g_xn * xi[a];
    }
    // vt = vr - vn n
    float g_vn = -dot3(g_vt, n);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g_vr[a] = g_vt[a];
        g_n[a] -= vn * g_vt[a];
    }
    // f += fp n - c vn n
    const float gFn = dot3(gF, n);
    g_fp += gFn;
    g_c -= vn * gFn;
    g_vn -= c * gFn;
#pragma unroll
    for (int a = 0; a < 3; ++a) g_n[a] += fp * gF[a] - c * vn * gF[a];
    // vn = vr . n
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        g_vr[a] += g_vn * n[a];
        g_n[a] += g_vn * vr[a];
    }
    // fp = clamp(kneg * dist); kneg = -k_n, kt = ratio k_n, c ~ sqrt(k_n)
    const float k = -P.kneg;
    float g_kneg = 0.f;
    g_dist = 0.f;
    if (fp_linear) {
        g_dist = P.kneg * g_fp;
        g_kneg = dist * g_fp;
    }
    g_k += -g_kneg + (P.kt / k) * g_kt + g_c * c / (2.f * k);
}

}  // namespace

// Snapshot of the current state and spring generation, by original id:
// dP[id], dV[id] = the grain's position/velocity float4s, dH / dC = its
// stored springs (slots below the count only).
__global__ __launch_bounds__(kBlock) void k_snap_save(const Ctl* __restrict__ ctl,
                                                      const float4* __restrict__ Tp,
                                                      const float4* __restrict__ Tv,
                                                      HistoryIO hist, float4* __restrict__ dP,
                                                      float4* __restrict__ dV,
                                                      float4* __restrict__ dH, int* __restrict__ dC,
                                                      int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // T position
    if (i >= n) return;
    const float4 p = Tp[i];
    const int id = __float_as_int(p.w);
    dP[id] = p;
    dV[id] = Tv[i];
    const int par = ctl->hist_parity;
    const int c = hist.cnt[par] ? hist.cnt[par][i] : 0;  // no springs under the viscous model
    dC[id] = c;
    for (int t = 0; t < c; ++t) dH[static_cast<size_t>(t) * n + id] = hist.buf[par][static_cast<size_t>(t) * n + i];
}

// Restores a snapshot: T = original order, keys and block partials for the
// next bin, springs into the current generation (whose index is then the
// original id).
__global__ __launch_bounds__(kBlock) void k_snap_load(const Params* __restrict__ Pp,
                                                      const Ctl* __restrict__ ctl,
                                                      BlockPart* __restrict__ parts,
                                                      const float4* __restrict__ sP,
                                                      const float4* __restrict__ sV,
                                                      const float4* __restrict__ sH,
                                                      const int* __restrict__ sC,
                                                      float4* __restrict__ Tp, float4* __restrict__ Tv,
                                                      uint32_t* __restrict__ key, HistoryIO hist,
                                                      int n) {
    const Params& P = *Pp;
    const int i = logical_block() * blockDim.x + threadIdx.x;
    int lo[3] = {kEmptyMin, kEmptyMin, kEmptyMin}, hi[3] = {kEmptyMax, kEmptyMax, kEmptyMax};
    int member = -1;
    if (i < n) {
        const float4 p = sP[i], v = sV[i];
        Tp[i] = p;
        Tv[i] = v;
        member = member_of(P, v.w);
        const int par = ctl->hist_parity;
        if (hist.cnt[par]) {  // no springs under the viscous model
            const int c = sC[i];
            hist.cnt[par][i] = c;
            for (int t = 0; t < c; ++t) hist.buf[par][static_cast<size_t>(t) * n + i] = sH[static_cast<size_t>(t) * n + i];
        }
        int cx, cy, cz;
        grain_cell(P, p.x, p.y, p.z, cx, cy, cz);
        key[i] = pack_cell(cx, cy, cz);
        lo[0] = hi[0] = cx;
        lo[1] = hi[1] = cy;
        lo[2] = hi[2] = cz;
    }
    store_block_part(parts, P.members > 1 ? P.pmem : nullptr, logical_block(), member, lo, hi, 0u);
}

// dL/dx, dL/dv of the final state from the cotangent of its CIC grids
// (4 channels per member, src/density_grid_kernel.cu:43-106) plus an optional
// direct state cotangent (original order, 6 floats per grain).
__global__ __launch_bounds__(kBlock) void k_adj_seed(const Params* __restrict__ Pp,
                                                     const float4* __restrict__ Tp,
                                                     const float4* __restrict__ Tv,
                                                     const float* __restrict__ grid_cot,
                                                     const float* __restrict__ state_cot,
                                                     float4* __restrict__ lx, float4* __restrict__ lv,
                                                     int n) {
    const Params& P = *Pp;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = Tp[i], v = Tv[i];
    const int id = __float_as_int(p.w);
    float gx[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
    if (state_cot) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            gx[a] = state_cot[6 * static_cast<size_t>(id) + a];
            gv[a] = state_cot[6 * static_cast<size_t>(id) + 3 + a];
        }
    }
    const int member = member_of(P, v.w);
    const float m = mat_mass(P, member, grain_mat(v.w));
    if (grid_cot && m != 0.f) {
        const size_t nodes = static_cast<size_t>(P.dgx) * P.dgy * P.dgz;
        const float* G = grid_cot + static_cast<size_t>(member) * 4 * nodes;
        const float x[3] = {p.x, p.y, p.z}, vel[3] = {v.x, v.y, v.z};
        const int size[3] = {P.dgx, P.dgy, P.dgz};
        int idx[3][2];
        float w[3][2], dw[3][2];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (size[a] <= 1) {
                idx[a][0] = idx[a][1] = 0;
                w[a][0] = 1.f; w[a][1] = 0.f;
                dw[a][0] = dw[a][1] = 0.f;
                continue;
            }
            const float h = (P.hi[a] - P.lo[a]) / static_cast<float>(size[a] - 1);
            const float u = (x[a] - P.lo[a]) / h;
            const float base = floorf(u);
            const float frac = u - base;
            const int b = sat_f2i(base);
            idx[a][0] = min(size[a] - 1, max(0, b));
            idx[a][1] = (b == 2147483647) ? (size[a] - 1) : min(size[a] - 1, max(0, b + 1));
            w[a][0] = 1.f - frac;
            w[a][1] = frac;
            dw[a][0] = -1.f / h;
            dw[a][1] = 1.f / h;
        }
#pragma unroll
        for (int cx = 0; cx < 2; ++cx)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    const size_t node = (static_cast<size_t>(idx[0][cx]) * P.dgy + idx[1][cy]) * P.dgz + idx[2][cz];
                    const float g0 = G[node], g1 = G[nodes + node], g2 = G[2 * nodes + node],
                                g3 = G[3 * nodes + node];
                    // d(m w)/dx + sum_b d(m v_b w)/dx
                    const float q = m * (g0 + vel[0] * g1 + vel[1] * g2 + vel[2] * g3);
                    const float wxy = w[0][cx] * w[1][cy];
                    gx[0] += q * dw[0][cx] * w[1][cy] * w[2][cz];
                    gx[1] += q * w[0][cx] * dw[1][cy] * w[2][cz];
                    gx[2] += q * wxy * dw[2][cz];
                    const float mw = m * wxy * w[2][cz];
                    gv[0] += mw * g1;
                    gv[1] += mw * g2;
                    gv[2] += mw * g3;
                }
    }
    lx[id] = make_float4(gx[0], gx[1], gx[2], 0.f);
    lv[id] = make_float4(gv[0], gv[1], gv[2], 0.f);
}

// The per-grain adjoint passes run one-wave blocks of L lanes per grain.  A
// grain's contacts fall into the stencil's three dx groups, and its share
// is summed per group: with L = 4, lanes 0, 1, 2 of the grain take dx
// groups 0, 1, 2 (lane 3 none) and the lanes' parts are added in lane
// order; with L = 1 the lane takes the groups one after another and adds
// their parts in the same order, so both give the same bits.  Four lanes
// run a grain's chains of dependent loads side by side, which pays while
// the grid is small: below kAdjLanesUpTo grains the passes run L = 4
// (8,192 grains: 33 -> 20 us for the contacts, 15 -> 9 us for the gather),
// above it L = 1 (a 171,425-grain batch has 2.6 one-lane waves per SIMD
// already, and four lanes made its backward step 11% slower).
constexpr int kAdjLanes = 4;
constexpr int kAdjLanesUpTo = 65536;
constexpr int kAdjBlock = 64;

// Pass A of the adjoint of one SIE step (see the file comment).  Grains in
// cell order (Sp/Sv, start: the bin of state s); h_in = the springs the
// forward step read; lxi_next[k][id] = dL/d(spring slot k written by the
// step); lxi_prev[t][id] receives dL/d(spring slot t read by the step).  The
// cotangent lists are indexed by original id.
// A dx group's touching candidates are counted before they are evaluated,
// which fixes every contact's spring slot and pair index (walls first, then
// the groups in order).  The share of the grain's force cotangent and of its
// parameters is summed per dx group (walls and the integrator's terms in
// group 0's part), and the parts are added in group order.  Every slot of
// the springs the step read is written once: a matched contact's
// dL/dxi_old, or zero.
template <int L>
__global__ __launch_bounds__(kAdjBlock) void k_adj_contacts(
        const Params* __restrict__ Pp, Ctl* ctl, const float4* __restrict__ Sp,
        const float4* __restrict__ Sv, const uint32_t* __restrict__ start,
        const int* __restrict__ hidx, const float4* __restrict__ h_in,
        const int* __restrict__ h_cnt, const int* __restrict__ mbox,
        const uint32_t* __restrict__ mcell, const float4* __restrict__ lx,
        const float4* __restrict__ lv, const float4* __restrict__ lxi_next,
        float4* __restrict__ lxi_prev, float4* __restrict__ own_x, float4* __restrict__ own_v,
        float4* __restrict__ cg_d, float4* __restrict__ cg_v, int* __restrict__ cg_cnt,
        int* __restrict__ cg_p, double2* __restrict__ gparam, int n) {
    static_assert(L == 1 || L == kAdjLanes, "adjoint pass variants");
    constexpr int G = kAdjBlock / L;  // grains per block
    const Params& P = *Pp;
    __shared__ uint32_t s_list[kContactList][kAdjBlock];
    const int lane = threadIdx.x;
    const int sub = lane & (L - 1);
    const int first = logical_block() * G;
    if (first >= n) return;  // the whole wave is past the last grain
    // Lanes past the last grain mirror it and write nothing.
    const bool live = first + lane / L < n;
    const int p = live ? first + lane / L : n - 1;
    const float4 pi = Sp[p];
    const float4 vi4 = Sv[p];
    const float x[3] = {pi.x, pi.y, pi.z};
    const float v[3] = {vi4.x, vi4.y, vi4.z};
    const int id = __float_as_int(pi.w);
    const int mat = grain_mat(vi4.w);
    const int member = member_of(P, vi4.w);
    const float mu = P.members == 1 ? P.mu : P.mmu[member];
    const float m = mat_mass(P, member, mat);

    // SIE: v' = v + dt F/m, x' = x + dt v'.
    const float4 lx4 = lx[p], lv4 = lv[p];
    const float lvt[3] = {lv4.x + P.dt * lx4.x, lv4.y + P.dt * lx4.y, lv4.z + P.dt * lx4.z};
    float gF[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) gF[a] = P.dt * lvt[a] / m;
    // the current dx group's part of the grain's share (group 0's starts
    // with the integrator's terms)
    float gx[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
    if (sub == 0) {
        gx[0] = lx4.x; gx[1] = lx4.y; gx[2] = lx4.z;
        gv[0] = lvt[0]; gv[1] = lvt[1]; gv[2] = lvt[2];
    }
    float gk = 0.f, gmu = 0.f;

    const int hs = hidx[p];  // the grain's index in the springs the step read
    const int h_old = h_cnt[hs];
    unsigned written = 0u;   // slots of the springs read that this lane wrote
    // Stored spring of `partner` and its slot (-1: a new contact, zero
    // spring).  Lists keep their order, so slot `guess` (the contact's slot
    // in the new list) normally holds it; otherwise the old list is searched.
    auto h_find = [&](int partner, int guess, float xi[3]) __attribute__((always_inline)) {
        const float4 hg = h_in[static_cast<size_t>(min(guess, kHist - 1)) * n + hs];
        if (guess < h_old && __float_as_int(hg.w) == partner) {
            xi[0] = hg.x; xi[1] = hg.y; xi[2] = hg.z;
            return guess;
        }
        xi[0] = xi[1] = xi[2] = 0.f;
        const int t = spring_slot(h_in, n, hs, h_old, partner);
        if (t >= 0) {
            const float4 h = h_in[static_cast<size_t>(t) * n + hs];
            xi[0] = h.x; xi[1] = h.y; xi[2] = h.z;
        }
        return t;
    };
    auto g_next = [&](int slot, float g[3]) __attribute__((always_inline)) {
        if (slot < kHist) {
            const float4 q = lxi_next[static_cast<size_t>(slot) * n + p];
            g[0] = q.x; g[1] = q.y; g[2] = q.z;
        } else {
            g[0] = g[1] = g[2] = 0.f;
        }
    };
    auto keep_xi = [&](int t, const float g_xi[3]) __attribute__((always_inline)) {
        if (t >= 0 && live) {
            lxi_prev[static_cast<size_t>(t) * n + hs] = make_float4(g_xi[0], g_xi[1], g_xi[2], 0.f);
            written |= 1u << t;
        }
    };

    // Walls, lower then upper per axis, as the forward, in group 0's part;
    // every lane counts them (they take the first spring slots).
    const float c_wall = mat_cwall(P, member, mat);
    int wc = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const float dist = side == 0 ? x[a] - P.wlo[a] - P.radius : P.whi[a] - x[a] - P.radius;
            if (dist < 0.f) {
                if (sub == 0) {
                    const float raw = P.kneg * dist;
                    const float fw = fmaxf(0.f, fminf(P.max_force, raw));
                    float nrm[3] = {0.f, 0.f, 0.f};
                    nrm[a] = side == 0 ? 1.f : -1.f;
                    float vn = 0.f;
                    vn += v[0] * nrm[0];
                    vn += v[1] * nrm[1];
                    vn += v[2] * nrm[2];
                    float xi[3], gX[3], g_n[3], g_vr[3], g_xi[3], g_dist;
                    const int t = h_find(wall_partner(a, side), wc, xi);
                    g_next(wc, gX);
                    contact_vjp(P, mu, nrm, v, vn, fw, fw == raw, dist, c_wall, xi, gF, gX, g_n, g_dist,
                                g_vr, g_xi, gk, gmu);
                    keep_xi(t, g_xi);
                    gx[a] += side == 0 ? g_dist : -g_dist;
#pragma unroll
                    for (int b = 0; b < 3; ++b) gv[b] += g_vr[b];
                }
                ++wc;
            }
        }
    }

    MemberCells mc;  // the member's cells as the step's bin laid them out
    mc.g = make_subgrid(mbox + 6 * member);
    mc.base = mcell[member];
    int cx, cy, cz;
    grain_cell(P, x[0], x[1], x[2], cx, cy, cz);
    const uint32_t kcap = static_cast<uint32_t>(P.kcap);
    // The forward pass's candidate walk (L = 4: this lane's dx group; L = 1:
    // the three groups in order, each listed entry tagged with its group in
    // bits 28-29): touching candidates (squared distance below touch2, exactly
    // the contact law's dist < 0) listed in LDS in stencil order; a lane with
    // more than kContactList of them walks again past the ones evaluated.
    int seen = 0, consumed = 0, listed = 0;
    uint32_t wtag = 0u;  // L = 1: the group being walked, in bits 28-29
    auto probe = [&](uint32_t q, float px, float py, float pz) __attribute__((always_inline)) {
        const float e0 = x[0] - px, e1 = x[1] - py, e2 = x[2] - pz;
        float s2 = 0.f;
        s2 += e0 * e0;
        s2 += e1 * e1;
        s2 += e2 * e2;
        if (s2 < P.touch2 && q != static_cast<uint32_t>(p)) {
            if (seen >= consumed && listed < kContactList) {
                s_list[listed][lane] = q | wtag;
                ++listed;
            }
            ++seen;
        }
    };
    auto column = [&](int c) __attribute__((always_inline)) {
        uint4 q4;
        __builtin_memcpy(&q4, start + mc.index(cx + c / 3 - 1, cy + c % 3 - 1, cz - 1), sizeof(uint4));
        return q4;
    };
    auto walk = [&]() __attribute__((always_inline)) {
        seen = 0;
        listed = 0;
        if (!live) return;
        if constexpr (L == 1) {
#pragma unroll 1
            for (int g = 0; g < 3; ++g) {
                wtag = static_cast<uint32_t>(g) << 28;
                walk_stencil<12, true>(GlobalPos{Sp}, kcap, 0u, 0xFFFFFFFFu, column, probe, g, g + 1);
            }
        } else if (sub < 3) {
            walk_stencil<12, true>(GlobalPos{Sp}, kcap, 0u, 0xFFFFFFFFu, column, probe, sub, sub + 1);
        }
    };
    walk();
    const int mine = seen;  // this lane's contacts
    int before = 0;         // L = 4: the grain's contacts in the lanes before this one
    int pairs = mine;       // the grain's pair contacts
    if constexpr (L > 1) {
#pragma unroll
        for (int k = 1; k < L; ++k) {
            const int o = __shfl(mine, lane - k, 64);
            if (sub >= k) before += o;
        }
        pairs = __shfl(before + mine, lane | (L - 1), 64);
    }
    // L = 1: the parts of the groups before group `grp` (whose part the
    // accumulators hold), summed in order.
    float tx[3] = {0.f, 0.f, 0.f}, tv[3] = {0.f, 0.f, 0.f};
    float tk = 0.f, tmu = 0.f;
    int grp = 0;
    auto fold = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            tx[a] = grp == 0 ? gx[a] : tx[a] + gx[a];
            tv[a] = grp == 0 ? gv[a] : tv[a] + gv[a];
            gx[a] = gv[a] = 0.f;
        }
        tk = grp == 0 ? gk : tk + gk;
        tmu = grp == 0 ? gmu : tmu + gmu;
        gk = gmu = 0.f;
        ++grp;
    };
    constexpr uint32_t kSlotBits = 0x0FFFFFFFu;
    for (;;) {
        uint32_t qn = 0;
        float4 pjn = make_float4(0.f, 0.f, 0.f, 0.f), vjn = pjn;
        if (listed > 0) {
            qn = s_list[0][lane];
            pjn = Sp[qn & kSlotBits];
            vjn = Sv[qn & kSlotBits];
        }
        for (int k = 0; k < listed; ++k) {
            const float4 pj = pjn, vj = vjn;
            const uint32_t qk = qn & kSlotBits;
            if constexpr (L == 1) {
                while (grp < static_cast<int>(qn >> 28)) fold();
            }
            // the next contact's loads, unconditional (the last one's again)
            qn = s_list[min(k + 1, listed - 1)][lane];
            pjn = Sp[qn & kSlotBits];
            vjn = Sv[qn & kSlotBits];
            const int kc = before + consumed + k;  // the contact's pair index
            const int slot = wc + kc;              // its spring slot in the new list
            float d0, d1, d2;
            const float norm = pair_norm(x, pj.x, pj.y, pj.z, d0, d1, d2);
            const float dist = norm - P.radius - P.radius;
            const int jd = __float_as_int(pj.w);
            const float nrm[3] = {d0 / norm, d1 / norm, d2 / norm};
            const float raw = P.kneg * dist;
            const float fp = fmaxf(0.f, fminf(P.max_force, raw));
            const float c = mat_cpair(P, member, mat, grain_mat(vj.w));
            const float vr[3] = {v[0] - vj.x, v[1] - vj.y, v[2] - vj.z};
            float vn = 0.f;
            vn += vr[0] * nrm[0];
            vn += vr[1] * nrm[1];
            vn += vr[2] * nrm[2];
            float xi[3], gX[3], g_n[3], g_vr[3], g_xi[3], g_dist;
            const int t = h_find(jd, slot, xi);
            g_next(slot, gX);
            contact_vjp(P, mu, nrm, vr, vn, fp, fp == raw, dist, c, xi, gF, gX, g_n, g_dist, g_vr,
                        g_xi, gk, gmu);
            keep_xi(t, g_xi);
            // n = d / |d|, dist = |d| - 2r
            const float gnn = dot3(g_n, nrm);
            float gd[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) gd[a] = (g_n[a] - gnn * nrm[a]) / norm + g_dist * nrm[a];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                gx[a] += gd[a];
                gv[a] += g_vr[a];
            }
            if (live && kc < kAdjContacts) {
                cg_d[static_cast<size_t>(kc) * n + p] = make_float4(gd[0], gd[1], gd[2], __int_as_float(jd));
                cg_v[static_cast<size_t>(kc) * n + p] = make_float4(g_vr[0], g_vr[1], g_vr[2], 0.f);
                cg_p[static_cast<size_t>(p) * kAdjContacts + kc] = static_cast<int>(qk);
            }
        }
        consumed += listed;
        if (consumed >= mine) break;
        walk();
    }
    // The grain's share: the groups' parts in order (a fourth, zero part
    // stands for lane 3 of four, which takes no group).
    float sx[3], sv[3], sk, smu;
    unsigned w_all = written;
    if constexpr (L == 1) {
        while (grp < 3) fold();
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            sx[a] = tx[a] + 0.f;
            sv[a] = tv[a] + 0.f;
        }
        sk = tk + 0.f;
        smu = tmu + 0.f;
    } else {
        auto grain_sum = [&](float a) __attribute__((always_inline)) {
            const float a1 = __shfl(a, lane + 1, 64), a2 = __shfl(a, lane + 2, 64), a3 = __shfl(a, lane + 3, 64);
            return ((a + a1) + a2) + a3;
        };
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            sx[a] = grain_sum(gx[a]);
            sv[a] = grain_sum(gv[a]);
        }
        sk = grain_sum(gk);
        smu = grain_sum(gmu);
        w_all |= __shfl(written, lane + 1, 64) | __shfl(written, lane + 2, 64) | __shfl(written, lane + 3, 64);
    }
    if (sub == 0 && live) {
        for (int t = 0; t < h_old; ++t) {
            if (!((w_all >> t) & 1u)) lxi_prev[static_cast<size_t>(t) * n + hs] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (pairs > kAdjContacts) ctl->error = kAdjErrContacts;
        cg_cnt[p] = min(pairs, kAdjContacts);
        own_x[p] = make_float4(sx[0], sx[1], sx[2], 0.f);
        own_v[p] = make_float4(sv[0], sv[1], sv[2], 0.f);
        double2 gp = gparam[id];
        gp.x += static_cast<double>(smu);
        gp.y += static_cast<double>(sk);
        gparam[id] = gp;
    }
}

// Pass B: dL/d(state s) of the grain at cell-sorted position j = its own
// share minus the stored shares of every pair contact (i, j) that grain i
// evaluated, written at j's position in the previous step's order
// (hidx[j]).  j's partners are summed in four parts, part k taking partners
// k, k + 4, ... (lane k of four per grain, or one lane part after part);
// the parts are added in order and the sum subtracted from the own share.
// Partner lists are grain-major (cg_p[i * kAdjContacts + t]) so that a search
// reads four partners per load.
template <int L>
__global__ __launch_bounds__(kAdjBlock) void k_adj_gather(const float4* __restrict__ own_x,
                                                       const float4* __restrict__ own_v,
                                                       const float4* __restrict__ cg_d,
                                                       const float4* __restrict__ cg_v,
                                                       const int* __restrict__ cg_cnt,
                                                       const int* __restrict__ cg_p,
                                                       const int* __restrict__ hidx,
                                                       float4* __restrict__ lx, float4* __restrict__ lv,
                                                       int n) {
    static_assert(L == 1 || L == kAdjLanes, "adjoint pass variants");
    constexpr int G = kAdjBlock / L;
    const int lane = threadIdx.x;
    const int sub = lane & (L - 1);
    const int first = blockIdx.x * G;
    if (first >= n) return;
    const bool live = first + lane / L < n;
    const int j = live ? first + lane / L : n - 1;
    const int cj = cg_cnt[j];
    const int* pj = cg_p + static_cast<size_t>(j) * kAdjContacts;
    float e[3] = {0.f, 0.f, 0.f}, f[3] = {0.f, 0.f, 0.f};  // the current part
    float te[3] = {0.f, 0.f, 0.f}, tf[3] = {0.f, 0.f, 0.f};  // the parts before it (L = 1)
#pragma unroll 1
    for (int k = L == 1 ? 0 : sub; k < (L == 1 ? kAdjLanes : sub + 1); ++k) {
#pragma unroll
        for (int a = 0; a < 3; ++a) e[a] = f[a] = 0.f;
        for (int u = k; u < cj; u += kAdjLanes) {
            const int i = pj[u];
            const int ci = cg_cnt[i];
            const int4* pi = reinterpret_cast<const int4*>(cg_p + static_cast<size_t>(i) * kAdjContacts);
            int t = -1;
            for (int b = 0; b < ci && t < 0; b += 8) {
                const int4 w0 = pi[b >> 2], w1 = pi[(b >> 2) + 1];
                const int c8[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (t < 0 && b + q < ci && c8[q] == j) t = b + q;
                }
            }
            if (t >= 0) {
                const float4 d = cg_d[static_cast<size_t>(t) * n + i];
                const float4 w = cg_v[static_cast<size_t>(t) * n + i];
                e[0] += d.x; e[1] += d.y; e[2] += d.z;
                f[0] += w.x; f[1] += w.y; f[2] += w.z;
            }
        }
        if constexpr (L == 1) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                te[a] = k == 0 ? e[a] : te[a] + e[a];
                tf[a] = k == 0 ? f[a] : tf[a] + f[a];
            }
        }
    }
    if constexpr (L > 1) {
        auto grain_sum = [&](float a) __attribute__((always_inline)) {
            const float a1 = __shfl(a, lane + 1, 64), a2 = __shfl(a, lane + 2, 64), a3 = __shfl(a, lane + 3, 64);
            return ((a + a1) + a2) + a3;
        };
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            te[a] = grain_sum(e[a]);
            tf[a] = grain_sum(f[a]);
        }
    }
    if (sub == 0 && live) {
        const float4 gx = own_x[j], gv = own_v[j];
        const int dst = hidx[j];
        lx[dst] = make_float4(gx.x - te[0], gx.y - te[1], gx.z - te[2], gx.w);
        lv[dst] = make_float4(gv.x - tf[0], gv.y - tf[1], gv.z - tf[2], gv.w);
    }
}

// dL/d(state) and the spring cotangents from original order (a checkpoint's
// order) into the cell order of a segment's last recomputed step (Sp: that
// step's cell-sorted positions, the original index in .w).
__global__ __launch_bounds__(kBlock) void k_adj_to_cells(const float4* __restrict__ Sp,
                                                         const float4* __restrict__ lx_id,
                                                         const float4* __restrict__ lv_id,
                                                         const float4* __restrict__ lxi_id,
                                                         float4* __restrict__ lx, float4* __restrict__ lv,
                                                         float4* __restrict__ lxi, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int id = __float_as_int(Sp[p].w);
    lx[p] = lx_id[id];
    lv[p] = lv_id[id];
    for (int t = 0; t < kHist; ++t) lxi[static_cast<size_t>(t) * n + p] = lxi_id[static_cast<size_t>(t) * n + id];
}

}  // namespace wd
