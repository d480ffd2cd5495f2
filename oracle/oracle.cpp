// CPU ORACLE -- test infrastructure only.
//
// A plain restatement of the reference's per-step algorithm, used by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the thing
// the GPU engine is checked (and timed) against.  It is never linked into,
// loaded by or dispatched to from the product library.
//
// Each function names the reference code it follows.  The FP32 paths keep
// the reference's operation order exactly (compile with -ffp-contract=off).
// Where the reference is nondeterministic (the in-cell order produced by its
// atomics, src/find_neighbors_kernel.cu:61-112, and its float atomic sums,
// src/density_grid_kernel.cu:72-80, src/energy_kernel.cu:80) the oracle uses
// the canonical form: grains of a cell in ascending original index, sums in
// double in ascending grain order.
//
// How each function is pinned (tests/test_oracle.py):
//   or_initial_state  golden lattice fixtures; closed-form lattice spacing
//   or_cell_*         brute-force recomputation; closed-form cell of a point
//   or_neighbor_rows  O(N^2) brute-force stencil sets; overflow closed form
//   or_rhs            two-body closed forms (spring, dashpot restitution,
//                     friction cap), Newton's third law, FP64 variant
//   or_steps          free-fall closed form, constant-velocity transport,
//                     Picard == SIE under constant force
//   or_energy         closed-form free-fall energy, piecewise contact energy
//   or_density_grid   node / cell-centre / momentum closed forms, mass
//                     conservation
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "../include/world_amd.h"

namespace {

struct Grid {
    int gx, gy, gz;
    float lo[3], hi[3];
};

Grid grid_of(const wd_config& c) {
    Grid g;
    g.gx = c.collision_grid_size_x;
    g.gy = c.collision_grid_size_y;
    g.gz = c.collision_grid_size_z;
    g.lo[0] = c.x_min; g.lo[1] = c.y_min; g.lo[2] = c.z_min;
    g.hi[0] = c.x_max; g.hi[1] = c.y_max; g.hi[2] = c.z_max;
    return g;
}

// static_cast<int>(float) as the GPU performs it: truncate, saturate, NaN -> 0.
int to_int_sat(float t) {
    if (std::isnan(t)) return 0;
    if (t >= 2147483648.0f) return 2147483647;
    if (t <= -2147483648.0f) return -2147483647 - 1;
    return static_cast<int>(t);
}

// src/find_neighbors_kernel.cu:11-22
int cell_of(float x, float lo, float hi, int g) {
    int c = to_int_sat((x - lo) / (hi - lo) * static_cast<float>(g));
    if (c < 0) c = 0;
    if (c > g - 1) c = g - 1;
    return c;
}

void cell3(const Grid& g, const float* p, int c[3]) {
    c[0] = cell_of(p[0], g.lo[0], g.hi[0], g.gx);
    c[1] = cell_of(p[1], g.lo[1], g.hi[1], g.gy);
    c[2] = cell_of(p[2], g.lo[2], g.hi[2], g.gz);
}

// src/find_neighbors_kernel.cu:28-31
long long flat(const Grid& g, int x, int y, int z) {
    return (static_cast<long long>(x) * g.gy + y) * g.gz + z;
}

// Canonical cell table: grains sorted by (flat cell, index), so a cell's
// members are a contiguous run in ascending index.  Sparse: its cost depends
// on the number of grains, not on the size of the domain grid.
struct Bins {
    Grid g;
    std::vector<long long> keys;  // sorted flat cells, one per grain
    std::vector<int> members;     // grain indices in the same order
    long long over = 0;           // cells holding more than k grains

    // [first, first + count) of cell f in members.
    void run(long long f, int& first, int& count) const {
        const auto lo = std::lower_bound(keys.begin(), keys.end(), f);
        const auto hi = std::upper_bound(lo, keys.end(), f);
        first = static_cast<int>(lo - keys.begin());
        count = static_cast<int>(hi - lo);
    }
};

void build_bins(const wd_config& c, int n, const float* st, Bins& b) {
    b.g = grid_of(c);
    std::vector<std::pair<long long, int>> kv(n);
    for (int i = 0; i < n; ++i) {
        int cc[3];
        cell3(b.g, st + 6 * static_cast<size_t>(i), cc);
        kv[i] = {flat(b.g, cc[0], cc[1], cc[2]), i};
    }
    std::sort(kv.begin(), kv.end());
    b.keys.resize(n);
    b.members.resize(n);
    b.over = 0;
    for (int i = 0; i < n; ++i) {
        b.keys[i] = kv[i].first;
        b.members[i] = kv[i].second;
    }
    for (int i = 0; i < n;) {
        int j = i;
        while (j < n && b.keys[j] == b.keys[i]) ++j;
        if (j - i > c.particles_per_cell) b.over++;
        i = j;
    }
}

// Visits the canonical neighbor row of grain i: the 27 stencil cells in
// dx -> dy -> dz order (src/find_neighbors_kernel.cu:130-158), each cell's
// first k members, skipping i itself.
template <typename F>
void for_neighbors(const Bins& b, int k, const float* p, int i, F&& fn) {
    int cc[3];
    cell3(b.g, p, cc);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int x = cc[0] + dx, y = cc[1] + dy, z = cc[2] + dz;
                if (x < 0 || x >= b.g.gx || y < 0 || y >= b.g.gy || z < 0 || z >= b.g.gz) continue;
                int first = 0, cnt = 0;
                b.run(flat(b.g, x, y, z), first, cnt);
                if (cnt > k) cnt = k;
                for (int s = 0; s < cnt; ++s) {
                    const int j = b.members[first + s];
                    if (j != i) fn(j);
                }
            }
}

// src/compute_rhs_kernel.cu:41-47 (evaluated in FP32 on the host).
float damping32(float e, float k, float m) {
    const float pi = 3.14159265358979323846f;
    if (e <= 0.f) return 2.f * std::sqrt(k * m);
    if (e >= 1.f) return 0.f;
    const float ln_e = std::log(e);
    return -2.f * std::sqrt(m * k) * ln_e / std::sqrt(pi * pi + ln_e * ln_e);
}
double damping64(double e, double k, double m) {
    const double pi = 3.14159265358979323846;
    if (e <= 0.0) return 2.0 * std::sqrt(k * m);
    if (e >= 1.0) return 0.0;
    const double ln_e = std::log(e);
    return -2.0 * std::sqrt(m * k) * ln_e / std::sqrt(pi * pi + ln_e * ln_e);
}

struct Phys {
    float g, r, fmax, kneg, k, e, mu;
    float kt, dt;
    bool history;
    float lower[3], upper[3];
    std::vector<float> mass;
};

// Contact history (extension): per grain, springs keyed by partner original
// index (walls -1..-6), in contact order, at most kHistSlots entries.
constexpr int kHistSlots = 24;
struct HEntry {
    int partner;
    float xi[3];
};
using HList = std::vector<HEntry>;

int wall_code(int a, int side) { return -(1 + 2 * a + side); }

template <typename T>
void hist_find(const HList* h, int partner, T xi[3]) {
    xi[0] = xi[1] = xi[2] = T(0);
    if (!h) return;
    for (const HEntry& e : *h) {
        if (e.partner == partner) {
            for (int a = 0; a < 3; ++a) xi[a] = static_cast<T>(e.xi[a]);
            return;
        }
    }
}

template <typename T>
void hist_push(HList* h, int partner, const T xi[3]) {
    if (!h || static_cast<int>(h->size()) >= kHistSlots) return;
    HEntry e;
    e.partner = partner;
    for (int a = 0; a < 3; ++a) e.xi[a] = static_cast<float>(xi[a]);
    h->push_back(e);
}

Phys phys_of(const wd_config& c) {
    Phys p;
    p.g = c.gravity;
    p.r = c.particle_radius;
    p.fmax = c.max_force;
    if (c.normal_stiffness > 0.f) {
        p.k = c.normal_stiffness;
        p.kneg = -c.normal_stiffness;
    } else {
        p.k = c.max_force / c.particle_radius;
        p.kneg = -c.max_force / c.particle_radius;
    }
    p.e = c.restitution;
    p.mu = c.friction;
    p.kt = c.tangential_stiffness_ratio * p.k;
    p.dt = c.dt;
    p.history = c.tangential_model == WD_TANGENTIAL_SPRING_HISTORY;
    p.lower[0] = c.wall_x_min; p.lower[1] = c.wall_y_min; p.lower[2] = c.floor_z;
    p.upper[0] = c.wall_x_max; p.upper[1] = c.wall_y_max; p.upper[2] = c.ceiling_z;
    p.mass.assign(c.masses, c.masses + c.num_masses);
    return p;
}

// True when fminf(a, b) returns a: b is NaN, or a < b with a not NaN.
template <typename T>
bool pick_first_min(T a, T b) {
    if (std::isnan(b)) return true;
    if (std::isnan(a)) return false;
    return a < b;
}

// Branch decisions of one FP32 evaluation, replayed by the FP64 one so that
// a contact / clamp / friction branch flip does not count as a discrepancy.
struct Branches {
    std::vector<uint8_t> bits;
    size_t pos = 0;
    bool recording = true;
    bool take(bool v) {
        if (recording) {
            bits.push_back(v ? 1 : 0);
            return v;
        }
        return bits[pos++] != 0;
    }
};

// src/compute_rhs_kernel.cu:66-78
template <typename T>
void friction(T f[3], const T vt[3], T limit, T c, Branches* br) {
    T s = 0;
    for (int a = 0; a < 3; ++a) s += vt[a] * vt[a];
    const T n = std::sqrt(s);
    const bool stop = br ? br->take(n < T(1e-8f)) : (n < T(1e-8f));
    if (stop) return;
    // fminf(limit, c*n): a NaN operand yields the other one.
    const T cn = c * n;
    const bool lim = br ? br->take(pick_first_min(limit, cn)) : pick_first_min(limit, cn);
    const T mag = lim ? limit : cn;
    for (int a = 0; a < 3; ++a) f[a] -= mag * vt[a] / n;
}

// Tangential spring with history (extension): rotate the stored spring into
// the tangent plane, add v_t dt, F_t = -k_t xi - c v_t, and on Coulomb slip
// scale F_t to mu * fp and reset the spring to match.  Same FP32 order as
// the GPU's spring_friction.
template <typename T>
bool spring_fric(T f[3], const T n[3], const T vr[3], T vn, T fp, T c, T xi[3], T kt, T mu, T dt,
                 Branches* br, int forced_slip = -1) {
    using std::sqrt;
    T vt[3];
    for (int a = 0; a < 3; ++a) vt[a] = vr[a] - vn * n[a];
    T xn = 0;
    for (int a = 0; a < 3; ++a) xn += xi[a] * n[a];
    for (int a = 0; a < 3; ++a) xi[a] = xi[a] - xn * n[a];
    for (int a = 0; a < 3; ++a) xi[a] = xi[a] + vt[a] * dt;
    T ft[3];
    for (int a = 0; a < 3; ++a) ft[a] = -(kt * xi[a]) - c * vt[a];
    T s = 0;
    for (int a = 0; a < 3; ++a) s += ft[a] * ft[a];
    const T fm = sqrt(s);
    const T lim = mu * fp;
    const T inv_kt = T(1) / kt;  // rounded once, as the engine's parameter block holds it
    const bool slip = forced_slip >= 0 ? forced_slip != 0 : (br ? br->take(fm > lim) : (fm > lim));
    if (slip) {
        const T scale = lim / fm;
        for (int a = 0; a < 3; ++a) {
            ft[a] = ft[a] * scale;
            xi[a] = -(ft[a] + c * vt[a]) * inv_kt;
        }
    }
    for (int a = 0; a < 3; ++a) f[a] += ft[a];
    return slip;
}

// max(0, min(fmax, kneg * d)) with fminf/fmaxf semantics (a NaN operand
// yields the other one) and the branch choices recorded / replayed.
template <typename T>
T spring(T fmax, T kneg, T d, Branches* br) {
    const T raw = kneg * d;
    const bool capped = br ? br->take(pick_first_min(fmax, raw)) : pick_first_min(fmax, raw);
    const T inner = capped ? fmax : raw;
    const bool neg = br ? br->take(!(inner > T(0))) : !(inner > T(0));
    return neg ? T(0) : inner;
}

// One contact of one grain in one step, as the FP32 forward took it: the
// partner (original index, or a wall code), whether the normal spring was
// clamped at max_force, and whether the tangential spring slipped.
struct RContact {
    int partner;
    uint8_t capped, slip;
};

// Force on grain i (src/compute_rhs_kernel.cu:81-212).  T = float follows the
// reference's FP32 order; T = double replays the FP32 branch decisions.
// `scale` receives sum |f_term| over every force term (tolerance scale).
template <typename T>
void grain_force(const wd_config& c, const Phys& ph, const Bins& b, int n, const float* st,
                 const int* mat, int i, T out_a[3], Branches* br, double* scale,
                 const HList* oldh = nullptr, HList* newh = nullptr,
                 std::vector<RContact>* rec = nullptr) {
    (void)n;
    const float* si = st + 6 * static_cast<size_t>(i);
    T pos[3], vel[3];
    for (int a = 0; a < 3; ++a) {
        pos[a] = static_cast<T>(si[a]);
        vel[a] = static_cast<T>(si[3 + a]);
    }
    const int mi = mat[i];
    const T m_i = static_cast<T>(ph.mass[mi]);
    const T kk = static_cast<T>(ph.k);
    const T fmax = static_cast<T>(ph.fmax);
    const T kneg = static_cast<T>(ph.kneg);
    const T r = static_cast<T>(ph.r);
    const T mu = static_cast<T>(ph.mu);
    T c_wall;
    if (sizeof(T) == 4) c_wall = static_cast<T>(damping32(ph.e, ph.k, ph.mass[mi]));
    else c_wall = static_cast<T>(damping64(ph.e, kk, m_i));
    const T kt = static_cast<T>(ph.kt), dt = static_cast<T>(ph.dt);
    T f[3] = {0, 0, 0};
    double sc = 0.0;
    f[2] -= m_i * static_cast<T>(ph.g);
    sc += std::fabs(static_cast<double>(m_i * static_cast<T>(ph.g)));
    for (int a = 0; a < 3; ++a) {
        T vt[3];
        for (int q = 0; q < 3; ++q) vt[q] = (q == a) ? T(0) : vel[q];
        const T dl = pos[a] - static_cast<T>(ph.lower[a]) - r;
        const T fl = spring(fmax, kneg, dl, br);
        f[a] += fl;
        sc += std::fabs(static_cast<double>(fl));
        const bool cl = br ? br->take(dl < T(0)) : (dl < T(0));
        if (cl) {
            f[a] -= c_wall * vel[a];
            if (ph.history) {
                T nrm[3] = {0, 0, 0};
                nrm[a] = T(1);
                T vn = 0;
                for (int q = 0; q < 3; ++q) vn += vel[q] * nrm[q];
                T xi[3];
                hist_find(oldh, wall_code(a, 0), xi);
                const bool slip = spring_fric(f, nrm, vel, vn, fl, c_wall, xi, kt, mu, dt, br);
                hist_push(newh, wall_code(a, 0), xi);
                if (rec) rec->push_back({wall_code(a, 0), static_cast<uint8_t>(fl == fmax), static_cast<uint8_t>(slip)});
            } else {
                friction(f, vt, mu * fl, c_wall, br);
            }
            sc += std::fabs(static_cast<double>(c_wall * vel[a])) + std::fabs(static_cast<double>(mu * fl));
        }
        const T du = static_cast<T>(ph.upper[a]) - pos[a] - r;
        const T fu = spring(fmax, kneg, du, br);
        f[a] -= fu;
        sc += std::fabs(static_cast<double>(fu));
        const bool cu = br ? br->take(du < T(0)) : (du < T(0));
        if (cu) {
            f[a] -= c_wall * vel[a];
            if (ph.history) {
                T nrm[3] = {0, 0, 0};
                nrm[a] = T(-1);
                T vn = 0;
                for (int q = 0; q < 3; ++q) vn += vel[q] * nrm[q];
                T xi[3];
                hist_find(oldh, wall_code(a, 1), xi);
                const bool slip = spring_fric(f, nrm, vel, vn, fu, c_wall, xi, kt, mu, dt, br);
                hist_push(newh, wall_code(a, 1), xi);
                if (rec) rec->push_back({wall_code(a, 1), static_cast<uint8_t>(fu == fmax), static_cast<uint8_t>(slip)});
            } else {
                friction(f, vt, mu * fu, c_wall, br);
            }
            sc += std::fabs(static_cast<double>(c_wall * vel[a])) + std::fabs(static_cast<double>(mu * fu));
        }
    }
    for_neighbors(b, c.particles_per_cell, si, i, [&](int j) {
        const float* sj = st + 6 * static_cast<size_t>(j);
        T d[3];
        T s = 0;
        for (int a = 0; a < 3; ++a) {
            d[a] = pos[a] - static_cast<T>(sj[a]);
            s += d[a] * d[a];
        }
        const T norm = std::sqrt(s);
        const T dist = norm - r - r;
        T nrm[3];
        for (int a = 0; a < 3; ++a) nrm[a] = d[a] / norm;
        const T fp = spring(fmax, kneg, dist, br);
        for (int a = 0; a < 3; ++a) f[a] += fp * nrm[a];
        sc += std::fabs(static_cast<double>(fp));
        const bool contact = br ? br->take(dist < T(0)) : (dist < T(0));
        if (contact) {
            const int mj = mat[j];
            T cp;
            if (sizeof(T) == 4) {
                const float msum = ph.mass[mi] + ph.mass[mj];
                const float mred = (msum > 0.f) ? (ph.mass[mi] * ph.mass[mj]) / msum : 0.f;
                cp = static_cast<T>(damping32(ph.e, ph.k, mred));
            } else {
                const double a_ = ph.mass[mi], b_ = ph.mass[mj];
                const double msum = a_ + b_;
                cp = static_cast<T>(damping64(ph.e, kk, (msum > 0) ? a_ * b_ / msum : 0.0));
            }
            T vr[3], vn = 0;
            for (int a = 0; a < 3; ++a) {
                vr[a] = vel[a] - static_cast<T>(sj[3 + a]);
                vn += vr[a] * nrm[a];
            }
            T vt[3];
            for (int a = 0; a < 3; ++a) {
                f[a] -= cp * vn * nrm[a];
                vt[a] = vr[a] - vn * nrm[a];
            }
            if (ph.history) {
                T xi[3];
                hist_find(oldh, j, xi);
                const bool slip = spring_fric(f, nrm, vr, vn, fp, cp, xi, kt, mu, dt, br);
                hist_push(newh, j, xi);
                if (rec) rec->push_back({j, static_cast<uint8_t>(fp == fmax), static_cast<uint8_t>(slip)});
            } else {
                friction(f, vt, mu * fp, cp, br);
            }
            sc += std::fabs(static_cast<double>(cp * vn)) + std::fabs(static_cast<double>(mu * fp));
        }
    });
    for (int a = 0; a < 3; ++a) out_a[a] = f[a] / static_cast<T>(ph.mass[mi]);
    if (scale) *scale = sc / static_cast<double>(ph.mass[mi]);
}

// src/energy_kernel.cu:22-28
template <typename T>
T contact_u(T delta, T r, T fmax, T k, T clamp) {
    if (delta <= T(0)) return T(0);
    if (delta <= clamp) return T(0.5f) * k * delta * delta;
    return T(0.5f) * fmax * clamp + fmax * (delta - clamp);
    (void)r;
}

// src/energy_kernel.cu:30-81 for one grain; pairs counted from the lower index.
template <typename T>
T grain_energy(const wd_config& c, const Phys& ph, const Bins& b, const float* st, const int* mat,
               int i, const HList* h = nullptr) {
    const float* si = st + 6 * static_cast<size_t>(i);
    const T r = static_cast<T>(ph.r), fmax = static_cast<T>(ph.fmax), k = static_cast<T>(ph.k);
    const T clamp = c.normal_stiffness > 0.f ? static_cast<T>(c.max_force / c.normal_stiffness) : r;
    T total = 0;
    T s = 0;
    for (int a = 0; a < 3; ++a) s += static_cast<T>(si[3 + a]) * static_cast<T>(si[3 + a]);
    const T m = static_cast<T>(ph.mass[mat[i]]);
    total += T(0.5f) * m * s;
    total += m * static_cast<T>(ph.g) * (static_cast<T>(si[2]) - static_cast<T>(ph.lower[2]));
    for (int a = 0; a < 3; ++a) {
        const T p = static_cast<T>(si[a]);
        total += contact_u<T>(r - (p - static_cast<T>(ph.lower[a])), r, fmax, k, clamp);
        total += contact_u<T>(r - (static_cast<T>(ph.upper[a]) - p), r, fmax, k, clamp);
    }
    const T kt = static_cast<T>(ph.kt);
    auto spring_u = [&](int partner) {
        T xi[3];
        hist_find(h, partner, xi);
        T s2 = 0;
        for (int a = 0; a < 3; ++a) s2 += xi[a] * xi[a];
        total += T(0.5f) * kt * s2;
    };
    if (ph.history) {  // stored tangential springs of the wall contacts
        for (int a = 0; a < 3; ++a) {
            const T p = static_cast<T>(si[a]);
            if (p - static_cast<T>(ph.lower[a]) - r < T(0)) spring_u(wall_code(a, 0));
            if (static_cast<T>(ph.upper[a]) - p - r < T(0)) spring_u(wall_code(a, 1));
        }
    }
    for_neighbors(b, c.particles_per_cell, si, i, [&](int j) {
        if (j <= i) return;
        const float* sj = st + 6 * static_cast<size_t>(j);
        T d2 = 0;
        for (int a = 0; a < 3; ++a) {
            const T d = static_cast<T>(si[a]) - static_cast<T>(sj[a]);
            d2 += d * d;
        }
        const T norm = std::sqrt(d2);
        total += contact_u<T>(T(2.f) * r - norm, r, fmax, k, clamp);
        if (ph.history && norm - r - r < T(0)) spring_u(j);
    });
    return total;
}

// src/density_grid_kernel.cu:25-41
template <typename T>
void axis_w(T p, T lo, T hi, int size, int& i0, int& i1, T& w0, T& w1) {
    if (size <= 1) {
        i0 = 0; i1 = 0; w0 = 1; w1 = 0;
        return;
    }
    const T h = (hi - lo) / static_cast<T>(size - 1);
    const T u = (p - lo) / h;
    const T base = std::floor(u);
    const T frac = u - base;
    const long long bb = sizeof(T) == 4 ? to_int_sat(static_cast<float>(base))
                                        : static_cast<long long>(std::isnan(base) ? 0 : base);
    auto clampi = [size](long long v) { return static_cast<int>(v < 0 ? 0 : (v > size - 1 ? size - 1 : v)); };
    i0 = clampi(bb);
    i1 = clampi(bb + 1);
    w0 = T(1) - frac;
    w1 = frac;
}

// A stateful oracle simulation: original-order FP32 state plus the contact
// history, advanced by take_step (src/particle_dynamics.cu:264-281).
struct OrSim {
    wd_config c;
    int n = 0;
    std::vector<float> st;
    std::vector<int> mat;
    std::vector<HList> hist;
    long long over = 0;
    float amax = 0.f;
};

void force_pass(OrSim& s, const Phys& ph, const std::vector<HList>& oldh, std::vector<HList>& newh,
                std::vector<float>& acc, int threads) {
    Bins b;
    build_bins(s.c, s.n, s.st.data(), b);
    s.over += b.over;
    newh.assign(s.n, HList());
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int i = 0; i < s.n; ++i) {
        grain_force<float>(s.c, ph, b, s.n, s.st.data(), s.mat.data(), i, &acc[3 * static_cast<size_t>(i)],
                           nullptr, nullptr, ph.history ? &oldh[i] : nullptr,
                           ph.history ? &newh[i] : nullptr);
    }
    s.amax = 0.f;
    for (int i = 0; i < s.n; ++i) {
        float q = 0.f;
        for (int a = 0; a < 3; ++a) q += acc[3 * i + a] * acc[3 * i + a];
        const float m = std::sqrt(q);
        if (m > s.amax) s.amax = m;
    }
}

void sim_steps(OrSim& s, int nsteps, int threads) {
    const Phys ph = phys_of(s.c);
    const float dt = s.c.dt;
    std::vector<float> acc(3 * static_cast<size_t>(s.n));
    std::vector<HList> newh;
    const bool picard = s.c.time_integrator == WD_BACKWARD_EULER_PICARD;
    for (int step = 0; step < nsteps; ++step) {
        if (!picard) {
            force_pass(s, ph, s.hist, newh, acc, threads);
            for (int i = 0; i < s.n; ++i) {
                float* q = s.st.data() + 6 * static_cast<size_t>(i);
                for (int a = 0; a < 3; ++a) q[3 + a] += dt * acc[3 * i + a];
                for (int a = 0; a < 3; ++a) q[a] += dt * q[3 + a];
            }
            if (ph.history) s.hist.swap(newh);
        } else {
            const std::vector<float> base(s.st);
            if (s.c.picard_iterations <= 0) continue;
            for (int k = 0; k < s.c.picard_iterations; ++k) {
                force_pass(s, ph, s.hist, newh, acc, threads);  // springs restart from step n
                for (int i = 0; i < s.n; ++i) {
                    float* q = s.st.data() + 6 * static_cast<size_t>(i);
                    const float* q0 = base.data() + 6 * static_cast<size_t>(i);
                    for (int a = 0; a < 3; ++a) {
                        const float vn = q0[3 + a] + dt * acc[3 * i + a];
                        q[3 + a] = vn;
                        q[a] = q0[a] + dt * vn;
                    }
                }
            }
            if (ph.history) s.hist.swap(newh);
        }
    }
}


// ---------------------------------------------------------------------------
// Frozen-branch replay: the checker of the adjoint.
//
// rec_forward runs the FP32 SIE forward with contact history (the same
// arithmetic as sim_steps) and records, per step and grain, the contacts in
// the order the force pass took them with their clamp and slip decisions.
// replay<T> re-runs that trajectory with every contact set and branch frozen
// to the record: T = float reproduces the forward bit for bit (a check of the
// replay itself), T = double is the FP64 function that central finite
// differences sample, T = Dual carries one forward-mode tangent (the JVP).

struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double a) : v(a), d(0.0) {}  // NOLINT: constants promote
    Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline Dual sqrt(Dual a) {
    const double r = std::sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
// Float primal with an FP64 tangent: the primal operations are the FP32
// forward's own (bit for bit), the tangent is their exact derivative
// evaluated at those FP32 values.  This is the Jacobian at the very points
// the GPU visits.
struct DualF {
    float v = 0.f;
    double d = 0.0;
    DualF() = default;
    DualF(double a) : v(static_cast<float>(a)), d(0.0) {}  // NOLINT: constants promote
    DualF(float a, double b) : v(a), d(b) {}
};
inline DualF operator+(DualF a, DualF b) { return {a.v + b.v, a.d + b.d}; }
inline DualF operator-(DualF a, DualF b) { return {a.v - b.v, a.d - b.d}; }
inline DualF operator-(DualF a) { return {-a.v, -a.d}; }
inline DualF operator*(DualF a, DualF b) {
    return {a.v * b.v, a.d * static_cast<double>(b.v) + static_cast<double>(a.v) * b.d};
}
inline DualF operator/(DualF a, DualF b) {
    const double bv = b.v;
    return {a.v / b.v, (a.d * bv - static_cast<double>(a.v) * b.d) / (bv * bv)};
}
inline DualF& operator+=(DualF& a, DualF b) { return a = a + b; }
inline DualF& operator-=(DualF& a, DualF b) { return a = a - b; }
inline bool operator>(DualF a, DualF b) { return a.v > b.v; }
inline bool operator<(DualF a, DualF b) { return a.v < b.v; }
inline DualF sqrt(DualF a) {
    const float r = std::sqrt(a.v);
    return {r, a.d / (2.0 * static_cast<double>(r))};
}
inline double val(DualF x) { return x.v; }
inline double tan_of(DualF x) { return x.d; }
inline Dual to_dual(DualF x) { return Dual(x.v, x.d); }
inline Dual to_dual(Dual x) { return x; }

inline double val(float x) { return x; }
inline double val(double x) { return x; }
inline double val(Dual x) { return x.v; }
inline double tan_of(float) { return 0.0; }
inline double tan_of(double) { return 0.0; }
inline double tan_of(Dual x) { return x.d; }

template <typename T>
T damping_t(double e, T k, double m) {
    using std::sqrt;
    const double pi = 3.14159265358979323846;
    if (e <= 0.0) return T(2.0) * sqrt(k * T(m));
    if (e >= 1.0) return T(0.0);
    const double ln_e = std::log(e);
    return T(-2.0 * ln_e / std::sqrt(pi * pi + ln_e * ln_e)) * sqrt(T(m) * k);
}

struct RStep {
    std::vector<int> off;  // n + 1
    std::vector<RContact> c;
};

struct Recording {
    wd_config c;
    int n = 0;
    std::vector<float> st0, stT;
    std::vector<int> mat;
    std::vector<RStep> steps;
};

void rec_forward(Recording& R, int nsteps, int threads) {
    OrSim s;
    s.c = R.c;
    s.n = R.n;
    s.st = R.st0;
    s.mat = R.mat;
    s.hist.assign(R.n, HList());
    const Phys ph = phys_of(s.c);
    const float dt = s.c.dt;
    std::vector<float> acc(3 * static_cast<size_t>(s.n));
    std::vector<HList> newh;
    std::vector<std::vector<RContact>> per(R.n);
    R.steps.assign(nsteps, RStep());
    for (int step = 0; step < nsteps; ++step) {
        Bins b;
        build_bins(s.c, s.n, s.st.data(), b);
        newh.assign(s.n, HList());
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
        for (int i = 0; i < s.n; ++i) {
            per[i].clear();
            grain_force<float>(s.c, ph, b, s.n, s.st.data(), s.mat.data(), i, &acc[3 * static_cast<size_t>(i)],
                               nullptr, nullptr, &s.hist[i], &newh[i], &per[i]);
        }
        RStep& rs = R.steps[step];
        rs.off.assign(s.n + 1, 0);
        for (int i = 0; i < s.n; ++i) rs.off[i + 1] = rs.off[i] + static_cast<int>(per[i].size());
        rs.c.resize(rs.off[s.n]);
        for (int i = 0; i < s.n; ++i) std::copy(per[i].begin(), per[i].end(), rs.c.begin() + rs.off[i]);
        for (int i = 0; i < s.n; ++i) {
            float* q = s.st.data() + 6 * static_cast<size_t>(i);
            for (int a = 0; a < 3; ++a) q[3 + a] += dt * acc[3 * i + a];
            for (int a = 0; a < 3; ++a) q[a] += dt * q[3 + a];
        }
        s.hist.swap(newh);
    }
    R.stT = s.st;
}

template <typename T>
struct RHist {
    int partner;
    T xi[3];
};

template <typename T>
void rhist_find(const std::vector<RHist<T>>& h, int partner, T xi[3]) {
    xi[0] = xi[1] = xi[2] = T(0.0);
    for (const RHist<T>& e : h) {
        if (e.partner == partner) {
            for (int a = 0; a < 3; ++a) xi[a] = e.xi[a];
            return;
        }
    }
}

// Parameters of the replayed function.  For T = float they are the forward's
// own FP32 constants; otherwise mu and k are the inputs and every derived
// constant (k_t, the damping table) is formed from k in T.
template <typename T>
struct RParams {
    T mu, k, kneg, kt, fmax, r, dt, g;
    T lower[3], upper[3];
    std::vector<T> mass, c_wall, c_pair;
    int nmat = 0;
};

template <typename T>
RParams<T> rparams(const wd_config& c, const Phys& ph, T mu, T k) {
    RParams<T> p;
    p.nmat = c.num_masses;
    p.mass.resize(p.nmat);
    for (int m = 0; m < p.nmat; ++m) p.mass[m] = T(ph.mass[m]);
    p.fmax = T(ph.fmax);
    p.r = T(ph.r);
    p.dt = T(ph.dt);
    p.g = T(ph.g);
    for (int a = 0; a < 3; ++a) {
        p.lower[a] = T(ph.lower[a]);
        p.upper[a] = T(ph.upper[a]);
    }
    p.c_wall.resize(p.nmat);
    p.c_pair.resize(static_cast<size_t>(p.nmat) * p.nmat);
    if constexpr (std::is_same<T, float>::value) {
        p.mu = T(ph.mu);
        p.k = T(ph.k);
        p.kneg = T(ph.kneg);
        p.kt = T(ph.kt);
        for (int a = 0; a < p.nmat; ++a) {
            p.c_wall[a] = T(damping32(ph.e, ph.k, ph.mass[a]));
            for (int b = 0; b < p.nmat; ++b) {
                const float msum = ph.mass[a] + ph.mass[b];
                const float mred = (msum > 0.f) ? (ph.mass[a] * ph.mass[b]) / msum : 0.f;
                p.c_pair[a * p.nmat + b] = T(damping32(ph.e, ph.k, mred));
            }
        }
    } else if constexpr (std::is_same<T, DualF>::value) {
        // The FP32 forward's constants; c ~ sqrt(k_n), so dc/dk_n = c / (2 k_n).
        p.mu = mu;
        p.k = k;
        p.kneg = T(0.0) - k;
        p.kt = T(static_cast<double>(c.tangential_stiffness_ratio)) * k;
        const float kv = static_cast<float>(val(k));
        const double dk = tan_of(k);
        auto damp = [&](float m) {
            const float cv = damping32(ph.e, kv, m);
            return DualF(cv, static_cast<double>(cv) * dk / (2.0 * static_cast<double>(kv)));
        };
        for (int a = 0; a < p.nmat; ++a) {
            p.c_wall[a] = damp(ph.mass[a]);
            for (int b = 0; b < p.nmat; ++b) {
                const float msum = ph.mass[a] + ph.mass[b];
                const float mred = (msum > 0.f) ? (ph.mass[a] * ph.mass[b]) / msum : 0.f;
                p.c_pair[a * p.nmat + b] = damp(mred);
            }
        }
    } else {
        p.mu = mu;
        p.k = k;
        p.kneg = T(0.0) - k;
        p.kt = T(static_cast<double>(c.tangential_stiffness_ratio)) * k;
        for (int a = 0; a < p.nmat; ++a) {
            p.c_wall[a] = damping_t<T>(ph.e, k, ph.mass[a]);
            for (int b = 0; b < p.nmat; ++b) {
                const double ma = ph.mass[a], mb = ph.mass[b];
                const double mred = (ma + mb > 0.0) ? ma * mb / (ma + mb) : 0.0;
                p.c_pair[a * p.nmat + b] = damping_t<T>(ph.e, k, mred);
            }
        }
    }
    return p;
}

// One grain's acceleration with the recorded contacts of one step, in the
// order and arithmetic of grain_force (contacts only: a candidate that does
// not touch adds exact zeros there).
template <typename T>
void replay_grain(const RParams<T>& P, const T* st, const int* mat, int i, const RContact* c, int nc,
                  const std::vector<RHist<T>>& oldh, std::vector<RHist<T>>& newh, T out_a[3]) {
    using std::sqrt;
    const T* si = st + 6 * static_cast<size_t>(i);
    T pos[3], vel[3];
    for (int a = 0; a < 3; ++a) {
        pos[a] = si[a];
        vel[a] = si[3 + a];
    }
    const int mi = mat[i];
    const T m_i = P.mass[mi];
    const T c_wall = P.c_wall[mi];
    T f[3] = {T(0.0f), T(0.0f), T(0.0f)};
    f[2] -= m_i * P.g;
    newh.clear();
    auto push = [&](int partner, const T xi[3]) {
        if (static_cast<int>(newh.size()) >= kHistSlots) return;
        RHist<T> e;
        e.partner = partner;
        for (int a = 0; a < 3; ++a) e.xi[a] = xi[a];
        newh.push_back(e);
    };
    for (int k = 0; k < nc; ++k) {
        const RContact& e = c[k];
        if (e.partner < 0) {
            const int a = (-e.partner - 1) / 2, side = (-e.partner - 1) % 2;
            const T d = side == 0 ? pos[a] - P.lower[a] - P.r : P.upper[a] - pos[a] - P.r;
            const T fw = e.capped ? P.fmax : P.kneg * d;
            if (side == 0) f[a] += fw;
            else f[a] -= fw;
            f[a] -= c_wall * vel[a];
            T nrm[3] = {T(0.0f), T(0.0f), T(0.0f)};
            nrm[a] = side == 0 ? T(1.0f) : T(-1.0f);
            T vn = T(0.0f);
            for (int q = 0; q < 3; ++q) vn += vel[q] * nrm[q];
            T xi[3];
            rhist_find(oldh, e.partner, xi);
            spring_fric<T>(f, nrm, vel, vn, fw, c_wall, xi, P.kt, P.mu, P.dt, nullptr, e.slip);
            push(e.partner, xi);
        } else {
            const T* sj = st + 6 * static_cast<size_t>(e.partner);
            T d[3];
            T s = T(0.0f);
            for (int a = 0; a < 3; ++a) {
                d[a] = pos[a] - sj[a];
                s += d[a] * d[a];
            }
            const T norm = sqrt(s);
            const T dist = norm - P.r - P.r;
            T nrm[3];
            for (int a = 0; a < 3; ++a) nrm[a] = d[a] / norm;
            const T fp = e.capped ? P.fmax : P.kneg * dist;
            for (int a = 0; a < 3; ++a) f[a] += fp * nrm[a];
            const T cp = P.c_pair[mi * P.nmat + mat[e.partner]];
            T vr[3], vn = T(0.0f);
            for (int a = 0; a < 3; ++a) {
                vr[a] = vel[a] - sj[3 + a];
                vn += vr[a] * nrm[a];
            }
            for (int a = 0; a < 3; ++a) f[a] -= cp * vn * nrm[a];
            T xi[3];
            rhist_find(oldh, e.partner, xi);
            spring_fric<T>(f, nrm, vr, vn, fp, cp, xi, P.kt, P.mu, P.dt, nullptr, e.slip);
            push(e.partner, xi);
        }
    }
    for (int a = 0; a < 3; ++a) out_a[a] = f[a] / m_i;
}

template <typename A, typename T>
A acc_of(T x) {
    if constexpr (std::is_same<A, Dual>::value && std::is_same<T, DualF>::value) return to_dual(x);
    else return A(x);
}

// CIC deposit of a T state (axis weights as axis_w), accumulated in A.
template <typename T, typename A>
void cic_t(const wd_config& c, int n, const T* st, const int* mat, A* out) {
    const int size[3] = {c.density_grid_size_x, c.density_grid_size_y, c.density_grid_size_z};
    const size_t nodes = static_cast<size_t>(size[0]) * size[1] * size[2];
    for (size_t k = 0; k < 4 * nodes; ++k) out[k] = A(0.0);
    const float lo[3] = {c.x_min, c.y_min, c.z_min}, hi[3] = {c.x_max, c.y_max, c.z_max};
    for (int i = 0; i < n; ++i) {
        const T* s = st + 6 * static_cast<size_t>(i);
        const float mf = c.masses[mat[i]];
        if (mf == 0.f) continue;
        const T m = T(mf);
        int idx[3][2];
        T w[3][2];
        for (int a = 0; a < 3; ++a) {
            if (size[a] <= 1) {
                idx[a][0] = idx[a][1] = 0;
                w[a][0] = T(1.0f);
                w[a][1] = T(0.0f);
                continue;
            }
            const T h = (T(hi[a]) - T(lo[a])) / T(static_cast<float>(size[a] - 1));
            const T u = (s[a] - T(lo[a])) / h;
            const double bv = std::floor(val(u));
            const T frac = u - T(static_cast<float>(bv));
            long long bb;
            if (sizeof(T) == 4 || std::is_same<T, DualF>::value) bb = to_int_sat(static_cast<float>(bv));
            else bb = std::isnan(bv) ? 0 : static_cast<long long>(bv);
            auto clampi = [&](long long v) { return static_cast<int>(v < 0 ? 0 : (v > size[a] - 1 ? size[a] - 1 : v)); };
            idx[a][0] = clampi(bb);
            idx[a][1] = clampi(bb + 1);
            w[a][0] = T(1.0f) - frac;
            w[a][1] = frac;
        }
        const T mom[3] = {m * s[3], m * s[4], m * s[5]};
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y)
                for (int z = 0; z < 2; ++z) {
                    const T ww = w[0][x] * w[1][y] * w[2][z];
                    if (val(ww) == 0.0 && tan_of(ww) == 0.0) continue;
                    const size_t node = (static_cast<size_t>(idx[0][x]) * size[1] + idx[1][y]) * size[2] + idx[2][z];
                    out[node] += acc_of<A>(m * ww);
                    for (int a = 0; a < 3; ++a) out[(a + 1) * nodes + node] += acc_of<A>(mom[a] * ww);
                }
    }
}

template <typename T>
void replay_run(const Recording& R, const RParams<T>& P, std::vector<T>& st, int nsteps, int threads) {
    const int n = R.n;
    std::vector<std::vector<RHist<T>>> oldh(n), newh(n);
    std::vector<T> acc(3 * static_cast<size_t>(n));
    for (int step = 0; step < nsteps; ++step) {
        const RStep& rs = R.steps[step];
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
        for (int i = 0; i < n; ++i) {
            replay_grain<T>(P, st.data(), R.mat.data(), i, rs.c.data() + rs.off[i], rs.off[i + 1] - rs.off[i],
                            oldh[i], newh[i], &acc[3 * static_cast<size_t>(i)]);
        }
        for (int i = 0; i < n; ++i) {
            T* q = st.data() + 6 * static_cast<size_t>(i);
            for (int a = 0; a < 3; ++a) q[3 + a] += P.dt * acc[3 * i + a];
            for (int a = 0; a < 3; ++a) q[a] += P.dt * q[3 + a];
        }
        oldh.swap(newh);
    }
}
}  // namespace

extern "C" {

// initialize_to_cube / two_particles / single_particle
// (src/particle_dynamics.cu:128-209); file loading is the caller's business.
int or_initial_state(const wd_config* c, float* st, int* mat, int cap, int* n_out) {
    const float r = c->particle_radius;
    if (c->init_type == WD_INIT_TWO_PARTICLES) {
        *n_out = 2;
        if (cap < 2) return -1;
        const float s[12] = {c->init_x0, c->init_y0, c->init_z0, 0, 0, 0,
                             c->init_x0, c->init_y0, c->init_z0 + c->two_particle_separation, 0, 0, 0};
        std::memcpy(st, s, sizeof(s));
        mat[0] = mat[1] = c->particle_material;
        *n_out = 2;
        return 0;
    }
    if (c->init_type == WD_INIT_SINGLE_PARTICLE) {
        *n_out = 1;
        if (cap < 1) return -1;
        const float s[6] = {c->init_x0, c->init_y0, c->init_z0, c->init_vx0, c->init_vy0, c->init_vz0};
        std::memcpy(st, s, sizeof(s));
        mat[0] = c->particle_material;
        *n_out = 1;
        return 0;
    }
    const float L[3] = {c->cube_length_x, c->cube_length_y, c->cube_length_z};
    int side[3];
    for (int a = 0; a < 3; ++a) side[a] = static_cast<int>(L[a] / (2 * r));
    const int n = side[0] * side[1] * side[2];
    *n_out = n;
    if (n > cap) return -1;
    std::mt19937 rj(c->init_seed), rp(c->init_perturbation_seed);
    std::uniform_real_distribution<float> uj(-c->init_jitter * r, c->init_jitter * r);
    std::uniform_real_distribution<float> up(-c->init_perturbation * r, c->init_perturbation * r);
    std::normal_distribution<float> gp(0.f, c->init_perturbation * r);
    const float o[3] = {c->init_x0, c->init_y0, c->init_z0};
    const float v[3] = {c->init_vx0, c->init_vy0, c->init_vz0};
    int p = 0;
    for (int i = 0; i < side[0]; ++i)
        for (int j = 0; j < side[1]; ++j)
            for (int k = 0; k < side[2]; ++k, ++p) {
                const int l[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    const float dj = c->init_jitter > 0.f ? uj(rj) : 0.f;
                    float dp = 0.f;
                    if (c->init_perturbation > 0.f) {
                        dp = c->perturbation_distribution == WD_PERTURB_GAUSSIAN ? gp(rp) : up(rp);
                    }
                    st[6 * p + a] = o[a] + static_cast<float>(l[a]) * 2.f * c->lattice_spacing * r + dj + dp;
                    st[6 * p + 3 + a] = v[a];
                }
                mat[p] = c->particle_material;
            }
    return 0;
}

// Flat cell index of every grain (particle_cell, src/find_neighbors_kernel.cu:35-40).
void or_cell_keys(const wd_config* c, int n, const float* st, long long* keys) {
    const Grid g = grid_of(*c);
    for (int i = 0; i < n; ++i) {
        int cc[3];
        cell3(g, st + 6 * static_cast<size_t>(i), cc);
        keys[i] = flat(g, cc[0], cc[1], cc[2]);
    }
}

// Canonical cell order: ids sorted by (flat cell, original index).
long long or_cell_order(const wd_config* c, int n, const float* st, int* ids, long long* keys) {
    Bins b;
    build_bins(*c, n, st, b);
    for (int p = 0; p < n; ++p) {
        ids[p] = b.members[p];
        int cc[3];
        cell3(b.g, st + 6 * static_cast<size_t>(ids[p]), cc);
        keys[p] = flat(b.g, cc[0], cc[1], cc[2]);
    }
    return b.over;
}

// Neighbor rows in the reference's flat -1-terminated format
// (src/find_neighbors_kernel.cu:119-161); returns the over-capacity cell count.
long long or_neighbor_rows(const wd_config* c, int n, const float* st, int* out) {
    Bins b;
    build_bins(*c, n, st, b);
    const size_t row = static_cast<size_t>(27) * c->particles_per_cell;
    for (int i = 0; i < n; ++i) {
        size_t w = 0;
        int* r = out + row * i;
        for_neighbors(b, c->particles_per_cell, st + 6 * static_cast<size_t>(i), i,
                      [&](int j) { if (w + 1 < row) r[w++] = j; });
        r[w] = -1;
    }
    return b.over;
}

// rhs = [v, F/m] (src/compute_rhs_kernel.cu:81-212).  precision 32: FP32 in
// the reference order.  precision 64: FP64 with the FP32 branch decisions;
// `scale` (may be null) gets the per-grain sum of |force term| / m.
void or_rhs(const wd_config* c, int n, const float* st, const int* mat, int precision,
            double* rhs, double* scale) {
    Bins b;
    build_bins(*c, n, st, b);
    const Phys ph = phys_of(*c);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        double* r = rhs + 6 * static_cast<size_t>(i);
        for (int a = 0; a < 3; ++a) r[a] = st[6 * static_cast<size_t>(i) + 3 + a];
        if (precision == 32) {
            float acc[3];
            grain_force<float>(*c, ph, b, n, st, mat, i, acc, nullptr, nullptr);
            for (int a = 0; a < 3; ++a) r[3 + a] = acc[a];
        } else {
            Branches br;
            float acc32[3];
            double sc = 0.0;
            grain_force<float>(*c, ph, b, n, st, mat, i, acc32, &br, &sc);
            br.recording = false;
            double acc[3];
            grain_force<double>(*c, ph, b, n, st, mat, i, acc, &br, nullptr);
            for (int a = 0; a < 3; ++a) r[3 + a] = acc[a];
            if (scale) scale[i] = sc;
        }
    }
}

void* or_sim_create(const wd_config* c, int n, const float* st, const int* mat) {
    OrSim* s = new OrSim();
    s->c = *c;
    s->n = n;
    s->st.assign(st, st + 6 * static_cast<size_t>(n));
    s->mat.assign(mat, mat + n);
    s->hist.assign(n, HList());
    return s;
}

void or_sim_destroy(void* h) { delete static_cast<OrSim*>(h); }

void or_sim_set_dt(void* h, float dt) { static_cast<OrSim*>(h)->c.dt = dt; }

void or_sim_steps(void* h, int nsteps, int threads) { sim_steps(*static_cast<OrSim*>(h), nsteps, threads); }

void or_sim_state(void* h, float* out) {
    const OrSim* s = static_cast<OrSim*>(h);
    std::memcpy(out, s->st.data(), sizeof(float) * s->st.size());
}

long long or_sim_overflow(void* h) { return static_cast<OrSim*>(h)->over; }

float or_sim_max_accel(void* h) { return static_cast<OrSim*>(h)->amax; }

// Stored springs in original order, kHistSlots per grain (zero-filled).
void or_sim_history(void* h, int* partner, float* xi, int* count) {
    const OrSim* s = static_cast<OrSim*>(h);
    for (int i = 0; i < s->n; ++i) {
        const HList& l = s->hist[i];
        count[i] = static_cast<int>(l.size());
        for (int k = 0; k < kHistSlots; ++k) {
            const size_t o = static_cast<size_t>(i) * kHistSlots + k;
            const bool v = k < static_cast<int>(l.size());
            partner[o] = v ? l[k].partner : 0;
            for (int a = 0; a < 3; ++a) xi[3 * o + a] = v ? l[k].xi[a] : 0.f;
        }
    }
}

// Replaces the stored springs (the layout or_sim_history writes), so that a
// simulation can continue from a state and springs taken mid-run.
void or_sim_set_history(void* h, const int* partner, const float* xi, const int* count) {
    OrSim* s = static_cast<OrSim*>(h);
    for (int i = 0; i < s->n; ++i) {
        HList& l = s->hist[i];
        l.clear();
        const int k_n = std::min(std::max(count[i], 0), kHistSlots);
        for (int k = 0; k < k_n; ++k) {
            const size_t o = static_cast<size_t>(i) * kHistSlots + k;
            HEntry e;
            e.partner = partner[o];
            for (int a = 0; a < 3; ++a) e.xi[a] = xi[3 * o + a];
            l.push_back(e);
        }
    }
}

// Energy of the current state including the stored tangential springs.
double or_sim_energy(void* h, int precision) {
    const OrSim* s = static_cast<OrSim*>(h);
    Bins b;
    build_bins(s->c, s->n, s->st.data(), b);
    const Phys ph = phys_of(s->c);
    double total = 0.0;
    for (int i = 0; i < s->n; ++i) {
        const HList* hl = ph.history ? &s->hist[i] : nullptr;
        total += precision == 32
                ? static_cast<double>(grain_energy<float>(s->c, ph, b, s->st.data(), s->mat.data(), i, hl))
                : grain_energy<double>(s->c, ph, b, s->st.data(), s->mat.data(), i, hl);
    }
    return total;
}

// nsteps of take_step in FP32 on the original-order state, in place, from an
// empty contact history.  Returns the over-capacity cell count summed over
// every neighbor pass; *max_accel gets max |a| of the last force pass
// (src/particle_dynamics.cu:328-344).
long long or_steps(const wd_config* c, int n, float* st, const int* mat, int nsteps, int threads,
                   float* max_accel) {
    OrSim s;
    s.c = *c;
    s.n = n;
    s.st.assign(st, st + 6 * static_cast<size_t>(n));
    s.mat.assign(mat, mat + n);
    s.hist.assign(n, HList());
    sim_steps(s, nsteps, threads);
    std::memcpy(st, s.st.data(), sizeof(float) * s.st.size());
    if (max_accel) *max_accel = s.amax;
    return s.over;
}

// Total energy (src/energy_kernel.cu:30-81).  precision 32: per-grain FP32
// totals in the reference order, summed in double by ascending index.
// precision 64: every term in double.
double or_energy(const wd_config* c, int n, const float* st, const int* mat, int precision) {
    Bins b;
    build_bins(*c, n, st, b);
    const Phys ph = phys_of(*c);
    double total = 0.0;
    for (int i = 0; i < n; ++i) {
        total += precision == 32 ? static_cast<double>(grain_energy<float>(*c, ph, b, st, mat, i))
                                 : grain_energy<double>(*c, ph, b, st, mat, i);
    }
    return total;
}

// Energy of each grain (FP32 reference order), for per-grain checks.
void or_grain_energy(const wd_config* c, int n, const float* st, const int* mat, float* out) {
    Bins b;
    build_bins(*c, n, st, b);
    const Phys ph = phys_of(*c);
    for (int i = 0; i < n; ++i) out[i] = grain_energy<float>(*c, ph, b, st, mat, i);
}

// CIC mass + momentum deposit (src/density_grid_kernel.cu:43-106), channel
// major [4][nx][ny][nz].  precision 32: FP32 weights and products as the
// reference forms them, summed in double.  precision 64: all in double.
void or_density_grid(const wd_config* c, int n, const float* st, const int* mat, int precision,
                     double* out) {
    const int sx = c->density_grid_size_x, sy = c->density_grid_size_y, sz = c->density_grid_size_z;
    const size_t nodes = static_cast<size_t>(sx) * sy * sz;
    std::memset(out, 0, sizeof(double) * 4 * nodes);
    const float lo[3] = {c->x_min, c->y_min, c->z_min}, hi[3] = {c->x_max, c->y_max, c->z_max};
    const int size[3] = {sx, sy, sz};
    for (int i = 0; i < n; ++i) {
        const float* s = st + 6 * static_cast<size_t>(i);
        const float m = c->masses[mat[i]];
        if (m == 0.f) continue;
        int idx[3][2];
        if (precision == 32) {
            float w[3][2];
            for (int a = 0; a < 3; ++a) axis_w<float>(s[a], lo[a], hi[a], size[a], idx[a][0], idx[a][1], w[a][0], w[a][1]);
            const float mom[3] = {m * s[3], m * s[4], m * s[5]};
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y)
                    for (int z = 0; z < 2; ++z) {
                        const float ww = w[0][x] * w[1][y] * w[2][z];
                        if (ww == 0.f) continue;
                        const size_t node = (static_cast<size_t>(idx[0][x]) * sy + idx[1][y]) * sz + idx[2][z];
                        out[node] += static_cast<double>(m * ww);
                        for (int a = 0; a < 3; ++a) out[(a + 1) * nodes + node] += static_cast<double>(mom[a] * ww);
                    }
        } else {
            double w[3][2];
            for (int a = 0; a < 3; ++a) axis_w<double>(s[a], lo[a], hi[a], size[a], idx[a][0], idx[a][1], w[a][0], w[a][1]);
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y)
                    for (int z = 0; z < 2; ++z) {
                        const double ww = w[0][x] * w[1][y] * w[2][z];
                        if (ww == 0.0) continue;
                        const size_t node = (static_cast<size_t>(idx[0][x]) * sy + idx[1][y]) * sz + idx[2][z];
                        out[node] += static_cast<double>(m) * ww;
                        for (int a = 0; a < 3; ++a) out[(a + 1) * nodes + node] += static_cast<double>(m) * s[3 + a] * ww;
                    }
        }
    }
}


// ---- frozen-branch replay ---------------------------------------------------

// Records nsteps of the FP32 SIE forward with contact history.
void* or_rec_create(const wd_config* c, int n, const float* st, const int* mat, int nsteps, int threads) {
    Recording* R = new Recording();
    R->c = *c;
    R->n = n;
    R->st0.assign(st, st + 6 * static_cast<size_t>(n));
    R->mat.assign(mat, mat + n);
    rec_forward(*R, nsteps, threads);
    return R;
}

void or_rec_destroy(void* h) { delete static_cast<Recording*>(h); }

void or_rec_state(void* h, float* out) {
    const Recording* R = static_cast<Recording*>(h);
    std::memcpy(out, R->stT.data(), sizeof(float) * R->stT.size());
}

// Contacts of one recorded step: per-grain counts, and (if partners is not
// null) the partners and flags (bit 0 clamp, bit 1 slip) in order.
long long or_rec_contacts(void* h, int step, int* counts, int* partners, int* flags) {
    const Recording* R = static_cast<Recording*>(h);
    const RStep& rs = R->steps[step];
    for (int i = 0; i < R->n; ++i) counts[i] = rs.off[i + 1] - rs.off[i];
    if (partners) {
        for (size_t k = 0; k < rs.c.size(); ++k) {
            partners[k] = rs.c[k].partner;
            flags[k] = rs.c[k].capped | (rs.c[k].slip << 1);
        }
    }
    return static_cast<long long>(rs.c.size());
}

// Replays the recorded trajectory from state (x0, v0) with parameters
// (mu, k_n) and the recorded branches.  precision 32: FP32 with the
// forward's own constants (mu and k_n ignored); 64: FP64; 65: FP64 with the
// tangent (tx0, tv0, tmu, tk) carried forward (JVP); 66: the FP32 forward's
// own primal values (bit for bit) with an FP64 tangent, i.e. the Jacobian at
// the FP32 trajectory.  Writes the final state and its CIC grid [4][nodes];
// for 65 and 66 also their tangents.
void or_rec_replay(void* h, int precision, const double* x0, const double* v0, double mu, double k,
                   const double* tx0, const double* tv0, double tmu, double tk, double* state_out,
                   double* dstate_out, double* rho_out, double* drho_out, int threads) {
    const Recording* R = static_cast<Recording*>(h);
    const int n = R->n;
    const Phys ph = phys_of(R->c);
    const size_t nodes = static_cast<size_t>(R->c.density_grid_size_x) * R->c.density_grid_size_y *
                         R->c.density_grid_size_z;
    const int nsteps = static_cast<int>(R->steps.size());
    if (precision == 32) {
        std::vector<float> st(6 * static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                st[6 * i + a] = static_cast<float>(x0[3 * i + a]);
                st[6 * i + 3 + a] = static_cast<float>(v0[3 * i + a]);
            }
        const RParams<float> P = rparams<float>(R->c, ph, 0.f, 0.f);
        replay_run<float>(*R, P, st, nsteps, threads);
        for (size_t q = 0; q < st.size(); ++q) state_out[q] = st[q];
        if (rho_out) cic_t<float, double>(R->c, n, st.data(), R->mat.data(), rho_out);
    } else if (precision == 64) {
        std::vector<double> st(6 * static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                st[6 * i + a] = x0[3 * i + a];
                st[6 * i + 3 + a] = v0[3 * i + a];
            }
        const RParams<double> P = rparams<double>(R->c, ph, mu, k);
        replay_run<double>(*R, P, st, nsteps, threads);
        std::memcpy(state_out, st.data(), sizeof(double) * st.size());
        if (rho_out) cic_t<double, double>(R->c, n, st.data(), R->mat.data(), rho_out);
    } else if (precision == 66) {
        std::vector<DualF> st(6 * static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                st[6 * i + a] = DualF(static_cast<float>(x0[3 * i + a]), tx0 ? tx0[3 * i + a] : 0.0);
                st[6 * i + 3 + a] = DualF(static_cast<float>(v0[3 * i + a]), tv0 ? tv0[3 * i + a] : 0.0);
            }
        const RParams<DualF> P = rparams<DualF>(R->c, ph, DualF(static_cast<float>(mu), tmu),
                                                DualF(static_cast<float>(k), tk));
        replay_run<DualF>(*R, P, st, nsteps, threads);
        for (size_t q = 0; q < st.size(); ++q) {
            state_out[q] = st[q].v;
            if (dstate_out) dstate_out[q] = st[q].d;
        }
        if (rho_out || drho_out) {
            std::vector<Dual> rho(4 * nodes);
            cic_t<DualF, Dual>(R->c, n, st.data(), R->mat.data(), rho.data());
            for (size_t q = 0; q < rho.size(); ++q) {
                if (rho_out) rho_out[q] = rho[q].v;
                if (drho_out) drho_out[q] = rho[q].d;
            }
        }
    } else {
        std::vector<Dual> st(6 * static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                st[6 * i + a] = Dual(x0[3 * i + a], tx0 ? tx0[3 * i + a] : 0.0);
                st[6 * i + 3 + a] = Dual(v0[3 * i + a], tv0 ? tv0[3 * i + a] : 0.0);
            }
        const RParams<Dual> P = rparams<Dual>(R->c, ph, Dual(mu, tmu), Dual(k, tk));
        replay_run<Dual>(*R, P, st, nsteps, threads);
        for (size_t q = 0; q < st.size(); ++q) {
            state_out[q] = st[q].v;
            if (dstate_out) dstate_out[q] = st[q].d;
        }
        if (rho_out || drho_out) {
            std::vector<Dual> rho(4 * nodes);
            cic_t<Dual, Dual>(R->c, n, st.data(), R->mat.data(), rho.data());
            for (size_t q = 0; q < rho.size(); ++q) {
                if (rho_out) rho_out[q] = rho[q].v;
                if (drho_out) drho_out[q] = rho[q].d;
            }
        }
    }
}
}  // extern "C"
