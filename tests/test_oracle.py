"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker:
closed forms, brute force and independent numpy restatements, each chosen so a
plausible slip (dropped term, wrong sign or index, swapped operand) fails."""
import math

import numpy as np
import pytest

from world_amd import _capi, workloads

F = np.float32


def single(x, y, z, vx=0.0, vy=0.0, vz=0.0, **kw):
    c = _capi.default_config()
    for k, v in kw.items():
        setattr(c, k, v)
    st = np.array([x, y, z, vx, vy, vz], np.float32)
    return c, st, np.array([1], np.int32)


# ---------------------------------------------------------------- lattice --

def numpy_cube(c):
    """initialize_to_cube (src/particle_dynamics.cu:165-209) restated with
    numpy's MT19937 (legacy seeding == std::mt19937(seed)) and libstdc++'s
    uniform_real_distribution<float>: a + (b - a) * (u / 2^32), in FP32."""
    r = F(c.particle_radius)
    side = [int(F(L) / (F(2) * r)) for L in (c.cube_length_x, c.cube_length_y, c.cube_length_z)]
    rj = np.random.RandomState(c.init_seed)._bit_generator
    a, b = F(-F(c.init_jitter) * r), F(F(c.init_jitter) * r)
    n = side[0] * side[1] * side[2]
    draws = rj.random_raw(3 * n).astype(np.float64).astype(np.float32) if c.init_jitter > 0 else None
    st = np.zeros((n, 6), np.float32)
    o = [F(c.init_x0), F(c.init_y0), F(c.init_z0)]
    p = 0
    for i in range(side[0]):
        for j in range(side[1]):
            for k in range(side[2]):
                for ax, l in enumerate((i, j, k)):
                    x = F(o[ax] + F(F(F(l) * F(2)) * F(c.lattice_spacing)) * r)
                    if draws is not None:
                        g = F(draws[3 * p + ax] / F(4294967296.0))
                        if g >= 1:
                            g = np.nextafter(F(1), F(0))
                        x = F(x + F(a + F(b - a) * g))
                    else:
                        x = F(x + F(0))
                    st[p, ax] = x
                p += 1
    return st


def test_lattice_matches_numpy_restatement(oracle_lib):
    for c in (_capi.default_config(), workloads.column_collapse_4k()):
        st, mat = oracle_lib.initial_state(c)
        want = numpy_cube(c)
        assert st.size == want.size
        np.testing.assert_array_equal(st.reshape(-1, 6), want)
        assert np.all(mat == c.particle_material)


def test_lattice_golden_fixture(oracle_lib):
    from pathlib import Path

    g = np.load(Path(__file__).parent / "golden" / "lattices.npz")
    for name, c in (("default_cube", _capi.default_config()),
                    ("column_4k", workloads.column_collapse_4k())):
        st, _ = oracle_lib.initial_state(c)
        np.testing.assert_array_equal(st.reshape(-1, 6), g[name])


def test_default_cube_count_and_spacing(oracle_lib):
    c = _capi.default_config()
    st, _ = oracle_lib.initial_state(c)
    st = st.reshape(-1, 6)
    assert st.shape[0] == 1000  # 0.2 / (2 * 0.01) = 10 per axis
    # No jitter: neighbours along z differ by 2 * 1.01 * r.
    np.testing.assert_allclose(st[1, 2] - st[0, 2], 2 * 1.01 * 0.01, rtol=1e-6)
    np.testing.assert_allclose(st[10, 1] - st[0, 1], 2 * 1.01 * 0.01, rtol=1e-6)


# ------------------------------------------------------------- cell keys --

def test_cell_keys_closed_form(oracle_lib):
    c = _capi.default_config()  # 32^3 over [-1, 1]^3: cell width 1/16
    pts = np.array([[-1.0, -1.0, -1.0], [0.0, 0.0, 0.0], [0.99, -0.99, 0.5],
                    [2.0, -5.0, 1.0], [-0.0624, 0.0626, 0.9999]], np.float32)
    st = np.zeros((len(pts), 6), np.float32)
    st[:, :3] = pts
    keys = oracle_lib.cell_keys(c, st)
    cells = [(0, 0, 0), (16, 16, 16), (31, 0, 24), (31, 0, 31), (15, 17, 31)]
    assert list(keys) == [(x * 32 + y) * 32 + z for x, y, z in cells]


def test_cell_order_is_sorted_by_key_then_index(oracle_lib):
    c = workloads.column_collapse_4k()
    st, _ = oracle_lib.initial_state(c)
    rng = np.random.default_rng(1)
    st = st.reshape(-1, 6).copy()
    st[:, :3] += rng.normal(0, 0.01, (len(st), 3)).astype(np.float32)
    ids, keys, _ = oracle_lib.cell_order(c, st)
    k_all = oracle_lib.cell_keys(c, st)
    want = np.lexsort((np.arange(len(st)), k_all))
    np.testing.assert_array_equal(ids, want)
    np.testing.assert_array_equal(keys, k_all[want])


# ------------------------------------------------------------- neighbors --

def brute_rows(c, st):
    """All j != i in the clamped 27-cell stencil, by O(N^2) enumeration."""
    g = np.array([c.collision_grid_size_x, c.collision_grid_size_y, c.collision_grid_size_z])
    lo = np.array([c.x_min, c.y_min, c.z_min], np.float32)
    hi = np.array([c.x_max, c.y_max, c.z_max], np.float32)
    t = ((st[:, :3] - lo) / (hi - lo)) * g.astype(np.float32)
    cell = np.clip(np.trunc(t).astype(np.int64), 0, g - 1)
    out = []
    for i in range(len(st)):
        d = np.abs(cell - cell[i]).max(axis=1)
        js = np.nonzero(d <= 1)[0]
        out.append(set(int(j) for j in js if j != i))
    return out, cell


def test_neighbor_rows_match_brute_force(oracle_lib):
    c = workloads.column_collapse_4k()
    st, _ = oracle_lib.initial_state(c)
    st = st.reshape(-1, 6).copy()
    rng = np.random.default_rng(2)
    st[:, :3] += rng.normal(0, 0.004, (len(st), 3)).astype(np.float32)
    rows, over = oracle_lib.neighbor_rows(c, st)
    assert over == 0
    want, cell = brute_rows(c, st)
    for i in range(0, len(st), 7):
        row = rows[i]
        got = row[: np.argmax(row == -1)]
        assert set(got.tolist()) == want[i]
        assert len(got) == len(want[i])
        # order: stencil cell (dx, dy, dz) lexicographic, then index
        rel = cell[got] - cell[i]
        order_key = ((rel[:, 0] + 1) * 9 + (rel[:, 1] + 1) * 3 + (rel[:, 2] + 1)) * 100000 + got
        assert np.all(np.diff(order_key) > 0)


def test_overflow_counts_cells_and_caps_rows(oracle_lib):
    c = _capi.default_config()
    c.collision_grid_size_x = c.collision_grid_size_y = c.collision_grid_size_z = 1
    c.particles_per_cell = 2
    st, _ = oracle_lib.initial_state(c)
    rows, over = oracle_lib.neighbor_rows(c, st)
    assert over == 1  # one cell, over capacity once per pass
    assert list(rows[0][:2]) == [1, -1]
    assert list(rows[1][:2]) == [0, -1]
    assert list(rows[5][:3]) == [0, 1, -1]


# ------------------------------------------------------------------ forces --

def two_grains(dz, v1=(0, 0, 0), v2=(0, 0, 0), **kw):
    c = _capi.default_config()
    c.gravity = 0.0
    for k, v in kw.items():
        setattr(c, k, v)
    st = np.array([[0, 0, 0, *v1], [0, 0, dz, *v2]], np.float32)
    return c, st, np.array([1, 1], np.int32)


def test_spring_force_closed_form(oracle_lib):
    c, st, mat = two_grains(0.019)  # overlap 1e-3 < r: k * overlap
    rhs, _ = oracle_lib.rhs(c, st, mat, 64)
    k = 100.0 / 0.01
    m = 0.04
    f = k * (0.02 - 0.019)
    # dashpot and friction vanish at rest; grain 0 pushed down, grain 1 up
    np.testing.assert_allclose(rhs[0, 3:], [0, 0, -f / m], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(rhs[1, 3:], [0, 0, f / m], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(rhs[:, :3], st[:, 3:])


def test_spring_clamps_at_max_force(oracle_lib):
    c, st, mat = two_grains(0.005)  # overlap 0.015 > r
    rhs, _ = oracle_lib.rhs(c, st, mat, 64)
    np.testing.assert_allclose(rhs[1, 5], 100.0 / 0.04, rtol=1e-6)


def test_dashpot_and_friction_cap(oracle_lib):
    # approaching along z with a large tangential slip along x
    c, st, mat = two_grains(0.019, v1=(0, 0, 0.5), v2=(3.0, 0, -0.5), friction=0.5)
    rhs, _ = oracle_lib.rhs(c, st, mat, 64)
    k, m, e = 1.0e4, 0.04, 0.3
    mred = m * m / (2 * m)
    ln = math.log(e)
    cdamp = -2 * math.sqrt(mred * k) * ln / math.sqrt(math.pi ** 2 + ln * ln)
    fs = k * 0.001
    vn = (0 - 0) * 0 + (0.5 - (-0.5)) * -1.0  # v_rel . n with n = (x0 - x1)/|.| = -z
    fz0 = -fs - cdamp * vn * -1.0
    # friction on grain 0 opposes its slip (v0 - v1 = -3 along x): +x, capped at mu*fs
    fx0 = min(0.5 * fs, cdamp * 3.0)
    np.testing.assert_allclose(rhs[0, 3:], [fx0 / m, 0, fz0 / m], rtol=1e-5)
    # Newton's third law (equal masses): accelerations are opposite
    np.testing.assert_allclose(rhs[1, 3:], -rhs[0, 3:], rtol=1e-5)


def test_walls_gravity_and_wall_friction(oracle_lib):
    # grain pressed 1e-3 into the floor of the default box, sliding along +y
    c, st, mat = single(0.0, 0.0, -1.0 + 0.01 - 1e-3, vy=2.0)
    rhs, _ = oracle_lib.rhs(c, st, mat, 64)
    m, k, g = 0.04, 1.0e4, 9.81
    fn = k * 1e-3
    ln = math.log(0.3)
    cw = -2 * math.sqrt(m * k) * ln / math.sqrt(math.pi ** 2 + ln * ln)
    fy = -min(0.5 * fn, cw * 2.0)
    np.testing.assert_allclose(rhs[0, 3:], [0, fy / m, (fn - m * g) / m], rtol=1e-5)


def test_fp32_rhs_tracks_fp64(oracle_lib):
    c = workloads.column_collapse_4k()
    st, mat = oracle_lib.initial_state(c)
    st = st.reshape(-1, 6).copy()
    rng = np.random.default_rng(3)
    st[:, :3] += rng.normal(0, 0.0015, (len(st), 3)).astype(np.float32)
    st[:, 3:] = rng.normal(0, 0.2, (len(st), 3)).astype(np.float32)
    r32, _ = oracle_lib.rhs(c, st, mat, 32)
    r64, scale = oracle_lib.rhs(c, st, mat, 64)
    err = np.linalg.norm(r32[:, 3:] - r64[:, 3:], axis=1)
    assert np.all(err <= 1e-5 * scale + 1e-30)
    assert np.count_nonzero(scale > 9.81 * 1.0001) > 100  # contacts were exercised


# -------------------------------------------------------------- stepping --

def test_free_fall_closed_form(oracle_lib):
    c, st, mat = single(0.0, 0.0, 0.5)
    c.dt = 1e-4
    out, _, _ = oracle_lib.steps(c, st, mat, 100)
    dz = -9.81 * 1e-4 ** 2 * 100 * 101 / 2
    assert abs((out[0, 2] - 0.5) - dz) < 1e-6


def test_constant_velocity_transport(oracle_lib):
    c = _capi.default_config()
    c.gravity = 0.0
    c.init_vx0, c.init_vy0, c.init_vz0 = 1.5, -0.5, 0.25
    st, mat = oracle_lib.initial_state(c)
    out, _, _ = oracle_lib.steps(c, st, mat, 50)
    moved = out[:, :3] - st.reshape(-1, 6)[:, :3]
    np.testing.assert_allclose(moved, np.broadcast_to([1.5 * 5e-3, -0.5 * 5e-3, 0.25 * 5e-3], moved.shape),
                               rtol=1e-3)


def test_picard_equals_sie_under_constant_force(oracle_lib):
    c, st, mat = single(0.1, 0.2, 0.5, vx=0.3)
    a, _, _ = oracle_lib.steps(c, st, mat, 40)
    c.time_integrator = _capi.WD_BACKWARD_EULER_PICARD
    b, _, _ = oracle_lib.steps(c, st, mat, 40)
    np.testing.assert_array_equal(a, b)


def test_restitution_of_a_head_on_collision(oracle_lib):
    # unclamped two-body collision: exit / entry speed = e (Tsuji relation)
    c, st, mat = two_grains(0.021, v1=(0, 0, 0.05), v2=(0, 0, -0.05), max_force=1e4, restitution=0.5)
    c.dt = 2e-6
    out, _, _ = oracle_lib.steps(c, st, mat, 20000)
    e = (out[1, 5] - out[0, 5]) / 0.1
    assert abs(e - 0.5) < 0.01


# ----------------------------------------------------------------- energy --

def test_energy_closed_forms(oracle_lib):
    c, st, mat = single(0.1, 0.2, 0.3, vx=1.0, vy=-2.0, vz=0.5)
    e = oracle_lib.energy(c, st, mat, 64)
    want = 0.5 * 0.04 * (1 + 4 + 0.25) + 0.04 * 9.81 * (0.3 - -1.0)
    assert e == pytest.approx(want, rel=1e-7)  # inputs are FP32 values
    assert oracle_lib.energy(c, st, mat, 32) == pytest.approx(want, rel=1e-6)
    # pair overlap d < r: 0.5 k d^2 once; d > r: clamp branch
    for gap, u in ((0.019, 0.5 * 1e4 * 1e-3 ** 2), (0.005, 0.5 * 100 * 0.01 + 100 * (0.015 - 0.01))):
        c2, st2, mat2 = two_grains(gap)
        st2[:, 2] += 0.0  # both well inside the box
        e2 = oracle_lib.energy(c2, st2, mat2, 64)
        assert e2 == pytest.approx(u, rel=1e-6)


# -------------------------------------------------------------------- CIC --

def cic_cfg(x, y, z, v=(0, 0, 0)):
    c = _capi.default_config()
    c.density_grid_size_x = c.density_grid_size_y = c.density_grid_size_z = 5
    st = np.array([[x, y, z, *v]], np.float32)
    return c, st, np.array([1], np.int32)


def test_cic_single_node(oracle_lib):
    c, st, mat = cic_cfg(-0.5, 0.0, 0.5)
    g = oracle_lib.density_grid(c, st, mat)
    node = (1 * 5 + 2) * 5 + 3
    assert g[node] == pytest.approx(0.04, rel=1e-6)
    assert np.count_nonzero(g[:125]) == 1


def test_cic_cell_centre_splits_eight_ways(oracle_lib):
    c, st, mat = cic_cfg(-0.75, -0.25, 0.25)
    g = oracle_lib.density_grid(c, st, mat)[:125]
    nz = g[g > 1e-9]
    assert len(nz) == 8
    np.testing.assert_allclose(nz, 0.04 / 8, rtol=1e-5)


def test_cic_momentum_and_mass_conservation(oracle_lib):
    c, st, mat = cic_cfg(0.13, -0.41, 0.77, v=(2.0, -3.0, 0.5))
    g = oracle_lib.density_grid(c, st, mat).reshape(4, -1).sum(axis=1)
    np.testing.assert_allclose(g, [0.04, 0.08, -0.12, 0.02], rtol=1e-6)


# ---------------------------------------------------- tangential history --

def history_cfg(c, ratio=2.0 / 7.0):
    c = c.copy()
    c.tangential_model = _capi.WD_TANGENTIAL_SPRING_HISTORY
    c.tangential_stiffness_ratio = ratio
    return c


def viscous_cfg(c):
    c = c.copy()
    c.tangential_model = _capi.WD_TANGENTIAL_VISCOUS
    return c


def test_history_without_friction_equals_viscous(oracle_lib):
    # mu = 0: both tangential laws add exact zeros
    c = viscous_cfg(workloads.column_collapse_4k(friction=0.0))
    st, mat = oracle_lib.initial_state(c)
    a, _, _ = oracle_lib.steps(c, st, mat, 30, threads=8)
    b, _, _ = oracle_lib.steps(history_cfg(c), st, mat, 30, threads=8)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("history", [False, True])
def test_sliding_grain_decelerates_at_mu_g(oracle_lib, history):
    # model_problems/sliding_friction.yaml: a grain sliding on the floor
    # decelerates at mu g; fit on vx > 0.1 as plot_friction_validation.py does
    c = workloads.sliding_grain(friction=0.5)
    c = history_cfg(c) if history else viscous_cfg(c)
    st, mat = oracle_lib.initial_state(c)
    sim = oracle_lib.Sim(c, st, mat)
    t, vx = [], []
    for k in range(180):
        sim.steps(100)
        t.append((k + 1) * 100 * c.dt)
        vx.append(sim.state()[0, 3])
    t, vx = np.array(t), np.array(vx)
    sel = vx > 0.1
    slope = np.polyfit(t[sel], vx[sel], 1)[0]
    assert slope == pytest.approx(-0.5 * 9.81, rel=0.01)


def test_history_spring_sticks_then_slips(oracle_lib):
    # A grain on the floor pushed sideways by a small constant velocity step:
    # the stored spring grows by v dt per step while k_t |xi| < mu F_n (stick)
    c = history_cfg(workloads.sliding_grain(friction=0.5, vx0=1e-4, dt=1e-4))
    c.init_z0 = workloads.RADIUS - workloads.GRAIN_MASS * 9.81 / workloads.K_N  # static overlap
    st, mat = oracle_lib.initial_state(c)
    sim = oracle_lib.Sim(c, st, mat)
    sim.steps(1)
    partner, xi, cnt = sim.history()
    assert cnt[0] == 1 and partner[0, 0] == -5  # the floor
    # one step of stick: xi = v_t dt (the spring starts empty)
    assert xi[0, 0, 0] == pytest.approx(1e-4 * 1e-4, rel=1e-5)
    assert xi[0, 0, 2] == 0.0
    # the stored spring's energy enters the total energy
    e_with = sim.energy(64)
    kt = 2.0 / 7.0 * workloads.K_N
    s = sim.state()[0]
    m = workloads.GRAIN_MASS
    ov = workloads.RADIUS - s[2]
    e_expected = (0.5 * m * float(np.dot(s[3:], s[3:])) + m * 9.81 * s[2]
                  + 0.5 * workloads.K_N * ov * ov + 0.5 * kt * float(np.dot(xi[0, 0], xi[0, 0])))
    assert e_with == pytest.approx(e_expected, rel=1e-6)


def test_resume_from_state_and_springs(oracle_lib):
    # A run split in two -- the second half started from the first half's
    # state and stored springs -- equals the unbroken run bit for bit; started
    # without the springs it differs (the springs carry the friction state).
    c = workloads.column(6, 6, 6, friction=0.5)
    st, mat = oracle_lib.initial_state(c)
    st = st.reshape(-1, 6).copy()
    rng = np.random.default_rng(8)
    st[:, :3] += rng.normal(0, 0.0015, (len(st), 3)).astype(np.float32)
    st[:, 3:] = rng.normal(0, 0.3, (len(st), 3)).astype(np.float32)
    whole = oracle_lib.Sim(c, st, mat)
    whole.steps(40)
    first = oracle_lib.Sim(c, st, mat)
    first.steps(20)
    partner, xi, cnt = first.history()
    assert cnt.sum() > 0 and np.any(xi != 0)
    resumed = oracle_lib.Sim(c, first.state(), mat)
    resumed.set_history(partner, xi, cnt)
    p2, x2, c2 = resumed.history()
    np.testing.assert_array_equal(c2, cnt)
    np.testing.assert_array_equal(p2, partner)
    np.testing.assert_array_equal(x2, xi)
    resumed.steps(20)
    np.testing.assert_array_equal(resumed.state(), whole.state())
    cold = oracle_lib.Sim(c, first.state(), mat)
    cold.steps(20)
    assert not np.array_equal(cold.state(), whole.state())


# -- frozen-branch replay: the checker of the adjoint -------------------------

def replay_case(oracle_lib, nsteps=120):
    c = workloads.column(6, 6, 6, friction=0.3)
    st, mat = oracle_lib.initial_state(c)
    st = st.reshape(-1, 6).copy()
    rng = np.random.default_rng(5)
    st[:, :3] += rng.normal(0, 0.0015, (len(st), 3)).astype(np.float32)
    st[:, 3:] = rng.normal(0, 0.2, (len(st), 3)).astype(np.float32)
    return c, st, mat, oracle_lib.Recording(c, st, mat, nsteps)


def test_replay_in_fp32_reproduces_the_forward(oracle_lib):
    c, st, mat, R = replay_case(oracle_lib)
    fin = R.final_state()
    want, _, _ = oracle_lib.steps(c, st, mat, R.nsteps)
    np.testing.assert_array_equal(fin, want)
    s32, rho32 = R.replay(32, st[:, :3], st[:, 3:])
    np.testing.assert_array_equal(s32.astype(np.float32), fin)
    np.testing.assert_array_equal(rho32, oracle_lib.density_grid(c, fin, mat, 32))
    cnt, partners, flags = R.contacts(R.nsteps - 1)
    assert cnt.sum() == len(partners) > 0
    assert (flags >> 1).any() and not (flags >> 1).all()  # both stick and slip occur


def test_replay_tangent_matches_central_differences(oracle_lib):
    """The forward-mode tangent of the FP64 replay against central differences
    of the same frozen-branch function: the error falls as h^2."""
    c, st, mat, R = replay_case(oracle_lib)
    x0, v0 = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    mu, k = float(c.friction), float(c.normal_stiffness)
    rng = np.random.default_rng(9)
    tx, tv = rng.normal(size=x0.shape) * 1e-3, rng.normal(size=v0.shape) * 1e-2
    tmu, tk = 0.1, 10.0
    _, _, dst, drho = R.replay(65, x0, v0, mu, k, tx, tv, tmu, tk)
    W = rng.normal(size=drho.shape)
    errs = []
    for h in (1e-6, 1e-7):
        sp, rp = R.replay(64, x0 + h * tx, v0 + h * tv, mu + h * tmu, k + h * tk)
        sm, rm = R.replay(64, x0 - h * tx, v0 - h * tv, mu - h * tmu, k - h * tk)
        errs.append(abs((W @ rp - W @ rm) / (2 * h) - W @ drho) / abs(W @ drho))
        fds = (sp - sm) / (2 * h)
        assert np.abs(fds - dst).max() <= 0.05 * np.abs(dst).max()
    assert errs[1] < 1e-7 and errs[1] < errs[0] / 30


def test_replay_tangent_closed_form_free_flight(oracle_lib):
    """No contacts: SIE gives x_T = x_0 + T dt v_0 + (gravity terms), so the
    tangent of x_T along v_0 is T dt and along mu or k_n is zero."""
    c = workloads.column(1, 1, 1)
    c.init_type = _capi.WD_INIT_SINGLE_PARTICLE
    st = np.array([[0.0, 0.0, 0.3, 0.1, -0.2, 0.5]], np.float32)
    mat = np.array([1], np.int32)
    R = oracle_lib.Recording(c, st, mat, 50)
    _, _, dst, _ = R.replay(65, st[:, :3], st[:, 3:], 0.5, 150.0, np.zeros((1, 3)),
                            np.array([[1.0, 2.0, 3.0]]), 1.0, 1.0)
    np.testing.assert_allclose(dst[0, :3], np.array([1.0, 2.0, 3.0]) * 50 * c.dt, rtol=1e-12)
    np.testing.assert_allclose(dst[0, 3:], [1.0, 2.0, 3.0], rtol=1e-12)


def test_replay_fp32_primal_tangent(oracle_lib):
    """Precision 66 keeps the FP32 forward's primal bit for bit; over a short
    window its tangent equals the FP64 tangent to FP32 accuracy."""
    c, st, mat, R = replay_case(oracle_lib, nsteps=10)
    x0, v0 = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    rng = np.random.default_rng(4)
    tx, tv = rng.normal(size=x0.shape), rng.normal(size=v0.shape)
    s66, r66, d66, dr66 = R.replay(66, x0, v0, float(c.friction), float(c.normal_stiffness), tx, tv, 0.3, 150.0)
    np.testing.assert_array_equal(s66.astype(np.float32), R.final_state())
    _, _, d65, dr65 = R.replay(65, x0, v0, float(c.friction), float(c.normal_stiffness), tx, tv, 0.3, 150.0)
    assert np.abs(d66 - d65).max() <= 1e-4 * np.abs(d65).max()
    W = rng.normal(size=dr65.shape)
    assert W @ dr66 == pytest.approx(W @ dr65, rel=1e-4)
