"""The benchmark configurations at their real sizes, against the CPU oracle.

  * the 32,768-grain column (semi-implicit Euler and backward-Euler-Picard
    with 3 iterations, mu 0.5 and 0, energy every step, 64^3 CIC): states,
    springs and max |a| bit for bit over the first steps from the lattice and
    over steps taken from a mid-collapse state and its springs, energies and
    the CIC grid within the order-independent sums' tolerance;
  * the 1,048,576-grain column (4 x 4 m box, 128^3 CIC): cell order and the
    first steps bit for bit, then a window into the collapse with whole-size
    invariants (stable, no over-capacity cell, deposited mass n m, energy
    bounded by its initial value) and steps from the densest part of the
    collapse bit for bit;
  * the 64-member uneven ensemble (2,048..32,768 grains per member): every
    member against its own oracle run, from the lattice and mid-run;
  * the 16-member uneven adjoint batch (4,096..16,384 grains, 100-step window,
    checkpoints every 10 steps): dot-product test of one member's gradient
    against the oracle's frozen-branch tangent.
Bit-exact for states, springs, cell order and max |a|; 1e-6 relative for the
CIC deposit (fixed-point sums against the oracle's double sums) and the
standalone energy; 1e-12 for the fused energy log (double sums in another
order); 1e-4 for the adjoint's dot products."""
import numpy as np
import pytest

from world_amd import _capi, workloads
from world_amd.sim import Ensemble, ParticleDynamics

pytestmark = pytest.mark.gpu

THREADS = 16


def oracle_from(oracle_lib, cfg, state, mat, hist=None):
    o = oracle_lib.Sim(cfg, state, mat)
    if hist is not None:
        o.set_history(*hist)
    return o


def assert_springs_equal(got, want, first=0, count=None, label=""):
    gp, gx, gc = got
    wp, wx, wc = want
    count = len(wc) if count is None else count
    gc = gc[first:first + count]
    np.testing.assert_array_equal(gc, wc, err_msg=f"{label} spring counts")
    gp = gp[first:first + count]
    gx = gx[first:first + count]
    mask = np.arange(gp.shape[1])[None, :] < wc[:, None]
    gpl = np.where(gp >= 0, gp - first, gp)
    np.testing.assert_array_equal(np.where(mask, gpl, 0), np.where(mask, wp, 0),
                                  err_msg=f"{label} spring partners")
    np.testing.assert_array_equal(np.where(mask[..., None], gx, 0), np.where(mask[..., None], wx, 0),
                                  err_msg=f"{label} spring displacements")


def assert_cic_close(got, want, n, mass):
    got = got.astype(np.float64)
    tol = 1e-6 * np.abs(want).max() + 1e-6 * np.abs(want)
    assert np.all(np.abs(got - want) <= tol)
    assert got[: got.size // 4].sum() == pytest.approx(n * mass, rel=1e-6)


# ---- 32,768-grain column --------------------------------------------------

CASES_32K = [(0.5, False), (0.0, False), (0.5, True), (0.0, True)]


@pytest.mark.parametrize("mu,picard", CASES_32K, ids=["sie-mu0.5", "sie-mu0", "bep3-mu0.5", "bep3-mu0"])
def test_column_32k_from_lattice(oracle_lib, mu, picard):
    c = workloads.column_collapse_32k(friction=mu, picard=picard)
    st0, mat = _capi.initial_state(c)
    steps = 10
    o = oracle_from(oracle_lib, c, st0, mat)
    energies = []
    for _ in range(steps):
        energies.append(o.energy(32))
        o.steps(1, THREADS)
    with ParticleDynamics(c) as sim:
        sim.enable_energy_log(steps + 4)
        sim.take_step(steps)
        got = sim.state()
        np.testing.assert_array_equal(got, o.state())
        assert_springs_equal(sim.history(), o.history(), label="32k")
        assert sim.compute_max_acceleration() == o.max_accel()
        np.testing.assert_allclose(sim.energy_log(), energies, rtol=1e-12)
        sim.compute_density_grid()
        want = oracle_lib.density_grid(c, got, mat, 32)
        assert_cic_close(sim.host_density_grid, want, sim.n, workloads.GRAIN_MASS)
        assert sim.grid_overflow_count() == o.overflow() == 0


@pytest.mark.parametrize("mu,picard", CASES_32K, ids=["sie-mu0.5", "sie-mu0", "bep3-mu0.5", "bep3-mu0"])
def test_column_32k_mid_collapse(oracle_lib, mu, picard):
    """Steps 801..810 (the column has fallen and spread; several contacts
    per grain): the oracle continues from the GPU's state and springs."""
    c = workloads.column_collapse_32k(friction=mu, picard=picard)
    _, mat = _capi.initial_state(c)
    with ParticleDynamics(c) as sim:
        sim.enable_energy_log(900)
        sim.take_step(800)
        st = sim.state()
        hist = sim.history()
        passes, contacts = sim.contact_stats()
        o = oracle_from(oracle_lib, c, st, mat, hist)
        e0 = o.energy(32)
        assert sim.compute_total_energy() == pytest.approx(e0, rel=1e-6)
        sim.take_step(10)
        o.steps(10, THREADS)
        got = sim.state()
        np.testing.assert_array_equal(got, o.state())
        assert_springs_equal(sim.history(), o.history(), label="32k mid")
        assert sim.compute_max_acceleration() == o.max_accel()
        log = sim.energy_log()
        assert log[800] == pytest.approx(e0, rel=1e-12)
        sim.compute_density_grid()
        want = oracle_lib.density_grid(c, got, mat, 32)
        assert_cic_close(sim.host_density_grid, want, sim.n, workloads.GRAIN_MASS)
        assert sim.is_stable() and sim.grid_overflow_count() == 0
    # the window is contact-rich: the variants that spread contacts over
    # lanes have work to spread
    assert contacts / (passes * sim.n) > 1.0


# ---- 1,048,576-grain column -------------------------------------------------

@pytest.fixture(scope="module")
def column_1m():
    c = workloads.column_collapse_1m()
    st0, mat = _capi.initial_state(c)
    return c, st0, mat


def test_column_1m_cell_order_and_first_steps(oracle_lib, column_1m):
    c, st0, mat = column_1m
    with ParticleDynamics(c) as sim:
        assert sim.n == 1 << 20
        ids, keys = sim.cell_order()
        oid, okeys, _ = oracle_lib.cell_order(c, st0)
        np.testing.assert_array_equal(ids, oid)
        np.testing.assert_array_equal(keys.astype(np.int64), okeys)
        sim.take_step(3)
        o = oracle_from(oracle_lib, c, st0, mat)
        o.steps(3, THREADS)
        np.testing.assert_array_equal(sim.state(), o.state())
        assert sim.compute_max_acceleration() == o.max_accel()


def test_column_1m_collapse_window(oracle_lib, column_1m):
    """Steps 1..420 with whole-size invariants, then steps 421..422 (the
    column's foot at its densest) bit for bit from the GPU's state and
    springs."""
    c, st0, mat = column_1m
    with ParticleDynamics(c) as sim:
        sim.enable_energy_log(512)
        e0 = sim.compute_total_energy()
        sim.take_step(420)
        assert sim.is_stable()
        assert sim.grid_overflow_count() == 0
        log = sim.energy_log()
        assert len(log) == 420
        assert log.max() <= e0 + 1e-4 * abs(e0)  # dissipative contacts: never above the start
        assert log[-1] < log[0]
        passes, contacts = sim.contact_stats()
        assert contacts / (passes * sim.n) > 1.0
        st = sim.state()
        sim.compute_density_grid()
        grid = sim.host_density_grid.astype(np.float64)
        nodes = grid.size // 4
        assert grid[:nodes].sum() == pytest.approx(sim.n * workloads.GRAIN_MASS, rel=1e-6)
        # momentum channels against the state's total momentum
        p = st[:, 3:].astype(np.float64).sum(axis=0) * workloads.GRAIN_MASS
        for a in range(3):
            assert grid[(a + 1) * nodes:(a + 2) * nodes].sum() == pytest.approx(
                p[a], rel=1e-4, abs=1e-6 * sim.n * workloads.GRAIN_MASS)
        hist = sim.history()
        o = oracle_from(oracle_lib, c, st, mat, hist)
        sim.take_step(2)
        o.steps(2, THREADS)
        np.testing.assert_array_equal(sim.state(), o.state())
        assert_springs_equal(sim.history(), o.history(), label="1m")


# ---- 64-member uneven ensemble ----------------------------------------------

@pytest.fixture(scope="module")
def ensemble_64():
    cs = workloads.uneven_ensemble(64)
    inits = [_capi.initial_state(c) for c in cs]
    return cs, inits


def test_uneven_ensemble_64_members_match_oracle(oracle_lib, ensemble_64):
    cs, inits = ensemble_64
    with Ensemble(cs) as ens:
        assert ens.members == 64
        counts = [cnt for _, cnt in ens.ranges]
        assert min(counts) >= 2048 - 64 and max(counts) <= 32768 + 2048
        ens.take_step(10)
        got = ens.member_states()
        hist = ens.history()
        for m, (c, (st0, mat)) in enumerate(zip(cs, inits)):
            o = oracle_from(oracle_lib, c, st0, mat)
            o.steps(10, THREADS)
            np.testing.assert_array_equal(got[m], o.state(), err_msg=f"member {m}")
            first, count = ens.ranges[m]
            assert_springs_equal(hist, o.history(), first, count, label=f"member {m}")


def test_uneven_ensemble_64_mid_run_matches_oracle(oracle_lib, ensemble_64):
    """Steps 301..303 of every member from the batch's own state and springs
    at step 300 (columns falling and spreading at different rates)."""
    cs, inits = ensemble_64
    with Ensemble(cs) as ens:
        ens.take_step(300)
        st = ens.member_states()
        hp, hx, hc = ens.history()
        ens.take_step(3)
        got = ens.member_states()
        hist = ens.history()
        assert ens.member_stable().all()
        assert (ens.member_overflow_counts() == 0).all()
        for m, (c, (_, mat)) in enumerate(zip(cs, inits)):
            first, count = ens.ranges[m]
            p = hp[first:first + count]
            p = np.where(p >= 0, p - first, p)
            o = oracle_from(oracle_lib, c, st[m], mat, (p, hx[first:first + count], hc[first:first + count]))
            o.steps(3, THREADS)
            np.testing.assert_array_equal(got[m], o.state(), err_msg=f"member {m}")
            assert_springs_equal(hist, o.history(), first, count, label=f"member {m}")


# ---- 16-member uneven adjoint batch -----------------------------------------

def test_batched_adjoint_16_dot_product(oracle_lib):
    """The configured 16-member batch, one 100-step window from the released
    columns with a checkpoint every 10 steps; two members (the smallest and
    member 7) against the oracle's recording of their own window."""
    cs = workloads.uneven_ensemble(16, seed=77, n_min=4096, n_max=16384, aspect=(2.0, 2.0),
                                   mu=(0.3, 0.3))
    nodes = cs[0].density_grid_size_x * cs[0].density_grid_size_y * cs[0].density_grid_size_z
    rng = np.random.default_rng(21)
    W = rng.normal(size=(16, 4 * nodes))
    with Ensemble(cs) as ens:
        st = ens.member_states()
        ens.adjoint_record(100, 10)
        fin = ens.member_states()
        gs, gp = ens.adjoint_backward(W.astype(np.float32))
        ranges = list(ens.ranges)
    assert sum(c for _, c in ranges) > 16 * 4096
    for m in (int(np.argmin([c for _, c in ranges])), 7):
        first, count = ranges[m]
        c = cs[m]
        _, mat = _capi.initial_state(c)
        R = oracle_lib.Recording(c, st[m], mat, 100, threads=THREADS)
        np.testing.assert_array_equal(R.final_state(), fin[m])
        x0, v0 = st[m][:, :3].astype(np.float64), st[m][:, 3:].astype(np.float64)
        mu, k = float(c.friction), float(c.normal_stiffness)
        tx = rng.normal(size=x0.shape)
        tv = rng.normal(size=v0.shape)
        _, _, _, drho = R.replay(66, x0, v0, mu, k, tx, tv, 0.3, 0.7, threads=THREADS)
        jvp = W[m].astype(np.float32).astype(np.float64) @ drho
        g = gs[first:first + count].astype(np.float64)
        vjp = (g[:, :3] * tx).sum() + (g[:, 3:] * tv).sum() + gp[m, 0] * 0.3 + gp[m, 1] * 0.7
        print("member", m, jvp, vjp, flush=True)
        assert vjp == pytest.approx(jvp, rel=1e-4)
