"""Batched uneven ensemble: every member of one batch handle evolves bit for
bit as the same configuration run alone.

The reference runs members sequentially and requires equal grain counts
(src/chaos.cpp:163-171).  Here members differ in grain count (8 to 4,096,
so that 256-grain blocks span up to four members), initialization and
friction; their states, contact springs and CIC grids must equal the single
runs' exactly, and the single runs are pinned to the oracle by
test_gpu_parity.py."""
import numpy as np
import pytest

from world_amd import _capi, workloads
from world_amd.sim import Ensemble, ParticleDynamics

pytestmark = pytest.mark.gpu


def members(integrator=_capi.WD_SEMI_IMPLICIT_EULER):
    cs = [
        workloads.column(3, 3, 3, friction=0.2),
        workloads.column(2, 2, 2, friction=0.0),
        workloads.column(4, 4, 5, friction=0.6),
        workloads.perturbed(workloads.column_collapse_4k(), 50e-6, 1001),
        workloads.column(7, 7, 20, friction=0.4, seed=77),
    ]
    # members may also differ in restitution and grain mass (the dataset
    # driver's design columns)
    cs[2].restitution = 0.8
    cs[4].set_masses([0.0, 2.5 * workloads.GRAIN_MASS])
    cs[4].restitution = 0.3
    for c in cs:
        c.time_integrator = integrator
    return cs


def jittered(c, seed):
    st, mat = _capi.initial_state(c)
    st = st.reshape(-1, 6).copy()
    rng = np.random.default_rng(seed)
    st[:, :3] += rng.normal(0, 0.0015, (len(st), 3)).astype(np.float32)
    st[:, 3:] = rng.normal(0, 0.2, (len(st), 3)).astype(np.float32)
    return st


def singles(cs, nsteps, states=None):
    out = []
    for m, c in enumerate(cs):
        with ParticleDynamics(c) as sim:
            if states is not None:
                sim.set_state(states[m])
            sim.take_step(nsteps)
            sim.compute_density_grid()
            out.append(dict(state=sim.state(), history=sim.history(),
                            grid=sim.host_density_grid.copy(), amax=sim.compute_max_acceleration(),
                            energy=sim.compute_total_energy()))
    return out


def check_members(ens, want):
    got = ens.member_states()
    partner, xi, cnt = ens.history()
    for m, (first, count) in enumerate(ens.ranges):
        np.testing.assert_array_equal(got[m], want[m]["state"], err_msg=f"member {m}")
        wp, wx, wc = want[m]["history"]
        np.testing.assert_array_equal(cnt[first:first + count], wc)
        for i in range(count):
            k = wc[i]
            p = partner[first + i, :k]
            p = np.where(p >= 0, p - first, p)
            np.testing.assert_array_equal(p, wp[i, :k])
            np.testing.assert_array_equal(xi[first + i, :k], wx[i, :k])
        np.testing.assert_array_equal(ens.member_density_grid(m), want[m]["grid"])
    assert ens.compute_max_acceleration() == max(w["amax"] for w in want)


@pytest.mark.parametrize("integrator", [_capi.WD_SEMI_IMPLICIT_EULER, _capi.WD_BACKWARD_EULER_PICARD])
def test_batch_from_initialization_equals_single_runs(integrator):
    cs = members(integrator)
    want = singles(cs, 40)
    with Ensemble(cs) as ens:
        assert ens.members == len(cs)
        assert [c for _, c in ens.ranges] == [len(w["state"]) for w in want]
        ens.take_step(40)
        check_members(ens, want)
        e = ens.compute_total_energy()
        assert e == pytest.approx(sum(w["energy"] for w in want), rel=1e-5)


def test_batch_contact_rich_equals_single_runs():
    cs = members()
    states = [jittered(c, 10 + m) for m, c in enumerate(cs)]
    want = singles(cs, 37, states)
    with Ensemble(cs) as ens:
        ens.set_state(np.concatenate(states))
        for k in (1, 16, 20):  # eager, one graph chunk, chunk + remainder
            ens.take_step(k)
        check_members(ens, want)


def test_batch_density_grids_block():
    cs = members()
    with Ensemble(cs) as ens:
        ens.take_step(5)
        allg = ens.density_grids()
        for m in range(ens.members):
            np.testing.assert_array_equal(allg[m], ens.member_density_grid(m))
        ens.compute_density_grid()
        np.testing.assert_array_equal(ens.host_density_grid, allg[0])


def test_member_overflow_and_stability():
    """Per-member lifetime overflow counts and stability flags equal each
    member's single-run diagnostics."""
    cs = members()
    # one grain per cell: cells that two grains share as the columns settle
    # overflow
    for c in cs:
        c.particles_per_cell = 1
    states = [jittered(c, 40 + m) for m, c in enumerate(cs)]
    states[0][0, 0] = 10.0  # member 0 leaves the domain: unstable
    want_over, want_stable = [], []
    for c, st in zip(cs, states):
        with ParticleDynamics(c) as sim:
            sim.set_state(st)
            sim.take_step(30)
            want_over.append(sim.grid_overflow_count())
            want_stable.append(sim.is_stable())
    with Ensemble(cs) as ens:
        ens.set_state(np.concatenate(states))
        ens.take_step(30)
        np.testing.assert_array_equal(ens.member_overflow_counts(), want_over)
        np.testing.assert_array_equal(ens.member_stable(), want_stable)
        assert ens.grid_overflow_count() == sum(want_over)
    assert sum(want_over) > 0 and not want_stable[0] and all(want_stable[1:])


def test_batch_parameter_setters_equal_created_parameters():
    """Friction per member and the normal stiffness set on a live batch
    handle (the autograd window's forward does this) give the same steps,
    bit for bit, as a batch created with those parameters.  (Setting them
    once dropped the member partials' buffer from the solver constants, and
    the next reduction read through a null pointer.)"""
    cs = members()
    want_cfgs = [c.copy() for c in cs]
    for m, c in enumerate(want_cfgs):
        c.friction = 0.1 * m
        c.normal_stiffness = 180.0
    with Ensemble(cs) as ens, Ensemble(want_cfgs) as want:
        for m, c in enumerate(want_cfgs):
            ens.set_friction(c.friction, m)
        ens.set_normal_stiffness(180.0)
        ens.take_step(25)
        want.take_step(25)
        np.testing.assert_array_equal(ens.state(), want.state())
        assert ens.member_stable().all()
