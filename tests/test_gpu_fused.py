"""Fused steps against the classic launch sequence and the CPU oracle.

A fused step is the force pass counting its grains into the next step's cells
(a grid of the last occupied box widened by one cell), the cell scan reducing
the pass's block partials in its last block, and the reorder.  It must give
the classic sequence's bits for every force variant, binning, integrator and
spring model; a grain that moves more than one cell in a step goes through the
scan's one-block fallback and must still give the oracle's bits; the two
sequences can alternate within a run."""
import numpy as np
import pytest

from world_amd import _capi, workloads
from world_amd.sim import Ensemble, ParticleDynamics

from .test_gpu_parity import FORCE_VARIANTS, assert_springs_equal, history_cfg, jittered, viscous_cfg

pytestmark = pytest.mark.gpu


def run(c, st, steps, fused, variant=None, binning=None, energy=False):
    with ParticleDynamics(c) as sim:
        sim.set_fused_steps(fused)
        if variant is not None:
            sim.set_force_variant(*variant)
        if binning is not None:
            sim.set_binning(binning)
        if energy:
            sim.enable_energy_log(steps)
        sim.set_state(st)
        sim.take_step(steps)
        out = {
            "state": sim.state(),
            "accel": sim.compute_max_acceleration(),
            "overflow": sim.grid_overflow_count(),
            "fused": sim.fused_steps(),
            "misses": sim.fused_misses(),
            "hist": sim.history() if c.tangential_model == _capi.WD_TANGENTIAL_SPRING_HISTORY else None,
            "energy": sim.energy_log() if energy else None,
            "contacts": sim.contact_stats(),
        }
    return out


def assert_same(got, want):
    np.testing.assert_array_equal(got["state"], want["state"])
    assert got["accel"] == want["accel"]
    assert got["overflow"] == want["overflow"]
    assert got["contacts"] == want["contacts"]
    if want["hist"] is not None:
        assert_springs_equal(got["hist"], want["hist"])
    if want["energy"] is not None:
        np.testing.assert_array_equal(got["energy"], want["energy"])


@pytest.mark.parametrize("binning", [_capi.WD_BINNING_SCATTER, _capi.WD_BINNING_BUCKETS])
@pytest.mark.parametrize("variant", FORCE_VARIANTS)
def test_fused_equals_classic_for_every_variant(variant, binning):
    c = workloads.column_collapse_4k(friction=0.5)
    st, _ = jittered(c, sigma_x=0.002, sigma_v=0.3, seed=61)
    classic = run(c, st, 30, False, variant, binning, energy=True)
    fused = run(c, st, 30, True, variant, binning, energy=True)
    assert fused["fused"] and not classic["fused"]
    assert fused["misses"] == 0
    assert_same(fused, classic)


@pytest.mark.parametrize("picard", [False, True])
@pytest.mark.parametrize("model", ["history", "viscous"])
def test_fused_trajectory_matches_oracle(oracle_lib, picard, model):
    c = workloads.column_collapse_4k(
        friction=0.5, integrator=_capi.WD_BACKWARD_EULER_PICARD if picard else _capi.WD_SEMI_IMPLICIT_EULER)
    c = history_cfg(c) if model == "history" else viscous_cfg(c)
    st, mat = jittered(c, sigma_x=0.002, sigma_v=0.3, seed=62)
    steps = 12 if picard else 40
    fused = run(c, st, steps, True, energy=True)
    classic = run(c, st, steps, False, energy=True)
    assert fused["misses"] == 0
    assert_same(fused, classic)
    want, over, amax = oracle_lib.steps(c, st, mat, steps, threads=8)
    np.testing.assert_array_equal(fused["state"], want)
    assert fused["overflow"] == over


def fast_grains(c, seed):
    """A jittered column with a few grains thrown sideways at 25 m/s: each
    moves 2.5 cells per step, out of the widened box of the step before."""
    st, mat = jittered(c, sigma_x=0.002, sigma_v=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(st), 12, replace=False)
    st[pick, 3] = np.where(np.arange(12) % 2 == 0, 25.0, -25.0).astype(np.float32)
    st[pick, 4] = 0.0
    st[pick, 5] = 0.0
    return st, mat


@pytest.mark.parametrize("binning", [_capi.WD_BINNING_SCATTER, _capi.WD_BINNING_BUCKETS])
def test_grain_past_the_widened_box_takes_the_fallback(oracle_lib, binning):
    c = workloads.column_collapse_4k(friction=0.5)
    st, mat = fast_grains(c, seed=63)
    fused = run(c, st, 20, True, binning=binning, energy=True)
    classic = run(c, st, 20, False, binning=binning, energy=True)
    assert fused["misses"] > 0  # the scan binned those steps on one block
    assert_same(fused, classic)
    want, over, _ = oracle_lib.steps(c, st, mat, 20, threads=8)
    np.testing.assert_array_equal(fused["state"], want)


def test_fused_and_classic_steps_alternate():
    """Fused and classic steps, deposits, energy and inspection calls in any
    order give the steps of one classic run."""
    c = workloads.column_collapse_4k(friction=0.5)
    st, _ = fast_grains(c, seed=64)
    want = run(c, st, 31, False)
    with ParticleDynamics(c) as sim:
        sim.set_state(st)
        sim.take_step(7)
        sim.compute_density_grid()
        sim.set_fused_steps(False)
        sim.take_step(5)
        sim.set_fused_steps(True)
        sim.compute_total_energy()
        sim.take_step(1)
        sim.cell_order()
        sim.take_step(9)
        sim.neighbor_rows()
        sim.take_step(9)
        np.testing.assert_array_equal(sim.state(), want["state"])
        assert_springs_equal(sim.history(), want["hist"])


def test_ensembles_run_the_classic_sequence():
    cfgs = [workloads.column(8, 8, 8, seed=s) for s in (1, 2)]
    with Ensemble(cfgs) as ens:
        assert not ens.fused_steps()


@pytest.mark.parametrize("picard", [False, True])
def test_overflow_totals_count_the_steps_taken(picard):
    """Cells over capacity (two grains per cell allowed): a fused step bins
    the next state ahead of time, but the totals count the neighbor passes of
    the steps taken (and of inspection calls) exactly as the classic sequence
    does, however the steps are split into calls."""
    c = workloads.column_collapse_4k(
        friction=0.5, integrator=_capi.WD_BACKWARD_EULER_PICARD if picard else _capi.WD_SEMI_IMPLICIT_EULER)
    c.particles_per_cell = 2
    st, _ = jittered(c, sigma_x=0.003, sigma_v=0.3, seed=65)
    totals = {}
    for fused in (False, True):
        with ParticleDynamics(c) as sim:
            sim.set_fused_steps(fused)
            sim.set_state(st)
            seen = []
            for k in (1, 4, 1, 7):
                sim.take_step(k)
                seen.append(sim.grid_overflow_count())
            sim.cell_order()  # an inspection bin counts as in the classic sequence
            seen.append(sim.grid_overflow_count())
            sim.take_step(3)
            seen.append(sim.grid_overflow_count())
            totals[fused] = (seen, sim.state())
    assert totals[False][0][-1] > 0
    assert totals[True][0] == totals[False][0]
    np.testing.assert_array_equal(totals[True][1], totals[False][1])


def test_eager_launches_equal_graph_replay():
    """Fused steps launched kernel by kernel (graph replay off) give the
    replayed steps' bits."""
    c = workloads.column_collapse_4k(friction=0.5)
    st, _ = jittered(c, sigma_x=0.002, sigma_v=0.3, seed=66)
    want = run(c, st, 25, True, energy=True)
    with ParticleDynamics(c) as sim:
        sim.set_graph_replay(False)
        sim.enable_energy_log(25)
        sim.set_state(st)
        sim.take_step(25)
        np.testing.assert_array_equal(sim.state(), want["state"])
        np.testing.assert_array_equal(sim.energy_log(), want["energy"])
        assert_springs_equal(sim.history(), want["hist"])
