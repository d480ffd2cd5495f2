"""Python mirror of the reference's ``ParticleDynamics`` class
(src/particle_dynamics.h:17-141) over the C ABI.

Same member names and semantics: public ``host_state`` / ``host_material`` /
``host_mass`` arrays, ``device_state.copy_from_host(host_state)`` to push a
host edit, ``dt`` re-read at every ``take_step``, FP32 ``time`` accumulated
one ``dt`` per step, and the on-demand diagnostics.  Used by the tests the way
the reference's GoogleTest suite uses the C++ class.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import WdConfig, check, lib, ptr

K_DIM = 3
K_STATE_STRIDE = 6
K_STENCIL_CELLS = 27
K_CHANNELS = 1 + K_DIM


class _DeviceVector:
    """Stand-in for ``DeviceVector<T>``: pushes a host array into the engine."""

    def __init__(self, sim: "ParticleDynamics", kind: str):
        self._sim = sim
        self._kind = kind

    def copy_from_host(self, host: np.ndarray) -> None:
        s = self._sim
        if self._kind == "state":
            a = np.ascontiguousarray(host, np.float32).reshape(-1)
            if a.size != K_STATE_STRIDE * s.n:
                raise ValueError("state must hold 6n floats")
            check(lib().wd_set_state(s._h, ptr(a)))
        else:
            a = np.ascontiguousarray(host, np.int32).reshape(-1)
            if a.size != s.n:
                raise ValueError("material must hold n ints")
            check(lib().wd_set_material(s._h, ptr(a)))


class ParticleDynamics:
    def __init__(self, config: WdConfig | None = None):
        self.config = (config or _capi.default_config()).copy()
        self._h = C.c_void_p()
        check(lib().wd_create(C.byref(self.config), C.byref(self._h)))
        self._setup()

    def _setup(self) -> None:
        n = C.c_int(0)
        check(lib().wd_num_grains(self._h, C.byref(n)))
        self.n = n.value
        self.dt = float(self.config.dt)
        self._dt_sent = self.dt
        self.particles_per_cell = self.config.particles_per_cell
        self.collision_grid_size_x = self.config.collision_grid_size_x
        self.collision_grid_size_y = self.config.collision_grid_size_y
        self.collision_grid_size_z = self.config.collision_grid_size_z
        self.density_grid_size_x = self.config.density_grid_size_x
        self.density_grid_size_y = self.config.density_grid_size_y
        self.density_grid_size_z = self.config.density_grid_size_z
        self.host_state = np.zeros(K_STATE_STRIDE * self.n, np.float32)
        self.host_material = np.zeros(self.n, np.int32)
        self.host_mass = np.array(self.config.get_masses(), np.float32)
        self.positions = np.zeros(K_DIM * self.n, np.float32)
        self.host_density_grid = np.zeros(K_CHANNELS * self.density_grid_nodes(), np.float32)
        self.device_state = _DeviceVector(self, "state")
        self.device_material = _DeviceVector(self, "material")
        if self.n:
            check(lib().wd_get_material(self._h, ptr(self.host_material)))
        self.unpack_state()

    # -- lifetime -------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().wd_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping -------------------------------------------------------
    def _push_dt(self) -> None:
        if self.dt != self._dt_sent:
            check(lib().wd_set_dt(self._h, C.c_float(self.dt)))
            self._dt_sent = self.dt

    def take_step(self, nsteps: int = 1) -> None:
        self._push_dt()
        check(lib().wd_step(self._h, int(nsteps)))

    @property
    def time(self) -> float:
        t = C.c_float(0)
        check(lib().wd_get_time(self._h, C.byref(t)))
        return t.value

    def sync(self) -> None:
        check(lib().wd_sync(self._h))

    # -- state access ---------------------------------------------------
    def unpack_state(self) -> None:
        if self.n:
            check(lib().wd_get_state(self._h, ptr(self.host_state)))
        self.positions = self.host_state.reshape(-1, K_STATE_STRIDE)[:, :K_DIM].reshape(-1).copy()

    def state(self) -> np.ndarray:
        self.unpack_state()
        return self.host_state.reshape(-1, K_STATE_STRIDE).copy()

    def set_state(self, state: np.ndarray) -> None:
        self.host_state[:] = np.asarray(state, np.float32).reshape(-1)
        self.device_state.copy_from_host(self.host_state)

    # -- diagnostics ----------------------------------------------------
    def density_grid_nodes(self) -> int:
        return self.density_grid_size_x * self.density_grid_size_y * self.density_grid_size_z

    def compute_density_grid(self) -> int:
        check(lib().wd_density_grid(self._h, ptr(self.host_density_grid)))
        return self.host_density_grid.size

    def compute_total_energy(self) -> float:
        e = C.c_float(0)
        check(lib().wd_total_energy(self._h, C.byref(e)))
        return e.value

    def enable_energy_log(self, capacity: int) -> None:
        check(lib().wd_energy_log_enable(self._h, int(capacity)))

    def energy_log(self) -> np.ndarray:
        rec = C.c_int(0)
        check(lib().wd_energy_log_read(self._h, None, 0, C.byref(rec)))
        out = np.zeros(max(rec.value, 1), np.float64)
        check(lib().wd_energy_log_read(self._h, ptr(out), rec.value, C.byref(rec)))
        return out[: rec.value]

    def enable_density_log(self, capacity: int) -> None:
        check(lib().wd_density_log_enable(self._h, int(capacity)))

    def record_density_grid(self) -> None:
        check(lib().wd_density_log_record(self._h))

    def density_log(self) -> np.ndarray:
        rec = C.c_int(0)
        check(lib().wd_density_log_read(self._h, None, 0, C.byref(rec)))
        out = np.zeros(max(rec.value, 1) * self.host_density_grid.size, np.float32)
        check(lib().wd_density_log_read(self._h, ptr(out), rec.value, C.byref(rec)))
        return out[: rec.value * self.host_density_grid.size].reshape(rec.value, self.host_density_grid.size)

    def profile_step(self, nsteps: int) -> dict[str, float]:
        """Mean ms per launch of each kernel of a step (HIP events on the
        engine's stream); advances the state by nsteps steps."""
        self._push_dt()
        ms = np.zeros(64, np.float32)
        cnt = C.c_int(0)
        names = C.create_string_buffer(2048)
        check(lib().wd_profile_step(self._h, int(nsteps), ptr(ms), ms.size, C.byref(cnt), names,
                                    len(names)))
        keys = names.value.decode().split(",") if cnt.value else []
        out: dict[str, float] = {}
        for k, v in zip(keys, ms[: cnt.value]):
            out[k] = out.get(k, 0.0) + float(v)
        return out

    def is_stable(self) -> bool:
        v = C.c_int(0)
        check(lib().wd_is_stable(self._h, C.byref(v)))
        return bool(v.value)

    def compute_max_acceleration(self) -> float:
        v = C.c_float(0)
        check(lib().wd_max_accel(self._h, C.byref(v)))
        return v.value

    def grid_overflow_count(self) -> int:
        v = C.c_int(0)
        check(lib().wd_overflow_count(self._h, C.byref(v)))
        return v.value

    def _stage(self, k: int) -> float:
        out = np.zeros(4, np.float32)
        check(lib().wd_stage_ms(self._h, ptr(out)))
        return float(out[k])

    def find_neighbors_wct(self) -> float:
        return self._stage(0)

    def compute_rhs_wct(self) -> float:
        return self._stage(1)

    def take_step_wct(self) -> float:
        return self._stage(2)

    def density_grid_wct(self) -> float:
        return self._stage(3)

    # -- inspection (parity tests) ----------------------------------------
    def cell_order(self) -> tuple[np.ndarray, np.ndarray]:
        ids = np.zeros(max(self.n, 1), np.int32)
        keys = np.zeros(max(self.n, 1), np.uint32)
        check(lib().wd_cell_order(self._h, ptr(ids), ptr(keys)))
        return ids[: self.n], keys[: self.n]

    def cell_order_noalloc(self) -> tuple[np.ndarray, np.ndarray]:
        """cell_order() staged through the library's fixed device array
        (wd_cell_order_noalloc): no device allocation per call."""
        ids = np.zeros(max(self.n, 1), np.int32)
        keys = np.zeros(max(self.n, 1), np.uint32)
        check(lib().wd_cell_order_noalloc(self._h, ptr(ids), ptr(keys)))
        return ids[: self.n], keys[: self.n]

    def neighbor_rows(self) -> np.ndarray:
        row = K_STENCIL_CELLS * self.particles_per_cell
        out = np.full(max(self.n, 1) * row, -1, np.int32)
        check(lib().wd_neighbor_rows(self._h, ptr(out), out.size))
        return out[: self.n * row].reshape(self.n, row)

    def history(self):
        """Stored contact springs (spring_history model), original order:
        partner ids [n, S], displacements [n, S, 3], counts [n]."""
        S = _capi.WD_HISTORY_SLOTS
        partner = np.zeros(max(self.n, 1) * S, np.int32)
        xi = np.zeros(max(self.n, 1) * S * 3, np.float32)
        cnt = np.zeros(max(self.n, 1), np.int32)
        check(lib().wd_get_history(self._h, ptr(partner), ptr(xi), ptr(cnt)))
        n = self.n
        return partner[: n * S].reshape(n, S), xi[: n * S * 3].reshape(n, S, 3), cnt[:n]

    # -- launch configuration (extension) ---------------------------------
    def set_force_block(self, block: int) -> None:
        """0: autotune on the next long take_step; 64 / 128 / 256: fixed,
        one lane per grain."""
        check(lib().wd_set_force_block(self._h, int(block)))

    def force_block(self) -> int:
        v = C.c_int(0)
        check(lib().wd_get_force_block(self._h, C.byref(v)))
        return v.value

    def set_force_variant(self, block: int, lanes: int) -> None:
        """(0, 0): autotune; (64|128|256, 0): one lane per grain; (64, 1|2|4):
        one-wave blocks with contacts spread over the lanes, `lanes` per grain."""
        check(lib().wd_set_force_variant(self._h, int(block), int(lanes)))

    def force_variant(self) -> tuple[int, int]:
        b, l = C.c_int(0), C.c_int(0)
        check(lib().wd_get_force_variant(self._h, C.byref(b), C.byref(l)))
        return b.value, l.value

    def set_scan_tile(self, cells: int) -> None:
        """0: tuned online with the force variant; 1024 / 2048 / 4096 cells
        per tile of the cell scan: fixed."""
        check(lib().wd_set_scan_tile(self._h, int(cells)))

    def scan_tile(self) -> int:
        v = C.c_int(0)
        check(lib().wd_get_scan_tile(self._h, C.byref(v)))
        return v.value

    def set_binning(self, mode: int) -> None:
        """0: tuned online; 1: scatter pass; 2: per-cell buckets (same cell
        order and results either way)."""
        check(lib().wd_set_binning(self._h, int(mode)))

    def binning(self) -> int:
        v = C.c_int(0)
        check(lib().wd_get_binning(self._h, C.byref(v)))
        return v.value

    def set_fused_steps(self, on: bool) -> None:
        """Fused steps (the default for a single simulation): the force pass
        counts its grains for the next bin and the cell scan reduces its
        partials, three launches per step instead of six; False selects the
        classic sequence.  Results are the same either way."""
        check(lib().wd_set_fused_steps(self._h, 1 if on else 0))

    def set_graph_replay(self, on: bool) -> None:
        """Steps replay captured graphs (default); False launches their
        kernels one by one (for tools that follow launches)."""
        check(lib().wd_set_graph_replay(self._h, 1 if on else 0))

    def fused_steps(self) -> bool:
        v = C.c_int(0)
        check(lib().wd_get_fused_steps(self._h, C.byref(v)))
        return bool(v.value)

    def fused_misses(self) -> int:
        """Fused steps whose grains were binned by the scan's one-block
        fallback (a grain moved more than one cell in the step)."""
        v = C.c_ulonglong(0)
        check(lib().wd_fused_misses(self._h, C.byref(v)))
        return v.value

    def set_cell_scale(self, scale: float) -> None:
        """Opt-in coarser collision cells (1: the configured grid)."""
        check(lib().wd_set_cell_scale(self._h, C.c_float(scale)))

    def cell_scale(self) -> tuple[float, tuple[int, int, int]]:
        """The cell scale and the collision grid it gives."""
        v = C.c_float(0)
        g = np.zeros(3, np.int32)
        check(lib().wd_get_cell_scale(self._h, C.byref(v), ptr(g)))
        return v.value, (int(g[0]), int(g[1]), int(g[2]))

    def tune_cell_scale(self, steps: int) -> float:
        """Times `steps` real steps on each of the scales 1, 1.25, 1.5 and
        keeps the fastest that overflowed no cell (advances 3 * steps)."""
        v = C.c_float(0)
        check(lib().wd_tune_cell_scale(self._h, int(steps), C.byref(v)))
        return v.value

    def contact_stats(self) -> tuple[int, int]:
        """(force passes, contacts summed over them) since construction."""
        p, c = C.c_ulonglong(0), C.c_ulonglong(0)
        check(lib().wd_contact_stats(self._h, C.byref(p), C.byref(c)))
        return p.value, c.value

    # -- parameters and the adjoint (extension) ----------------------------
    def set_friction(self, mu: float, member: int = 0) -> None:
        check(lib().wd_set_friction(self._h, int(member), C.c_float(mu)))

    def set_normal_stiffness(self, k_n: float) -> None:
        check(lib().wd_set_normal_stiffness(self._h, C.c_float(k_n)))

    def adjoint_record(self, nsteps: int, checkpoint_every: int) -> None:
        self._push_dt()
        check(lib().wd_adjoint_record(self._h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, grid_cotangent=None, state_cotangent=None):
        """dL/d(initial state) [n, 6] and per-member [dL/dmu, dL/dk_n]."""
        members = getattr(self, "members", 1)
        g = None
        if grid_cotangent is not None:
            g = np.ascontiguousarray(grid_cotangent, np.float32).reshape(-1)
            if g.size != members * K_CHANNELS * self.density_grid_nodes():
                raise ValueError("grid cotangent must hold members * 4 * nodes floats")
        sc = None
        if state_cotangent is not None:
            sc = np.ascontiguousarray(state_cotangent, np.float32).reshape(-1)
            if sc.size != K_STATE_STRIDE * self.n:
                raise ValueError("state cotangent must hold 6n floats")
        gs = np.zeros(K_STATE_STRIDE * max(self.n, 1), np.float32)
        gp = np.zeros(2 * members, np.float64)
        check(lib().wd_adjoint_backward(self._h, None if g is None else ptr(g),
                                        None if sc is None else ptr(sc), ptr(gs), ptr(gp)))
        return gs[: K_STATE_STRIDE * self.n].reshape(self.n, K_STATE_STRIDE), gp.reshape(members, 2)

    # -- device buffers: torch tensors on the handle's GPU ----------------
    def _device_tensor(self, t, size: int, what: str):
        import torch

        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise TypeError(f"{what} must be a contiguous float32 tensor on the GPU")
        if t.numel() != size:
            raise ValueError(f"{what} must hold {size} floats")
        return C.c_void_p(t.data_ptr())

    def set_state_device(self, state) -> None:
        """The grain state from a device tensor of 6n floats (original
        order), without a host copy."""
        check(lib().wd_set_state_device(self._h, self._device_tensor(state, K_STATE_STRIDE * self.n, "state")))

    def density_grids_device(self, out) -> None:
        """Every member's CIC grids of the current state into a device tensor
        of members * 4 * nodes floats."""
        members = getattr(self, "members", 1)
        size = members * K_CHANNELS * self.density_grid_nodes()
        check(lib().wd_batch_density_grid_device(self._h, -1, self._device_tensor(out, size, "out")))

    def adjoint_backward_device(self, grid_cotangent=None, state_cotangent=None):
        """adjoint_backward with device tensors: the cotangents in, dL/d(initial
        state) [n, 6] out on the device; [dL/dmu, dL/dk_n] per member (host)."""
        import torch

        members = getattr(self, "members", 1)
        g = None if grid_cotangent is None else self._device_tensor(
            grid_cotangent, members * K_CHANNELS * self.density_grid_nodes(), "grid cotangent")
        sc = None if state_cotangent is None else self._device_tensor(
            state_cotangent, K_STATE_STRIDE * self.n, "state cotangent")
        if grid_cotangent is None and state_cotangent is None:
            raise ValueError("adjoint_backward_device needs a grid or a state cotangent")
        dev = (grid_cotangent if grid_cotangent is not None else state_cotangent).device
        gs = torch.empty((self.n, K_STATE_STRIDE), dtype=torch.float32, device=dev)
        gp = np.zeros(2 * members, np.float64)
        check(lib().wd_adjoint_backward_device(self._h, g, sc, C.c_void_p(gs.data_ptr()), ptr(gp)))
        return gs, gp.reshape(members, 2)

    def rhs(self) -> np.ndarray:
        out = np.zeros(K_STATE_STRIDE * max(self.n, 1), np.float32)
        check(lib().wd_rhs(self._h, ptr(out)))
        return out[: K_STATE_STRIDE * self.n].reshape(self.n, K_STATE_STRIDE)


class Ensemble(ParticleDynamics):
    """E members advanced together by one handle (wd_batch_create).

    The reference runs ensemble members one after another in separate
    ParticleDynamics objects and requires equal grain counts
    (src/chaos.cpp:163-171); here members may differ in grain count,
    initialization and friction, and each evolves exactly as it would alone.
    The inherited arrays (host_state, host_material, ...) hold the members'
    grains back to back; member_* accessors slice one member out."""

    def __init__(self, configs):
        configs = [c.copy() for c in configs]
        if not configs:
            raise ValueError("an ensemble needs at least one member")
        arr = (WdConfig * len(configs))(*configs)
        self.configs = configs
        self.config = configs[0].copy()
        self._h = C.c_void_p()
        check(lib().wd_batch_create(arr, len(configs), C.byref(self._h)))
        self._setup()
        self.members = len(configs)
        self.ranges = []
        for m in range(self.members):
            first, count = C.c_int(0), C.c_int(0)
            check(lib().wd_batch_member_range(self._h, m, C.byref(first), C.byref(count)))
            self.ranges.append((first.value, count.value))

    def member_state(self, m: int) -> np.ndarray:
        """State of member m (its grains in their own original order)."""
        first, count = self.ranges[m]
        return self.state()[first:first + count]

    def member_states(self) -> list[np.ndarray]:
        st = self.state()
        return [st[f:f + c] for f, c in self.ranges]

    def member_density_grid(self, m: int) -> np.ndarray:
        out = np.zeros(K_CHANNELS * self.density_grid_nodes(), np.float32)
        check(lib().wd_batch_density_grid(self._h, int(m), ptr(out)))
        return out

    def member_overflow_counts(self) -> np.ndarray:
        out = np.zeros(self.members, np.int32)
        check(lib().wd_batch_overflow_counts(self._h, ptr(out)))
        return out

    def member_stable(self) -> np.ndarray:
        out = np.zeros(self.members, np.int32)
        check(lib().wd_batch_is_stable(self._h, ptr(out)))
        return out.astype(bool)

    def density_grids(self) -> np.ndarray:
        """All members' CIC grids, [members, 4 * nodes]."""
        out = np.zeros(self.members * K_CHANNELS * self.density_grid_nodes(), np.float32)
        check(lib().wd_batch_density_grid(self._h, -1, ptr(out)))
        return out.reshape(self.members, -1)
