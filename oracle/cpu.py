"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker (and the timed CPU baseline),
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_LIB_PATH = HERE / "_ref" / "libref_config.so"

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        from world_amd._capi import WdConfig  # struct layout only

        h = C.CDLL(str(LIB_PATH))
        P = C.POINTER(WdConfig)
        vp = C.c_void_p
        h.or_initial_state.argtypes = [P, vp, vp, C.c_int, C.POINTER(C.c_int)]
        h.or_initial_state.restype = C.c_int
        h.or_cell_keys.argtypes = [P, C.c_int, vp, vp]
        h.or_cell_keys.restype = None
        h.or_cell_order.argtypes = [P, C.c_int, vp, vp, vp]
        h.or_cell_order.restype = C.c_longlong
        h.or_neighbor_rows.argtypes = [P, C.c_int, vp, vp]
        h.or_neighbor_rows.restype = C.c_longlong
        h.or_rhs.argtypes = [P, C.c_int, vp, vp, C.c_int, vp, vp]
        h.or_rhs.restype = None
        h.or_steps.argtypes = [P, C.c_int, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
        h.or_steps.restype = C.c_longlong
        h.or_energy.argtypes = [P, C.c_int, vp, vp, C.c_int]
        h.or_energy.restype = C.c_double
        h.or_grain_energy.argtypes = [P, C.c_int, vp, vp, vp]
        h.or_grain_energy.restype = None
        h.or_density_grid.argtypes = [P, C.c_int, vp, vp, C.c_int, vp]
        h.or_density_grid.restype = None
        h.or_sim_create.argtypes = [P, C.c_int, vp, vp]
        h.or_sim_create.restype = vp
        h.or_sim_destroy.argtypes = [vp]
        h.or_sim_destroy.restype = None
        h.or_sim_set_dt.argtypes = [vp, C.c_float]
        h.or_sim_set_dt.restype = None
        h.or_sim_steps.argtypes = [vp, C.c_int, C.c_int]
        h.or_sim_steps.restype = None
        h.or_sim_state.argtypes = [vp, vp]
        h.or_sim_state.restype = None
        h.or_sim_overflow.argtypes = [vp]
        h.or_sim_overflow.restype = C.c_longlong
        h.or_sim_max_accel.argtypes = [vp]
        h.or_sim_max_accel.restype = C.c_float
        h.or_sim_history.argtypes = [vp, vp, vp, vp]
        h.or_sim_history.restype = None
        h.or_sim_set_history.argtypes = [vp, vp, vp, vp]
        h.or_sim_set_history.restype = None
        h.or_sim_energy.argtypes = [vp, C.c_int]
        h.or_sim_energy.restype = C.c_double
        h.or_rec_create.argtypes = [P, C.c_int, vp, vp, C.c_int, C.c_int]
        h.or_rec_create.restype = vp
        h.or_rec_destroy.argtypes = [vp]
        h.or_rec_destroy.restype = None
        h.or_rec_state.argtypes = [vp, vp]
        h.or_rec_state.restype = None
        h.or_rec_contacts.argtypes = [vp, C.c_int, vp, vp, vp]
        h.or_rec_contacts.restype = C.c_longlong
        h.or_rec_replay.argtypes = [vp, C.c_int, vp, vp, C.c_double, C.c_double, vp, vp,
                                    C.c_double, C.c_double, vp, vp, vp, vp, C.c_int]
        h.or_rec_replay.restype = None
        _lib = h
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def _st(state) -> np.ndarray:
    return np.ascontiguousarray(state, np.float32).reshape(-1)


def initial_state(cfg) -> tuple[np.ndarray, np.ndarray]:
    n = C.c_int(0)
    lib().or_initial_state(C.byref(cfg), None, None, 0, C.byref(n))
    st = np.zeros(6 * max(n.value, 1), np.float32)
    mat = np.zeros(max(n.value, 1), np.int32)
    rc = lib().or_initial_state(C.byref(cfg), _p(st), _p(mat), n.value, C.byref(n))
    assert rc == 0
    return st[: 6 * n.value], mat[: n.value]


def cell_keys(cfg, state) -> np.ndarray:
    st = _st(state)
    n = st.size // 6
    out = np.zeros(max(n, 1), np.int64)
    lib().or_cell_keys(C.byref(cfg), n, _p(st), _p(out))
    return out[:n]


def cell_order(cfg, state) -> tuple[np.ndarray, np.ndarray, int]:
    st = _st(state)
    n = st.size // 6
    ids = np.zeros(max(n, 1), np.int32)
    keys = np.zeros(max(n, 1), np.int64)
    over = lib().or_cell_order(C.byref(cfg), n, _p(st), _p(ids), _p(keys))
    return ids[:n], keys[:n], int(over)


def neighbor_rows(cfg, state) -> tuple[np.ndarray, int]:
    st = _st(state)
    n = st.size // 6
    row = 27 * cfg.particles_per_cell
    out = np.full(max(n, 1) * row, -1, np.int32)
    over = lib().or_neighbor_rows(C.byref(cfg), n, _p(st), _p(out))
    return out[: n * row].reshape(n, row), int(over)


def rhs(cfg, state, material, precision: int = 32) -> tuple[np.ndarray, np.ndarray]:
    st = _st(state)
    n = st.size // 6
    mat = np.ascontiguousarray(material, np.int32)
    out = np.zeros(6 * max(n, 1), np.float64)
    scale = np.zeros(max(n, 1), np.float64)
    lib().or_rhs(C.byref(cfg), n, _p(st), _p(mat), precision, _p(out), _p(scale))
    return out[: 6 * n].reshape(n, 6), scale[:n]


def steps(cfg, state, material, nsteps: int, threads: int = 1) -> tuple[np.ndarray, int, float]:
    st = _st(state).copy()
    n = st.size // 6
    mat = np.ascontiguousarray(material, np.int32)
    amax = C.c_float(0)
    over = lib().or_steps(C.byref(cfg), n, _p(st), _p(mat), int(nsteps), int(threads), C.byref(amax))
    return st.reshape(n, 6), int(over), amax.value


def energy(cfg, state, material, precision: int = 32) -> float:
    st = _st(state)
    mat = np.ascontiguousarray(material, np.int32)
    return float(lib().or_energy(C.byref(cfg), st.size // 6, _p(st), _p(mat), precision))


def grain_energy(cfg, state, material) -> np.ndarray:
    st = _st(state)
    n = st.size // 6
    mat = np.ascontiguousarray(material, np.int32)
    out = np.zeros(max(n, 1), np.float32)
    lib().or_grain_energy(C.byref(cfg), n, _p(st), _p(mat), _p(out))
    return out[:n]


def density_grid(cfg, state, material, precision: int = 32) -> np.ndarray:
    st = _st(state)
    n = st.size // 6
    mat = np.ascontiguousarray(material, np.int32)
    nodes = cfg.density_grid_size_x * cfg.density_grid_size_y * cfg.density_grid_size_z
    out = np.zeros(4 * nodes, np.float64)
    lib().or_density_grid(C.byref(cfg), n, _p(st), _p(mat), precision, _p(out))
    return out


HIST_SLOTS = 24


class Sim:
    """Stateful oracle simulation: FP32 state in original order plus the
    contact history of the spring_history extension."""

    def __init__(self, cfg, state, material):
        st = _st(state)
        self.n = st.size // 6
        self.cfg = cfg.copy()
        mat = np.ascontiguousarray(material, np.int32)
        self._h = lib().or_sim_create(C.byref(self.cfg), self.n, _p(st), _p(mat))

    def close(self):
        if self._h:
            lib().or_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_dt(self, dt: float) -> None:
        lib().or_sim_set_dt(self._h, dt)

    def steps(self, k: int, threads: int = 8) -> None:
        lib().or_sim_steps(self._h, int(k), int(threads))

    def state(self) -> np.ndarray:
        out = np.zeros(6 * max(self.n, 1), np.float32)
        lib().or_sim_state(self._h, _p(out))
        return out[: 6 * self.n].reshape(self.n, 6)

    def overflow(self) -> int:
        return int(lib().or_sim_overflow(self._h))

    def max_accel(self) -> float:
        return float(lib().or_sim_max_accel(self._h))

    def energy(self, precision: int = 32) -> float:
        return float(lib().or_sim_energy(self._h, precision))

    def history(self):
        partner = np.zeros(max(self.n, 1) * HIST_SLOTS, np.int32)
        xi = np.zeros(max(self.n, 1) * HIST_SLOTS * 3, np.float32)
        cnt = np.zeros(max(self.n, 1), np.int32)
        lib().or_sim_history(self._h, _p(partner), _p(xi), _p(cnt))
        n = self.n
        return (partner[: n * HIST_SLOTS].reshape(n, HIST_SLOTS),
                xi[: n * HIST_SLOTS * 3].reshape(n, HIST_SLOTS, 3), cnt[:n])

    def set_history(self, partner, xi, count) -> None:
        """Replaces the stored springs (the layout history() returns), to
        continue from a state and springs taken mid-run."""
        p = np.ascontiguousarray(partner, np.int32).reshape(-1)
        x = np.ascontiguousarray(xi, np.float32).reshape(-1)
        c = np.ascontiguousarray(count, np.int32).reshape(-1)
        assert p.size == self.n * HIST_SLOTS and x.size == 3 * p.size and c.size == self.n
        lib().or_sim_set_history(self._h, _p(p), _p(x), _p(c))


class Recording:
    """An FP32 SIE trajectory with contact history whose per-step contact sets
    and clamp / slip decisions are recorded, and replayed with those branches
    frozen: in FP32 (bit-identical to the forward), FP64 (the function that
    central finite differences sample), FP64 with one forward-mode tangent
    (the JVP), or the FP32 primal with an FP64 tangent (the JVP at the exact
    points of the FP32 trajectory)."""

    def __init__(self, cfg, state, material, nsteps: int, threads: int = 8):
        st = _st(state)
        self.n = st.size // 6
        self.cfg = cfg.copy()
        self.nsteps = int(nsteps)
        self.mat = np.ascontiguousarray(material, np.int32)
        self.nodes = (cfg.density_grid_size_x * cfg.density_grid_size_y *
                      cfg.density_grid_size_z)
        self._h = lib().or_rec_create(C.byref(self.cfg), self.n, _p(st), _p(self.mat),
                                      self.nsteps, int(threads))

    def close(self):
        if self._h:
            lib().or_rec_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def final_state(self) -> np.ndarray:
        out = np.zeros(6 * max(self.n, 1), np.float32)
        lib().or_rec_state(self._h, _p(out))
        return out[: 6 * self.n].reshape(self.n, 6)

    def contacts(self, step: int):
        cnt = np.zeros(max(self.n, 1), np.int32)
        total = lib().or_rec_contacts(self._h, int(step), _p(cnt), None, None)
        partners = np.zeros(max(total, 1), np.int32)
        flags = np.zeros(max(total, 1), np.int32)
        lib().or_rec_contacts(self._h, int(step), _p(cnt), _p(partners), _p(flags))
        return cnt[: self.n], partners[:total], flags[:total]

    def replay(self, precision: int, x0, v0, mu: float = 0.0, k: float = 0.0, tx0=None,
               tv0=None, tmu: float = 0.0, tk: float = 0.0, threads: int = 8):
        """Final state [n, 6], CIC grid [4 * nodes] and, for precision 65
        (FP64 primal) or 66 (the FP32 primal of the forward, bit for bit, with
        an FP64 tangent), their tangents."""
        x0 = np.ascontiguousarray(x0, np.float64).reshape(-1)
        v0 = np.ascontiguousarray(v0, np.float64).reshape(-1)
        assert x0.size == 3 * self.n and v0.size == 3 * self.n
        st = np.zeros(6 * max(self.n, 1), np.float64)
        rho = np.zeros(4 * self.nodes, np.float64)
        dst = np.zeros_like(st) if precision in (65, 66) else None
        drho = np.zeros_like(rho) if precision in (65, 66) else None
        tx = None if tx0 is None else np.ascontiguousarray(tx0, np.float64).reshape(-1)
        tv = None if tv0 is None else np.ascontiguousarray(tv0, np.float64).reshape(-1)
        lib().or_rec_replay(self._h, int(precision), _p(x0), _p(v0), float(mu), float(k),
                            None if tx is None else _p(tx), None if tv is None else _p(tv),
                            float(tmu), float(tk), _p(st), None if dst is None else _p(dst),
                            _p(rho), None if drho is None else _p(drho), int(threads))
        st = st[: 6 * self.n].reshape(self.n, 6)
        if precision in (65, 66):
            return st, rho, dst[: 6 * self.n].reshape(self.n, 6), drho
        return st, rho


# -- the reference's own YAML loader (built from the reference sources) ------

class RefConfigView(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("grid", C.c_int * 3),
        ("particles_per_cell", C.c_int),
        ("threads_per_block", C.c_int),
        ("density", C.c_int * 3),
        ("domain", C.c_float * 6),
        ("physics", C.c_float * 11),
        ("num_masses", C.c_int),
        ("masses", C.c_float * 16),
        ("init_type", C.c_char * 64),
        ("init_state_path", C.c_char * 512),
        ("init", C.c_float * 10),
        ("seeds", C.c_uint * 2),
        ("cube", C.c_float * 3),
        ("two_particle_separation", C.c_float),
        ("particle_material", C.c_int),
        ("time_integrator", C.c_char * 64),
        ("picard_iterations", C.c_int),
        ("kernel_timing", C.c_int),
        ("profiling", C.c_int * 2),
        ("dataset_sim_time", C.c_float),
        ("dataset_checkpoints", C.c_int),
        ("model_problem_sim_time", C.c_float),
    ]


def ref_loader_available() -> bool:
    return REF_LIB_PATH.exists()


def ref_load_config(path: str) -> RefConfigView:
    h = C.CDLL(str(REF_LIB_PATH))
    h.ref_load_config.argtypes = [C.c_char_p, C.POINTER(RefConfigView)]
    h.ref_load_config.restype = C.c_int
    out = RefConfigView()
    h.ref_load_config(str(path).encode(), C.byref(out))
    return out
