"""ctypes binding of libworld_amd.so (the C ABI declared in include/world_amd.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libworld_amd.so"

WD_MAX_MATERIALS = 16
WD_PATH_MAX = 512
WD_HISTORY_SLOTS = 24
WD_NOALLOC_MAX_GRAINS = 1 << 20
WD_BINNING_TUNED, WD_BINNING_SCATTER, WD_BINNING_BUCKETS = 0, 1, 2

WD_INIT_CUBE, WD_INIT_TWO_PARTICLES, WD_INIT_SINGLE_PARTICLE, WD_INIT_FILE = range(4)
WD_SEMI_IMPLICIT_EULER, WD_BACKWARD_EULER_PICARD = 0, 1
WD_TANGENTIAL_VISCOUS, WD_TANGENTIAL_SPRING_HISTORY = 0, 1
WD_PERTURB_UNIFORM, WD_PERTURB_GAUSSIAN = 0, 1
WD_OK, WD_ERR_ARG, WD_ERR_HIP, WD_ERR_IO, WD_ERR_STATE, WD_ERR_DEVICE = 0, -1, -2, -3, -4, -5


class WdConfig(C.Structure):
    """Field-for-field mirror of ``wd_config``."""

    _fields_ = [
        ("dt", C.c_float),
        ("collision_grid_size_x", C.c_int),
        ("collision_grid_size_y", C.c_int),
        ("collision_grid_size_z", C.c_int),
        ("particles_per_cell", C.c_int),
        ("threads_per_block", C.c_int),
        ("density_grid_size_x", C.c_int),
        ("density_grid_size_y", C.c_int),
        ("density_grid_size_z", C.c_int),
        ("x_min", C.c_float), ("x_max", C.c_float),
        ("y_min", C.c_float), ("y_max", C.c_float),
        ("z_min", C.c_float), ("z_max", C.c_float),
        ("gravity", C.c_float), ("particle_radius", C.c_float), ("max_force", C.c_float),
        ("floor_z", C.c_float), ("ceiling_z", C.c_float),
        ("wall_x_min", C.c_float), ("wall_x_max", C.c_float),
        ("wall_y_min", C.c_float), ("wall_y_max", C.c_float),
        ("restitution", C.c_float), ("friction", C.c_float),
        ("tangential_model", C.c_int),
        ("tangential_stiffness_ratio", C.c_float),
        ("normal_stiffness", C.c_float),
        ("num_masses", C.c_int),
        ("masses", C.c_float * WD_MAX_MATERIALS),
        ("init_type", C.c_int),
        ("init_state_path", C.c_char * WD_PATH_MAX),
        ("init_x0", C.c_float), ("init_y0", C.c_float), ("init_z0", C.c_float),
        ("init_vx0", C.c_float), ("init_vy0", C.c_float), ("init_vz0", C.c_float),
        ("init_jitter", C.c_float),
        ("init_seed", C.c_uint),
        ("init_perturbation", C.c_float),
        ("init_perturbation_seed", C.c_uint),
        ("cube_length_x", C.c_float), ("cube_length_y", C.c_float), ("cube_length_z", C.c_float),
        ("two_particle_separation", C.c_float),
        ("particle_material", C.c_int),
        ("lattice_spacing", C.c_float),
        ("perturbation_distribution", C.c_int),
        ("time_integrator", C.c_int),
        ("picard_iterations", C.c_int),
        ("kernel_timing", C.c_int),
    ]

    def set_masses(self, masses) -> None:
        masses = list(masses)
        if not 1 <= len(masses) <= WD_MAX_MATERIALS:
            raise ValueError("masses must hold 1..16 entries")
        self.num_masses = len(masses)
        for i, m in enumerate(masses):
            self.masses[i] = m

    def get_masses(self) -> list[float]:
        return [self.masses[i] for i in range(self.num_masses)]

    def copy(self) -> "WdConfig":
        out = WdConfig()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(WdConfig))
        return out


class WdError(RuntimeError):
    pass


# (name, restype, argtypes)
_SIGNATURES = [
    ("wd_last_error", C.c_char_p, []),
    ("wd_version", C.c_char_p, []),
    ("wd_config_size", C.c_size_t, []),
    ("wd_config_default", C.c_int, [C.POINTER(WdConfig)]),
    ("wd_config_load", C.c_int, [C.c_char_p, C.POINTER(WdConfig)]),
    ("wd_config_grain_count", C.c_int, [C.POINTER(WdConfig), C.POINTER(C.c_int)]),
    ("wd_initial_state", C.c_int, [C.POINTER(WdConfig), C.c_void_p, C.c_void_p, C.c_int]),
    ("wd_create", C.c_int, [C.POINTER(WdConfig), C.POINTER(C.c_void_p)]),
    ("wd_create_from_state", C.c_int, [C.POINTER(WdConfig), C.c_int, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_void_p)]),
    ("wd_destroy", C.c_int, [C.c_void_p]),
    ("wd_num_grains", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_get_state", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_set_state", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_get_material", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_set_material", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_set_dt", C.c_int, [C.c_void_p, C.c_float]),
    ("wd_get_time", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("wd_set_time", C.c_int, [C.c_void_p, C.c_float]),
    ("wd_step", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_sync", C.c_int, [C.c_void_p]),
    ("wd_density_grid", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_total_energy", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("wd_energy_log_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_energy_log_read", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("wd_density_log_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_density_log_record", C.c_int, [C.c_void_p]),
    ("wd_density_log_read", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("wd_profile_step", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                  C.c_char_p, C.c_size_t]),
    ("wd_max_accel", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("wd_is_stable", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_overflow_count", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_stage_ms", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_cell_order", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("wd_cell_order_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("wd_set_scan_tile", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_set_cell_scale", C.c_int, [C.c_void_p, C.c_float]),
    ("wd_get_cell_scale", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_void_p]),
    ("wd_tune_cell_scale", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float)]),
    ("wd_get_scan_tile", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_set_binning", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_get_binning", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_set_fused_steps", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_set_graph_replay", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_get_fused_steps", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_fused_misses", C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong)]),
    ("wd_neighbor_rows", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("wd_rhs", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_get_history", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("wd_batch_create", C.c_int, [C.POINTER(WdConfig), C.c_int, C.POINTER(C.c_void_p)]),
    ("wd_batch_members", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_batch_member_range", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]),
    ("wd_batch_density_grid", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("wd_batch_overflow_counts", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_batch_is_stable", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_set_friction", C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    ("wd_set_normal_stiffness", C.c_int, [C.c_void_p, C.c_float]),
    ("wd_adjoint_record", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("wd_set_force_variant", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("wd_get_force_variant", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("wd_set_force_block", C.c_int, [C.c_void_p, C.c_int]),
    ("wd_get_force_block", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("wd_contact_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    ("wd_adjoint_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    ("wd_set_state_device", C.c_int, [C.c_void_p, C.c_void_p]),
    ("wd_batch_density_grid_device", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("wd_adjoint_backward_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]),
]

EXPORTED = [name for name, _, _ in _SIGNATURES]

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m world_amd.build` "
                "(the engine has no fallback implementation)")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_LOCAL)
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.wd_config_size() != C.sizeof(WdConfig):
            raise ImportError("wd_config layout mismatch between Python binding and library")
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().wd_last_error().decode(errors="replace")
        raise WdError(f"world_amd error {rc}: {msg}")


def ptr(a: np.ndarray) -> C.c_void_p:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def default_config() -> WdConfig:
    cfg = WdConfig()
    check(lib().wd_config_default(C.byref(cfg)))
    return cfg


def load_config(path: str | os.PathLike) -> WdConfig:
    cfg = WdConfig()
    check(lib().wd_config_load(str(path).encode(), C.byref(cfg)))
    return cfg


def grain_count(cfg: WdConfig) -> int:
    n = C.c_int(0)
    check(lib().wd_config_grain_count(C.byref(cfg), C.byref(n)))
    return n.value


def initial_state(cfg: WdConfig) -> tuple[np.ndarray, np.ndarray]:
    n = C.c_int(0)
    check(lib().wd_config_grain_count(C.byref(cfg), C.byref(n)))
    st = np.zeros(6 * max(n.value, 1), np.float32)
    mat = np.zeros(max(n.value, 1), np.int32)
    check(lib().wd_initial_state(C.byref(cfg), ptr(st), ptr(mat), n.value))
    return st[: 6 * n.value].copy(), mat[: n.value].copy()
