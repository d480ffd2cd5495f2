"""MI355X-native granular dynamics engine (gfx950 HIP kernels behind a C ABI).

The compute path is ``libworld_amd.so``; this package holds its ctypes binding
(``_capi``), a Python mirror of the reference's ``ParticleDynamics`` API
(``sim``), workload builders (``workloads``) and the build script (``build``).
"""
from ._capi import WdConfig, WdError, default_config, initial_state, load_config  # noqa: F401

__all__ = ["WdConfig", "WdError", "default_config", "initial_state", "load_config"]
