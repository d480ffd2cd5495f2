"""The C-ABI library without a GPU: it loads, exports every entry point that
include/world_amd.h declares, agrees with the Python struct layout, and its
host-only entry points (configuration, initial state) match the oracle."""
import re
from pathlib import Path

import numpy as np

from world_amd import _capi, workloads

HEADER = Path(__file__).resolve().parents[1] / "include" / "world_amd.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_capi.EXPORTED) == names


def test_struct_layout_matches_library():
    import ctypes as C

    assert _capi.lib().wd_config_size() == C.sizeof(_capi.WdConfig)
    assert _capi.lib().wd_version().decode().startswith("world_amd")


def test_host_initial_state_matches_oracle(oracle_lib):
    for c in (_capi.default_config(), workloads.column_collapse_4k(), workloads.column_8k()):
        st, mat = _capi.initial_state(c)
        st2, mat2 = oracle_lib.initial_state(c)
        np.testing.assert_array_equal(st, st2)
        np.testing.assert_array_equal(mat, mat2)


def test_gaussian_perturbation_statistics():
    c = workloads.perturbed(workloads.column_collapse_4k(), 50e-6, 1001)
    base, _ = _capi.initial_state(workloads.column_collapse_4k())
    st, _ = _capi.initial_state(c)
    d = (st - base).reshape(-1, 6)[:, :3].astype(np.float64)
    assert abs(d.mean()) < 5e-6
    assert abs(d.std() - 50e-6) < 3e-6


def test_file_initializer_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    st = rng.normal(size=(7, 6)).astype(np.float32)
    p = tmp_path / "grains.bin"
    with open(p, "wb") as f:
        np.array([7], np.int32).tofile(f)
        st.tofile(f)
    c = _capi.default_config()
    c.init_type = _capi.WD_INIT_FILE
    c.init_state_path = str(p).encode()
    got, mat = _capi.initial_state(c)
    np.testing.assert_array_equal(got.reshape(7, 6), st)
    assert np.all(mat == 1)
    # short file is an error, not a crash
    with open(p, "wb") as f:
        np.array([9], np.int32).tofile(f)
        st.tofile(f)
    try:
        _capi.initial_state(c)
    except _capi.WdError as e:
        assert "short" in str(e)
    else:
        raise AssertionError("short state file accepted")


def test_batch_rejects_incompatible_members_before_device_work():
    import ctypes as C

    a = workloads.column(3, 3, 3, friction=0.1)
    b = workloads.column(4, 4, 4, friction=0.5)
    b.gravity = 3.7  # members may differ only in initialization, friction, restitution, masses
    arr = (_capi.WdConfig * 2)(a, b)
    h = C.c_void_p()
    rc = _capi.lib().wd_batch_create(arr, 2, C.byref(h))
    assert rc == _capi.WD_ERR_ARG and not h.value
    assert b"ensemble members differ" in _capi.lib().wd_last_error()
    rc = _capi.lib().wd_batch_create(arr, 0, C.byref(h))
    assert rc == _capi.WD_ERR_ARG


def test_uneven_ensemble_workload():
    cs = workloads.uneven_ensemble()
    assert len(cs) == 64
    n = []
    for m, c in enumerate(cs):
        n.append(_capi.grain_count(c))
        assert c.init_perturbation_seed == 1000 + m
        assert 0.0 <= c.friction <= 0.6
    assert 2048 * 0.8 < min(n) and max(n) <= 32768 * 1.1 and len(set(n)) > 32
