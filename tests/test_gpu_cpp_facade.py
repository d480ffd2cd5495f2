"""The reference's GoogleTest cases, compiled against the drop-in C++ class
(include/world/particle_dynamics.h) and run as a program."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).parent / "cpp" / "particle_dynamics_test"


def test_cpp_facade_reference_suite():
    assert BIN.exists(), "build the C++ façade test with `python -m world_amd.build`"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 12
