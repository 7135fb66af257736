"""The C++17 facade (include/bvh/collision.hpp) driven like the reference's
collision_object tests: a narrowphase functor over patch pairs, the built-in
narrowphase across three objects, local_patches and od validation."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "facade_test")


@pytest.mark.gpu
def test_cpp_facade():
    assert os.path.exists(EXE), "tests/cpp/facade_test not built: run __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "facade_test: ok" in r.stdout
