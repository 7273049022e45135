"""The reference-named C++ surface (include/tree_layouts.hpp) drives all three
layouts and their queries from a reference-style driver."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hpp_driver")


def test_cpp_header_driver():
    assert os.path.exists(BIN), "tests/cpp/hpp_driver is not built (run __graft_entry__.build())"
    out = subprocess.run([BIN, "1000003"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")
