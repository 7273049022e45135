"""The reference-named C++ surface (include/tree_layouts.hpp) drives all three
layouts and their queries from a reference-style driver (ITERS runs,
printTimings lines).  The driver checks the BST and B-tree layouts slot for
slot and every query answer itself; the mvEB layouts it reports by digest,
checked here against the oracle.  The involution programs' names
(namespace involutions) are checked the same way."""
import os
import re
import subprocess

import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hpp_driver")


def test_cpp_header_driver():
    assert os.path.exists(BIN), "tests/cpp/hpp_driver is not built (run __graft_entry__.build())"
    print("[test] hpp_driver 1000003", flush=True)
    out = subprocess.run([BIN, "1000003"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")
    digests = re.findall(r"^digest (\w+) (\d+) ([0-9a-f]{16})$", out.stdout, re.M)
    assert len(digests) == 3, out.stdout
    for _, n, hexd in digests:
        assert int(hexd, 16) == ol.digest(ol.layout_ranks("mveb", int(n))), (n, hexd)
    inv = re.findall(r"^digest-inv (\w+) (\d+) ([0-9a-f]{16})$", out.stdout, re.M)
    assert len(inv) == 3, out.stdout
    for _, n, hexd in inv:
        assert int(hexd, 16) == ol.digest(ol.layout_ranks("mveb", int(n))), (n, hexd)
    assert len(re.findall(r"^involutions layout \d \w+ n=\d+: [0-9.]+ ms, 0 slots differ$",
                          out.stdout, re.M)) == 6, out.stdout
    # printTimings / printQueryTimings: n mean sd tp_mean tp_sd then ITERS times
    lines = re.findall(r"^layout \d \w+ permute: (.*)$", out.stdout, re.M)
    assert lines and all(len(line.split()) == 5 + 10 for line in lines), out.stdout
    qlines = re.findall(r"^layout \d \w+ query: (.*)$", out.stdout, re.M)
    assert qlines and all(len(line.split()) == 6 + 10 for line in qlines), out.stdout
