"""The uneven batch's single persistent launch (k_batch and its sub-task
executors in tl_batch.hip), run on the CPU: the product source is compiled
for the host over a shim of the HIP surface (tests/emu/), its kernels run one
workgroup at a time with every lane a thread, on host memory, under
AddressSanitizer, and every array of a mixed batch (BST / B-tree b=8 / mvEB,
int32 and uint64 keys, hits and misses) must come out with the oracle's
layout and answers.  The emulator also aborts when the lanes of a workgroup
pass different numbers of barriers, or the lanes of a wave different numbers
of ballots.

The 60-array case mixes key widths among arrays that use the stash, which
is how this emulation found that the executors addressed every array's stash
region in units of its own key width (4-byte arrays overlapped their
neighbours' regions); regions are now 8-byte words.  The 240-array batch of
the GPU tests (about 4 minutes per LDS budget) runs with TL_EMU_FULL=1."""
import os
import subprocess

import pytest

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


@pytest.fixture(scope="module")
def emulator():
    subprocess.run(["make", "-C", EMU, "-j4"], check=True, capture_output=True)
    return os.path.join(EMU, "_build", "emu")


def _run(emulator, count, log2max, lds, qmax, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([emulator, str(count), str(log2max), str(lds), str(qmax), str(seed)],
                       capture_output=True, text=True, timeout=1500, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert f"0 of {count} arrays differ from the oracle" in r.stderr, r.stderr[-4000:]


def test_persistent_executors_mixed_widths(emulator):
    _run(emulator, 60, 16, 0, 256, 7)


@pytest.mark.skipif(os.environ.get("TL_EMU_FULL") != "1",
                    reason="the 240-array batch takes about 4 minutes per LDS budget; "
                           "set TL_EMU_FULL=1")
@pytest.mark.parametrize("lds", [0, 32768])
def test_persistent_executors_240_array_batch(emulator, lds):
    _run(emulator, 240, 18, lds, 4096, 5)


@pytest.fixture(scope="module")
def planned():
    subprocess.run(["make", "-C", EMU, "-j4", "_build/emu_planned"], check=True,
                   capture_output=True)
    return os.path.join(EMU, "_build", "emu_planned")


def _run_planned(planned, count, log2max, seed, one=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    if one:
        env["EMU_ONE"] = one
    r = subprocess.run([planned, str(count), str(log2max), "0", "0", str(seed)],
                       capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert f"0 of {count} arrays differ from the oracle" in r.stderr, r.stderr[-4000:]


def test_planned_path_mixed_batch(planned):
    """The planned path (tl_batch_run's multi-array launches and batched
    in-LDS launches, built with the emulated tl_permute.hip) on a mixed batch
    of 40 arrays of up to 2^19 keys, permutations only: every array equals
    the oracle's layout"""
    _run_planned(planned, 40, 19, 7)


@pytest.mark.parametrize("one", ["2 u64 300007", "2 i32 1048576", "2 u64 1048575",
                                 "0 u64 300007", "0 i32 1048577", "0 u64 65536"])
def test_planned_path_fused_forms(planned, one):
    """mvEB arrays whose top-level gather is a fused sparse slide (the leaf
    blocks' own gathers in the multi-array chained moves, their roots laid
    out by the multi-array part-roots launch), and BST arrays kept one place
    right (the level heads' copies, the top read one place right, the
    multi-array shifted level reversal)"""
    _run_planned(planned, 1, 19, 7, one)

