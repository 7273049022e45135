"""The BST of 2^D keys kept one place right, on the CPU: the bit-reversal
tiles of tl_kernels.hpp in every form (plain; phase one absorbing row ends;
the levels shifted in place and unfolded; one workgroup per tile number, a
looping grid, tile numbers XCD by XCD) and the level heads copy, compiled for
the host over the HIP shim in tests/emu/, every lane a thread, on exactly
sized buffers under AddressSanitizer.  The forward sequence must equal the
plain phases of the keys after the leaf with the leaf last, and the inverse
sequence must restore the input (tests/emu/emu_bitrev.cpp)."""
import os
import subprocess

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


def test_folded_bitrev_tiles_emulated():
    subprocess.run(["make", "-C", EMU, "-j4", "_build/emu_bitrev"], check=True,
                   capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([os.path.join(EMU, "_build", "emu_bitrev")], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "[emu-bitrev] 0 wrong in all" in r.stderr, r.stderr[-4000:]
    assert r.stderr.count("fold 0 wrong, unfold 0 wrong") == 30, r.stderr[-4000:]
