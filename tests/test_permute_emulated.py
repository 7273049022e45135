"""The single-array permutations on the CPU: the product's tl_permute.hip
compiled for the host over the HIP shim in tests/emu/ (emu_permute.cpp), every
plan launch run one workgroup at a time with every lane a thread, on exactly
sized buffers under AddressSanitizer.  Chained moves run in ticket order, so
every wait is met when reached.  Each case permutes 0..n-1 through the
library's entry point, compares every slot with the CPU oracle, runs the
inverse entry point and checks the identity.  The mvEB cases include the
fused sparse slides (a leaf block's own gather riding in the top level's
chained moves) of perfect and non-perfect trees, u32 and u64; the classic
vEB's fused slides differ only in the slot function, and their plans are
replayed in test_plan.py (its host oracle is too slow for this file).  The
BST cases cover the tree kept one place right (2^16 keys, and 300,007 keys:
gather, rotation, then the same phases).  The involution family
(tl_involution.hip) runs the same way: its three permutations against the
oracle's layouts (B-trees of b = 1, 2, 3, 8, 16; the mvEB's global rounds and
in-LDS base cases) and its shuffle, unshuffle and rotation primitives against
their closed forms."""
import os
import subprocess

import pytest

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


@pytest.fixture(scope="module")
def emulator():
    subprocess.run(["make", "-C", EMU, "-j4", "_build/emu_permute"], check=True,
                   capture_output=True)
    return os.path.join(EMU, "_build", "emu_permute")


@pytest.mark.parametrize("layout,width,sizes", [
    ("mveb", 8, [300007, 1048575]),
    ("mveb", 4, [300007]),
    ("bst", 8, [65536, 300007]),
    ("btree", 4, [300007]),
    ("bst_inv", 8, [4097, 65537]),
    ("btree_inv", 4, [4914, 30000]),
    ("mveb_inv", 4, [70000]),
    ("mveb_inv", 8, [4095]),
    ("prims", 8, [5040]),
])
def test_permute_emulated(emulator, layout, width, sizes):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([emulator, layout, str(width)] + [str(n) for n in sizes],
                       capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    if layout == "prims":
        assert r.stderr.count(": 0 wrong") == len(sizes), r.stderr[-4000:]
    elif layout.endswith("_inv"):
        # the B-tree cases run b = 1, 2, 3, 8, 16 per size
        per = 5 if layout == "btree_inv" else 1
        assert r.stderr.count(", 0 wrong") == per * len(sizes), r.stderr[-4000:]
    else:
        assert r.stderr.count("0 wrong, inverse 0 wrong") == len(sizes), r.stderr[-4000:]


def test_modinv_partner_exact_to_2p40():
    """The shuffle passes' modular-inverse partner (tl_modinv.hpp, built for
    the host) against Python's exact arithmetic, x -> g * (r * (x / g)^-1 mod
    (m / g)) with g = gcd(x, m) and the inverse from pow(., -1, .), at moduli
    below 2^32 and up to 2^40 (arrays of 2^32 keys and more, which no test
    permutes), with common factors of x and m and multipliers from 1 to m - 1,
    and every x of the moduli below 200."""
    import math

    import numpy as np
    subprocess.run(["make", "-C", EMU, "_build/emu_modinv"], check=True, capture_output=True)
    rng = np.random.default_rng(11)
    cases = []
    fixed = [2**32 - 1, 2**32, 2**32 + 1, 3 * 2**31, 2**40 - 1, 2**40 - 2, 17**9 - 1]
    for i in range(4000):
        m = fixed[i % len(fixed)] if i < 70 else int(
            rng.integers(2**32, 2**40) if i % 2 else rng.integers(3, 2**32))
        x = int(rng.integers(1, m))
        if i % 3 == 0:   # a common factor with m
            f = math.gcd(m, int(rng.integers(2, 2**12)))
            x = max(1, (x // f) * f) % m or f
        r = [17, math.isqrt(m), m - 1, int(rng.integers(1, m))][i % 4]
        cases.append((x, m, r))
    for m in range(2, 200):   # every x of small moduli, m / g down to 2
        for x in range(1, m):
            cases.append((x, m, [1, 17, m - 1][x % 3]))
    inp = "".join(f"{x} {m} {r}\n" for x, m, r in cases)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    out = subprocess.run([os.path.join(EMU, "_build", "emu_modinv")], input=inp, env=env,
                         capture_output=True, text=True, check=True, timeout=120).stdout.split()
    assert len(out) == len(cases)
    for (x, m, r), got in zip(cases, out):
        g = math.gcd(x, m)
        mg = m // g
        want = g * ((r * pow(x // g, -1, mg)) % mg)
        assert int(got) == want, (x, m, r)
