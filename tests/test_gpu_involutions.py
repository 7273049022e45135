"""GPU parity of the involution family (tl_permute_*_involutions) and its
primitives.

The three permutations must give exactly the oracle's layouts (the reference's
involution programs and its gather / cycle-leader programs lay out the same
trees).  The sizes cover perfect and non-perfect trees of every layout, the
B-tree sizes n in [(b+1)^k, (b+1)^k + b - 1] whose leaf shuffle is empty (the
reference's loop over [1, n - 1) does not end there), the mvEB's in-LDS base
cases and its global rounds.  The primitives are checked against their closed
forms: the k-way shuffle of n = d*k keys is the transpose of a k x d matrix,
its inverse the transpose of a d x k one, the k-way unshuffle of k^d - 1 keys a
stable partition by (i+1) mod k, the rotation numpy's roll.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

TDT = {"u32": torch.int32, "u64": torch.int64}

SMALL = [0, 1, 2, 3, 4, 7, 8, 15, 16, 17, 18, 31, 32, 33, 63, 100, 255, 256, 288, 289, 290,
         304, 305, 1000, 1023, 4095, 4096, 4097, 4912, 4913, 4914, 8191, 20000]
MEDIUM = [65535, 65536, 83520, 83521, 83536, 100003, 262143, 300007, 1048575, 1048576, 2000003]

CASES = [("bst", 1), ("btree", 16), ("btree", 8), ("btree", 3), ("btree", 1), ("mveb", 0)]


def _run(layout, b, kind, n, dev):
    A = torch.arange(n, dtype=TDT[kind], device=dev)
    tl_amd.permute_involutions(layout, A, b=b, kind=kind)
    torch.cuda.synchronize()
    got = A.cpu().numpy().astype(np.int64)
    exp = ol.layout_ranks(layout, n, b).astype(np.int64)
    return got, exp


@pytest.mark.gpu
@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_involutions_small(layout, b, kind, dev):
    for n in SMALL:
        print(f"[test] involutions {layout} b={b} {kind} n={n}", flush=True)
        got, exp = _run(layout, b, kind, n, dev)
        assert np.array_equal(got, exp), (layout, b, kind, n)


@pytest.mark.gpu
@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_involutions_medium(layout, b, kind, dev):
    for n in MEDIUM:
        print(f"[test] involutions {layout} b={b} {kind} n={n}", flush=True)
        got, exp = _run(layout, b, kind, n, dev)
        assert np.array_equal(got, exp), (layout, b, kind, n)


@pytest.mark.gpu
@pytest.mark.parametrize("layout,b,n", [("bst", 1, (1 << 24) + 12345), ("btree", 16, 17 ** 6 + 99),
                                        ("btree", 16, (1 << 24) + 12345), ("mveb", 0, (1 << 24) - 1),
                                        ("mveb", 0, (1 << 24) + 12345)])
def test_involutions_large(layout, b, n, dev):
    print(f"[test] involutions {layout} b={b} u64 n={n}", flush=True)
    got, exp = _run(layout, b, "u64", n, dev)
    assert np.array_equal(got, exp), (layout, b, n)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_shuffle_primitives(kind, dev):
    for n, k in [(12, 3), (12, 4), (1000, 8), (1 << 16, 16), (17 * 4097, 17), (3 * 1000003, 3),
                 (1 << 22, 1024), (7, 7), (5, 1)]:
        d = n // k
        print(f"[test] shuffle_dk {kind} n={n} k={k}", flush=True)
        a = np.arange(n, dtype=np.int64)
        A = torch.arange(n, dtype=TDT[kind], device=dev)
        tl_amd.shuffle_dk(A, k, kind=kind)
        assert np.array_equal(A.cpu().numpy().astype(np.int64), a.reshape(k, d).T.ravel()), (n, k)
        A = torch.arange(n, dtype=TDT[kind], device=dev)
        tl_amd.unshuffle_dk(A, k, kind=kind)
        assert np.array_equal(A.cpu().numpy().astype(np.int64), a.reshape(d, k).T.ravel()), (n, k)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_unshuffle_and_rotation(kind, dev):
    for k, dd in [(2, 1), (2, 10), (3, 7), (17, 3), (17, 5), (256, 2), (4, 11)]:
        n = k ** dd - 1
        print(f"[test] unshuffle {kind} n={n} k={k}", flush=True)
        i = np.arange(n, dtype=np.int64)
        exp = np.concatenate([i[(i + 1) % k == r] for r in range(k)])
        A = torch.arange(n, dtype=TDT[kind], device=dev)
        tl_amd.unshuffle(A, k, kind=kind)
        assert np.array_equal(A.cpu().numpy().astype(np.int64), exp), (n, k)
    for n in [1, 2, 1000, (1 << 22) + 3]:
        for k in sorted({0, 1, n // 3, n // 2, n - 1, n}):
            print(f"[test] shift_right {kind} n={n} k={k}", flush=True)
            A = torch.arange(n, dtype=TDT[kind], device=dev)
            tl_amd.shift_right(A, k, kind=kind)
            assert np.array_equal(A.cpu().numpy().astype(np.int64),
                                  np.roll(np.arange(n, dtype=np.int64), k)), (n, k)


def test_primitive_arguments_rejected_on_the_host():
    """Sizes the primitives do not define fail before any device work (the
    pointer is never dereferenced, so this runs without a GPU)."""
    L = tl_amd.lib()
    u64 = tl_amd.DTYPE["u64"]
    assert L.tl_shuffle_dk(1 << 20, 10, 3, u64, None, None) == tl_amd.TL_EINVAL
    assert L.tl_unshuffle_dk(1 << 20, 10, 0, u64, None, None) == tl_amd.TL_EINVAL
    assert L.tl_unshuffle(1 << 20, 10, 3, u64, None, None) == tl_amd.TL_EINVAL
    assert L.tl_unshuffle(1 << 20, 26, 1, u64, None, None) == tl_amd.TL_EINVAL
    assert L.tl_shift_right(1 << 20, 10, 11, u64, None, None) == tl_amd.TL_EINVAL
    assert L.tl_permute_btree_involutions(1 << 20, 10, 0, u64, None, None) == tl_amd.TL_EINVAL
    assert "multiple of k" in L.tl_last_error().decode() or "b must be" in \
        L.tl_last_error().decode()
