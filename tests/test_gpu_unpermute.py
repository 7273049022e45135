"""GPU parity of the inverse permutations (layout -> sorted order).

The oracle's layout of the ranks 0..n-1 goes back to 0..n-1 exactly, and
arbitrary data survives the forward and then the inverse permutation, at the
sizes of every plan shape the forward tests cover; the large sizes start from
the forward permutation on the GPU, whose parity the forward tests establish.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu

TDT = {"u32": torch.int32, "u64": torch.int64}
SMALL = [0, 1, 2, 3, 4, 7, 8, 15, 16, 17, 31, 63, 100, 255, 256, 1000, 1023, 4095, 4097,
         8191, 8192, 8193, 16383, 16385, 20000]
MEDIUM = [65535, 65536, 83520, 83521, 100003, 262143, 300007, 1048575, 1048576, 1419856,
          1419857, 2000003]
CASES = [("bst", 1), ("btree", 16), ("btree", 8), ("btree", 2), ("btree", 3), ("btree", 32),
         ("btree", 1), ("mveb", 0), ("veb", 0)]


def _check(layout, b, kind, n, dev, rng):
    ranks = ol.layout_ranks(layout, n, b).astype(np.int64)
    A = torch.from_numpy(ranks).to(TDT[kind]).to(dev)
    tl_amd.unpermute(layout, A, b=b, kind=kind)
    torch.cuda.synchronize()
    assert np.array_equal(A.cpu().numpy().astype(np.int64), np.arange(n)), (layout, b, kind, n)
    hi = 1 << 30 if kind == "u32" else 1 << 60
    X = torch.from_numpy(rng.integers(-hi, hi, size=n)).to(TDT[kind]).to(dev)
    Y = X.clone()
    tl_amd.permute(layout, Y, b=b, kind=kind)
    tl_amd.unpermute(layout, Y, b=b, kind=kind)
    torch.cuda.synchronize()
    assert torch.equal(X, Y), ("round trip", layout, b, kind, n)


@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_unpermute_small(layout, b, kind, dev):
    rng = np.random.default_rng(100 + b)
    for n in SMALL:
        _check(layout, b, kind, n, dev, rng)


@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_unpermute_medium(layout, b, kind, dev):
    rng = np.random.default_rng(200 + b)
    for n in MEDIUM:
        _check(layout, b, kind, n, dev, rng)


@pytest.mark.parametrize("layout,b", [("btree", 16), ("btree", 8), ("btree", 2), ("bst", 1)])
@pytest.mark.parametrize("row_stash", ["0", "1000"])
def test_unpermute_stash_budgets(layout, b, row_stash, dev, monkeypatch):
    """Row spreads off (LDS + unit-merge inverses only) and a tiny budget."""
    monkeypatch.setenv("TL_ROW_STASH_MAX", row_stash)
    rng = np.random.default_rng(300 + b)
    for n in MEDIUM + [17 ** 5 - 1, 3_000_017]:
        _check(layout, b, "u64", n, dev, rng)


@pytest.mark.parametrize("layout,b,n", [("bst", 1, (1 << 24) - 1), ("bst", 1, (1 << 24) + 12345),
                                        ("btree", 16, 17 ** 6 - 1), ("btree", 16, 30_000_001),
                                        ("btree", 8, 9 ** 7 + 5), ("mveb", 0, (1 << 24) - 1),
                                        ("mveb", 0, 30_000_001), ("veb", 0, (1 << 23) - 1),
                                        ("veb", 0, 30_000_001)])
def test_unpermute_large(layout, b, n, dev):
    A = torch.arange(n, dtype=torch.int64, device=dev)
    tl_amd.permute(layout, A, b=b, kind="u64")
    tl_amd.unpermute(layout, A, b=b, kind="u64")
    torch.cuda.synchronize()
    assert torch.equal(A, torch.arange(n, dtype=torch.int64, device=dev)), (layout, b, n)
