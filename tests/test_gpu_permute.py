"""GPU parity of the in-place permutations against the oracle.

Small and medium sizes are compared slot by slot with the oracle's restated
reference algorithm; every size class of the plan (in-LDS, row slides,
multi-level gathers with front partitions, bit-reversal tiles, chained moves)
is hit.
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


def _run(layout, b, kind, n, dev):
    A = torch.arange(n, dtype=TDT[kind], device=dev)
    tl_amd.permute(layout, A, b=b, kind=kind)
    torch.cuda.synchronize()
    got = A.cpu().numpy().astype(np.int64)
    exp = ol.layout_ranks(layout, n, b).astype(np.int64)
    return got, exp


CASES = [("bst", 1), ("btree", 16), ("btree", 8), ("btree", 2), ("btree", 3), ("btree", 32),
         ("btree", 1), ("mveb", 0), ("veb", 0)]


@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_permute_small(layout, b, kind, dev):
    for n in SMALL:
        got, exp = _run(layout, b, kind, n, dev)
        assert np.array_equal(got, exp), (layout, b, kind, n)


@pytest.mark.parametrize("layout,b", CASES)
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_permute_medium(layout, b, kind, dev):
    for n in MEDIUM:
        got, exp = _run(layout, b, kind, n, dev)
        assert np.array_equal(got, exp), (layout, b, kind, n)


@pytest.mark.parametrize("layout,b", [("btree", 16), ("btree", 8), ("btree", 2), ("bst", 1)])
@pytest.mark.parametrize("row_stash", ["0", "1000"])
def test_permute_stash_budgets(layout, b, row_stash, dev, monkeypatch):
    """Row slides off (hierarchical LDS + unit-merge gathers only) and a tiny
    row-slide budget (both kinds of gather in one plan)."""
    monkeypatch.setenv("TL_ROW_STASH_MAX", row_stash)
    for n in MEDIUM + [17 ** 5 - 1, 3_000_017]:
        got, exp = _run(layout, b, "u64", n, dev)
        assert np.array_equal(got, exp), (layout, b, row_stash, n)


@pytest.mark.parametrize("layout,b", [("btree", 16), ("btree", 8), ("bst", 1)])
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_rotation_walk_block(layout, b, kind, dev, monkeypatch):
    """Rotation lane walks in workgroups of 256 lanes (the default is 1024)
    give the same layouts"""
    monkeypatch.setenv("TL_WALK_BLOCK", "256")
    for n in MEDIUM + [3_000_017, 17 ** 6 - 1 + 12345]:
        got, exp = _run(layout, b, kind, n, dev)
        assert np.array_equal(got, exp), (layout, b, kind, n)


@pytest.mark.parametrize("tile_k", ["5", "6"])
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_bst_bitrev_tile_sizes(tile_k, kind, dev, monkeypatch):
    """32x32 (the default) and 64x64 bit-reversal tiles give the same BST
    layout, and the inverse restores sorted order"""
    monkeypatch.setenv("TL_BITREV_K", tile_k)
    for n in [4095, 16383, 16385, 65535, 300007, 1048575, 1048576, 3_000_017]:
        got, exp = _run("bst", 1, kind, n, dev)
        assert np.array_equal(got, exp), (tile_k, kind, n)
        A = torch.from_numpy(got.astype(np.int32 if kind == "u32" else np.int64)).to(dev)
        tl_amd.unpermute("bst", A, kind=kind)
        torch.cuda.synchronize()
        assert np.array_equal(A.cpu().numpy().astype(np.int64), np.arange(n)), (tile_k, kind, n)


@pytest.mark.parametrize("grid", ["8", "32"])
@pytest.mark.parametrize("tile_k", ["5", "6"])
@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_bst_bitrev_looped_grid(grid, tile_k, kind, dev, monkeypatch):
    """Bit-reversal tiles taken in a loop by a grid of fewer workgroups give
    the same BST layout (both phases), and the inverse restores sorted order"""
    monkeypatch.setenv("TL_BITREV_GRID", grid)
    monkeypatch.setenv("TL_BITREV_K", tile_k)
    for n in [4095, 65535, 300007, 1048575, 1048576, 3_000_017, (1 << 23) - 1, 1 << 23]:
        got, exp = _run("bst", 1, kind, n, dev)
        assert np.array_equal(got, exp), (grid, tile_k, kind, n)
        A = torch.from_numpy(got.astype(np.int32 if kind == "u32" else np.int64)).to(dev)
        tl_amd.unpermute("bst", A, kind=kind)
        torch.cuda.synchronize()
        assert np.array_equal(A.cpu().numpy().astype(np.int64), np.arange(n)), (grid, kind, n)


@pytest.mark.parametrize("fold", ["0", "1"])
@pytest.mark.parametrize("tile_k", ["5", "6"])
@pytest.mark.parametrize("kind", ["u32", "u64"])
@pytest.mark.parametrize("xcd", ["0", "1"])
def test_bst_leaf_fold(fold, tile_k, kind, xcd, dev, monkeypatch):
    """Trees of 2^h keys: the one leaf's move folded into phase one (row ends
    through the stash) and the unfolded plan (leaf gather, then the tiles) both
    give the oracle's layout, and the inverse restores sorted order; tiles
    numbered in workgroup order or XCD by XCD"""
    monkeypatch.setenv("TL_BST_FOLD", fold)
    monkeypatch.setenv("TL_BITREV_K", tile_k)
    monkeypatch.setenv("TL_BITREV_XCD", xcd)
    for n in [1 << 15, 1 << 16, 1 << 17, 1 << 20, 1 << 23, (1 << 23) - 1, 3_000_017]:
        got, exp = _run("bst", 1, kind, n, dev)
        assert np.array_equal(got, exp), (fold, tile_k, kind, n)
        A = torch.from_numpy(got.astype(np.int32 if kind == "u32" else np.int64)).to(dev)
        tl_amd.unpermute("bst", A, kind=kind)
        torch.cuda.synchronize()
        assert np.array_equal(A.cpu().numpy().astype(np.int64), np.arange(n)), (fold, kind, n)


@pytest.mark.parametrize("layout,b,n", [("bst", 1, (1 << 24) - 1), ("bst", 1, (1 << 24) + 12345),
                                        ("bst", 1, 1 << 24), ("bst", 1, 1 << 26),
                                        ("btree", 16, 17 ** 6 - 1), ("btree", 16, 30_000_001),
                                        ("btree", 8, 9 ** 7 + 5), ("mveb", 0, (1 << 24) - 1),
                                        ("mveb", 0, 30_000_001), ("veb", 0, (1 << 23) - 1),
                                        ("veb", 0, 30_000_001)])
def test_permute_large(layout, b, n, dev):
    got, exp = _run(layout, b, "u64", n, dev)
    assert np.array_equal(got, exp), (layout, b, n)


def test_permute_keys_not_ranks(dev):
    """Layouts move keys, not ranks: a sorted key set lands where its ranks do."""
    rng = np.random.default_rng(7)
    keys = np.cumsum(rng.integers(1, 64, size=123457, dtype=np.int64))
    for layout, b in (("bst", 1), ("btree", 16), ("mveb", 0)):
        A = torch.from_numpy(keys.copy()).to(dev)
        tl_amd.permute(layout, A, b=b, kind="u64")
        exp = keys[ol.layout_ranks(layout, keys.size, b).astype(np.int64)]
        assert np.array_equal(A.cpu().numpy(), exp)


def test_permute_invalid_args(dev):
    A = torch.arange(10, dtype=torch.int64, device=dev)
    with pytest.raises(tl_amd.TLError):
        tl_amd.permute_btree(A, 0, kind="u64")
