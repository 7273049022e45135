"""Host-side checks of the gfx950 library that need no GPU: the C ABI loads
and exports every declared entry point, and every permutation plan (dry run,
bounds-checked by the library) replays to the oracle's layout."""
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
import plan_emulator as pe
import tl_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = tl_amd.lib()
    hdr = open(os.path.join(ROOT, "include", "tree_layouts.h")).read()
    names = set(re.findall(r"\b(tl_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 12
    for name in sorted(names):
        assert hasattr(L, name), name
    assert b"gfx950" in L.tl_version()


def test_invalid_arguments_fail_without_device():
    with pytest.raises(tl_amd.TLError):
        tl_amd.plan_export("btree", 100, 0)
    with pytest.raises(tl_amd.TLError):
        tl_amd.plan_export("btree", 100, 5000)


@pytest.mark.parametrize("layout,bs", [("btree", (16, 8, 2, 3, 32, 1)), ("bst", (1,)),
                                       ("mveb", (0,))])
@pytest.mark.parametrize("lds", [256, 2048])
@pytest.mark.parametrize("row_stash", [None, "0", "40"])
def test_plans_replay_to_oracle(layout, bs, lds, row_stash, monkeypatch):
    """row_stash: the library's default stash budget for row slides, none (the
    hierarchical LDS + unit-merge gathers), or 40 separators (both kinds in
    one plan)"""
    if row_stash is not None:
        monkeypatch.setenv("TL_ROW_STASH_MAX", row_stash)
    sizes = [1, 2, 3, 7, 16, 17, 100, 255, 256, 1000, 1023, 4095, 5000, 17 ** 3 - 1, 17 ** 3,
             33333, 65535, 65536]
    for b in bs:
        for n in sizes:
            buf = tl_amd.plan_export(layout, n, b, "u64", lds)
            A = np.arange(n, dtype=np.int64)
            pe.replay(buf, A)
            assert np.array_equal(A, ol.layout_ranks(layout, n, b).astype(np.int64)), \
                (layout, b, lds, n)


def test_full_size_plans_are_in_bounds():
    for layout, b, n in [("btree", 16, 1 << 30), ("btree", 16, 10 ** 8), ("bst", 1, 1 << 30),
                         ("bst", 1, 3 << 28), ("mveb", 0, (1 << 29) - 1),
                         ("mveb", 0, 3 * 10 ** 8), ("mveb", 0, 1 << 30), ("btree", 8, 1 << 22)]:
        for kind in ("u32", "u64"):
            buf = tl_amd.plan_export(layout, n, b, kind)
            assert int(buf[0]) > 0
