"""Pins the CPU oracle (restated reference algorithms) before it is used as
the GPU checker:
  * against golden fixtures written from the compiled reference CPU code;
  * against constructions that share no code with it (in-order DFS fill of
    the node tree, structural mvEB recursion, brute-force search);
  * against the compiled reference itself when it was built here.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLD, "layout_digests.json")) as f:
        return json.load(f)


def test_small_layouts_match_reference_fixtures():
    data = np.load(os.path.join(GOLD, "layouts_small.npz"))
    assert len(data.files) > 300
    for key in data.files:
        layout, b, n = key.split("_")
        b, n = int(b[1:]), int(n[1:])
        got = ol.layout_ranks(layout, n, b)
        assert np.array_equal(got, data[key].astype(np.uint64)), key


def test_layout_digests_match_reference_fixtures(digests):
    checked = 0
    for key, want in digests.items():
        parts = key.split("_")
        if parts[0] == "veb" or key.endswith("_u32"):
            continue
        layout, b, n = parts[0], int(parts[1][1:]), int(parts[2][1:])
        if n > (1 << 22):
            continue   # the configuration sizes are in the slow test below
        assert ol.digest(ol.layout_ranks(layout, n, b)) == want, key
        checked += 1
    assert checked > 20000


@pytest.mark.slow
def test_config_size_digests(digests):
    a = np.arange((1 << 20) - 1, dtype=np.uint32)
    ol.permute("bst", a)
    assert ol.digest(a.astype(np.uint64)) == digests["bst_b1_n1048575_u32"]
    assert ol.digest(ol.layout_ranks("btree", 10 ** 8, 16)) == digests["btree_b16_n100000000"]


def test_closed_form_equals_inorder_construction():
    """The rank->slot closed form (used inside the GPU kernels) and the oracle's
    restated algorithm both equal the in-order DFS fill of the node tree."""
    o = ol.oracle()
    for b in (1, 2, 3, 4, 7, 8, 15, 16, 17, 31, 32, 33, 64):
        for n in list(range(1, 400)) + [1000, 4095, 4096, 4097, 33 ** 3]:
            ino = ol.inorder_ranks(n, b)
            pos = np.array([o.or_btree_pos(x, n, b) for x in range(n)], dtype=np.int64)
            cf = np.empty(n, dtype=np.uint64)
            cf[pos] = np.arange(n, dtype=np.uint64)
            assert np.array_equal(cf, ino), (b, n)
            assert np.array_equal(ol.layout_ranks("btree", n, b), ino), (b, n)
    for n in range(1, 2000):
        assert np.array_equal(ol.layout_ranks("bst", n), ol.inorder_ranks(n, 1)), n


def test_small_trees_below_b_are_single_nodes():
    """The reference's CPU B-tree recurses forever for n < b
    (cpu/src/btree-cycles.h:51-58); a single partial node is sorted order."""
    for b in (8, 16, 32):
        for n in range(0, b + 1):
            assert np.array_equal(ol.layout_ranks("btree", n, b), np.arange(n, dtype=np.uint64))


def test_float_height_edge_is_defined():
    """At b=32, n=33^3 the reference's double log10 height rounds down and the
    reference crashes; the integer height gives the in-order layout."""
    n, b = 33 ** 3, 32
    assert np.array_equal(ol.layout_ranks("btree", n, b), ol.inorder_ranks(n, b))


def test_mveb_restatement_equals_structural_recursion():
    for n in list(range(1, 1500)) + [4095, 4096, 65535, 65536, 100000, 131071]:
        assert np.array_equal(ol.layout_ranks("mveb", n), ol.mveb_struct_ranks(n)), n


def test_mveb_equals_classic_veb_where_splits_agree(digests):
    """The modified vEB rounds the bottom depth up to a power of two
    (gpu/src/mveb-cycles.cuh:240-247); where the classic split is already a
    power of two at every level the two layouts coincide."""
    for n in (127, 32767, 65535):
        assert ol.digest(ol.layout_ranks("mveb", n)) == digests[f"veb_classic_n{n}"], n
    r = ol.ref()
    if r is not None:
        a = np.arange((1 << 20) - 1, dtype=np.uint64)
        r.ref_permute_veb_u64(a.ctypes.data_as(ol.VP), a.size, 8)
        assert not np.array_equal(a, ol.layout_ranks("mveb", a.size))


def test_mveb_table_search_finds_every_key():
    """Invariant that pins the mvEB layout without a CPU reference: with the
    table set built by the GPU split rule, the reference's position recurrence
    finds every key at its own slot; absent keys answer n."""
    rng = np.random.default_rng(3)
    for n in list(range(1, 300)) + [1023, 1024, 4097, 65535, 65536, 99999, 262144, 1000003]:
        keys = np.arange(1, 2 * n + 1, 2, dtype=np.uint64)
        A = ol.permute("mveb", keys.copy())
        ans = ol.search("mveb", A, A, "u64")
        assert np.array_equal(ans, np.arange(n, dtype=np.uint64)), n
        miss = np.arange(0, 2 * n + 2, 2, dtype=np.uint64)
        assert (ol.search("mveb", A, miss, "u64") == n).all(), n


def test_build_table_hand_checked():
    """n = 7 (d = 3): table[1] = {L=3,R=1,D=0}, table[2] = {L=1,R=1,D=1}; the
    mvEB table equals the reference's classic table for d <= 4 and d = 8."""
    d, nt, ts, tab, idx, tidx = ol.tableset(7)
    assert (d, nt, ts) == (3, 1, 3)
    assert tab[1] == (3, 1, 0) and tab[2] == (1, 1, 1)
    with open(os.path.join(GOLD, "veb_tables_classic.json")) as f:
        classic = json.load(f)
    for dd in (1, 2, 3, 4, 8):
        _, _, _, tab, _, _ = ol.tableset((1 << dd) - 1)
        assert [list(t) for t in tab] == classic[str(dd)], dd


def test_tableset_of_config3_nonperfect():
    """n = 3e8: four structures of depths 29/16/7/3 (gpu/test/query-mveb-fixedN.cu:42-165)."""
    n = 3 * 10 ** 8
    d, nt, ts, tab, idx, tidx = ol.tableset(n)
    assert (d, nt, ts) == (29, 4, 55)
    assert idx == [n - 38114, n - 119, n - 7, n]
    assert tidx == [0, 29, 45, 52]


def test_queries_match_reference_fixtures():
    data = np.load(os.path.join(GOLD, "queries.npz"))
    tags = sorted({k.rsplit("_", 1)[0] for k in data.files})
    assert len(tags) == 6
    for tag in tags:
        layout, b, n = tag.split("_")
        b = int(b[1:])
        A, q, ans = data[tag + "_A"], data[tag + "_q"], data[tag + "_ans"]
        assert np.array_equal(ol.search(layout, A, q, "u64", b), ans), tag


def test_queries_brute_force_all_kinds():
    """Hit -> slot of the key, miss -> n, for every layout and key type."""
    rng = np.random.default_rng(11)
    for kind, dt in (("u32", np.uint32), ("u64", np.uint64), ("i32", np.int32), ("i64", np.int64)):
        for layout, b in (("bst", 1), ("btree", 16), ("btree", 5), ("btree", 1), ("mveb", 0)):
            for n in (1, 2, 9, 100, 777, 4096):
                lo = -1000 if kind[0] == "i" else 0
                keys = np.sort(rng.choice(np.arange(lo, lo + 10 * n), size=n, replace=False)).astype(dt)
                A = ol.permute(layout, keys.copy(), b)
                q = np.arange(lo - 3, lo + 10 * n + 3).astype(dt)
                ans = ol.search(layout, A, q, kind, b)
                slot = {int(v): i for i, v in enumerate(A)}
                exp = np.array([slot.get(int(v), n) for v in q], dtype=np.uint64)
                assert np.array_equal(ans, exp), (kind, layout, b, n)


def test_classic_veb_matches_reference_fixtures():
    """The classic (CPU) vEB restatement against the compiled reference's
    digests (every n <= 4096, 2^k-2 .. 2^k+1 up to 2^24, a few others)."""
    with open(os.path.join(GOLD, "veb_classic_digests.json")) as f:
        dig = json.load(f)
    assert len(dig) > 4000
    for key, want in dig.items():
        n = int(key.split("_n")[1])
        if n > 1 << 20:
            continue   # the largest sizes run in the slow test below
        assert ol.digest(ol.layout_ranks("veb", n, 0, p=8 if n > 4096 else 1)) == want, n


@pytest.mark.slow
def test_classic_veb_large_fixtures():
    with open(os.path.join(GOLD, "veb_classic_digests.json")) as f:
        dig = json.load(f)
    for key, want in dig.items():
        n = int(key.split("_n")[1])
        if n > 1 << 20:
            assert ol.digest(ol.layout_ranks("veb", n, 0, p=8)) == want, n


def test_classic_veb_equals_structural_recursion():
    for n in list(range(1, 600)) + [4095, 4096, 5000, 65535, 65536, 100003, 1 << 20]:
        assert np.array_equal(ol.layout_ranks("veb", n, 0), ol.classic_struct_ranks(n)), n


def test_classic_tables_match_reference_fixtures():
    with open(os.path.join(GOLD, "veb_tables_classic.json")) as f:
        want = json.load(f)
    for d, exp in want.items():
        d = int(d)
        t = (ol.VebTable * d)()
        ol.oracle().or_build_table_classic(t, (1 << d) - 1, d, 0)
        assert [[int(e.L), int(e.R), int(e.D)] for e in t] == exp, d


def test_classic_table_search_finds_every_key():
    """Every key of the classic layout is found at its own slot with the
    classic table set; keys in the gaps answer n."""
    for n in list(range(1, 300)) + [1000, 4095, 4096, 5000, 65535, 65536, 100003]:
        keys = np.arange(1, n + 1, dtype=np.uint64) * 3
        A = keys[ol.layout_ranks("veb", n, 0).astype(np.int64)]
        assert np.array_equal(ol.search("veb", A, A, "u64"), np.arange(n, dtype=np.uint64)), n
        assert (ol.search("veb", A, keys + 1, "u64") == n).all(), n


def test_binary_search_matches_reference_fixtures():
    data = np.load(os.path.join(GOLD, "binary_search.npz"))
    tags = sorted({k.rsplit("_", 1)[0] for k in data.files})
    assert len(tags) == 8
    for tag in tags:
        A, q, ans = data[tag + "_A"], data[tag + "_q"], data[tag + "_ans"]
        assert np.array_equal(ol.search("binary", A, q, "u64"), ans), tag


def test_binary_search_brute_force_all_kinds():
    """Hit -> index, miss -> index of the predecessor (numpy's searchsorted
    from the right, minus one), no predecessor -> n; n = 0 answers 0."""
    rng = np.random.default_rng(12)
    for kind, dt in (("u32", np.uint32), ("u64", np.uint64), ("i32", np.int32), ("i64", np.int64)):
        for n in (0, 1, 2, 3, 9, 100, 777, 4096):
            lo = -1000 if kind[0] == "i" else 5
            keys = np.sort(rng.choice(np.arange(lo, lo + 10 * n + 1), size=n,
                                      replace=False)).astype(dt)
            q = np.arange(lo - 5, lo + 10 * n + 6).astype(dt)
            ans = ol.search("binary", keys, q, kind)
            pred = np.searchsorted(keys, q, side="right").astype(np.int64) - 1
            exp = np.where(pred < 0, n, pred).astype(np.uint64)
            assert np.array_equal(ans, exp), (kind, n)
    # extremes of the key range
    for kind, dt in (("u64", np.uint64), ("i64", np.int64), ("u32", np.uint32), ("i32", np.int32)):
        info = np.iinfo(dt)
        keys = np.array([info.min, info.min + 1, 0 if info.min < 0 else 7, info.max - 1, info.max],
                        dtype=dt)
        q = np.array([info.min, info.max, info.max - 2, info.min + 2], dtype=dt)
        exp = np.searchsorted(keys, q, side="right") - 1
        assert np.array_equal(ol.search("binary", keys, q, kind), exp.astype(np.uint64)), kind


def test_parallel_restatement_equals_serial():
    for layout, b in (("bst", 1), ("btree", 16), ("btree", 8)):
        for n in (1000, 65536, 100003, 1 << 20):
            a1 = np.arange(n, dtype=np.uint64)
            a8 = np.arange(n, dtype=np.uint64)
            ol.permute(layout, a1, b, p=1)
            ol.permute(layout, a8, b, p=8)
            assert np.array_equal(a1, a8), (layout, b, n)


@pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="reference not compiled in this environment")
def test_live_reference_binary_search_agrees():
    rng = np.random.default_rng(13)
    for n in (1, 2, 5, 64, 1000, 65537, 1 << 20):
        keys = np.sort(rng.choice(np.arange(1, 4 * n + 10, dtype=np.uint64), size=n,
                                  replace=False))
        q = np.concatenate([keys[rng.integers(0, n, size=3000)],
                            rng.integers(int(keys[0]), 4 * n + 20, size=3000).astype(np.uint64)])
        ans = np.zeros(q.size, dtype=np.uint64)
        ol.ref().ref_search_binary_u64(keys.ctypes.data_as(ol.VP), n, q.ctypes.data_as(ol.VP),
                                       q.size, ans.ctypes.data_as(ol.VP), 8)
        assert np.array_equal(ol.search("binary", keys, q, "u64"), ans), n


@pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="reference not compiled in this environment")
def test_live_reference_classic_veb_agrees():
    for n in (1, 2, 3, 7, 100, 127, 128, 1000, 32767, 32768, 99999, 300007):
        a = np.arange(n, dtype=np.uint64)
        ol.ref().ref_permute_veb_u64(a.ctypes.data_as(ol.VP), n, 8)
        assert np.array_equal(a, ol.layout_ranks("veb", n, 0)), n


@pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="reference not compiled in this environment")
def test_live_reference_agrees():
    for b in (1, 2, 8, 16, 32):
        for n in (b, b + 1, 500, 5000, 17 ** 4, 17 ** 4 + 1, 300007):
            if (b, n) == (32, 33 ** 3):
                continue
            a = np.arange(n, dtype=np.uint64)
            ol.ref().ref_permute_btree_u64(a.ctypes.data_as(ol.VP), n, b, 8)
            assert np.array_equal(a, ol.layout_ranks("btree", n, b)), (b, n)


# ---- the involution family ----------------------------------------------

def _prim(name, a, *args):
    o = ol.oracle()
    getattr(o, f"or_{name}_{ol.width_suffix(a)}")(a.ctypes.data_as(ol.VP), *args)
    return a


@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
def test_oracle_involution_primitives_closed_forms(dt):
    """The restated shuffle primitives against what the mathematics fixes:
    the k-way shuffle of n = d*k keys is the transpose of a k x d matrix, its
    inverse the transpose of a d x k one, the unshuffle of k^d - 1 keys the
    stable partition by (i+1) mod k, the rotation numpy's roll."""
    for k, d in [(1, 9), (2, 1), (2, 7), (3, 5), (4, 4), (17, 6), (16, 15), (5, 100)]:
        n = k * d
        a = np.arange(n)
        got = _prim("shuffle_dk", np.arange(n, dtype=dt), k, n)
        assert np.array_equal(got, a.reshape(k, d).T.ravel()), (k, d)
        got = _prim("unshuffle_dk", np.arange(n, dtype=dt), k, n)
        assert np.array_equal(got, a.reshape(d, k).T.ravel()), (k, d)
    for k, d in [(2, 1), (2, 9), (3, 5), (17, 3), (16, 2), (4, 2), (256, 2)]:
        n = k ** d - 1
        i = np.arange(n)
        exp = np.concatenate([i[(i + 1) % k == r] for r in range(k)])
        assert np.array_equal(_prim("unshuffle", np.arange(n, dtype=dt), k, n, d), exp), (k, d)
    for n in (1, 2, 10, 1001):
        for k in sorted({0, 1, n // 3, n - 1, n}):
            got = _prim("rotate_right", np.arange(n, dtype=dt), n, k)
            assert np.array_equal(got, np.roll(np.arange(n), k)), (n, k)
    # digit reversal and the modular involution on their own
    o = ol.oracle()
    assert o.or_rev_base_d(int("1203", 4), 4, 4) == int("3021", 4)
    assert o.or_rev_base_b(int("1203", 4), 4, 4, 2) == int("1230", 4)
    for m in (12, 17, 100, 255):
        for x in range(1, m):
            y = o.or_involution(1, x, m)
            assert o.or_involution(1, y, m) == x                 # an involution
            g = np.gcd(x, m)
            assert (x * y) % m == (g * g) % m, (x, m)            # y = g^2 / x (mod m)


def test_oracle_involution_permutes_match_layouts():
    """The restated involution programs lay out exactly the trees of the
    independent in-order and structural constructions (family 2), including
    the B-tree sizes where the reference's empty shuffle does not end."""
    sizes = list(range(0, 400)) + [1000, 4095, 4096, 4913, 4914, 4929, 20000, 65535, 65536,
                                   70000, 83521, 100003]
    for n in sizes:
        a = ol.permute_involutions("bst", np.arange(n, dtype=np.uint64))
        assert np.array_equal(a, ol.layout_ranks("bst", n, 1)), ("bst", n)
        a = ol.permute_involutions("mveb", np.arange(n, dtype=np.uint64))
        assert np.array_equal(a, ol.layout_ranks("mveb", n)), ("mveb", n)
        for b in (1, 2, 3, 8, 16):
            if n > 20000 and b != 16:
                continue
            a = ol.permute_involutions("btree", np.arange(n, dtype=np.uint32), b)
            assert np.array_equal(a, ol.layout_ranks("btree", n, b).astype(np.uint32)), (b, n)


def _btree_inv_hangs(n, b):
    """n in [(b+1)^h, (b+1)^h + b - 2]: the reference's B-tree involution
    program shuffles an empty range there, and its loop over [1, n - 1) with
    n = 0 does not end (cpu/src/involutions.h:162-164)"""
    p = b + 1
    while p * (b + 1) <= n:
        p *= b + 1
    return p <= n <= p + b - 2


@pytest.mark.skipif(not os.path.exists(ol.REF_SO), reason="reference not compiled in this environment")
def test_live_reference_involutions_agree():
    r = ol.ref()
    for n in (1, 2, 3, 6, 7, 8, 100, 1000, 4095, 4096, 65535, 100003, 300007):
        a = np.arange(n, dtype=np.uint64)
        r.ref_permute_bst_inv_u64(a.ctypes.data_as(ol.VP), n, 1)
        assert np.array_equal(a, ol.permute_involutions("bst", np.arange(n, dtype=np.uint64))), n
    for b in (1, 2, 8, 16):
        for n in (b + 1, 100, 500, 5000, 17 ** 4 - 1, 17 ** 4 + 20, 100003):
            if _btree_inv_hangs(n, b):
                continue
            a = np.arange(n, dtype=np.uint64)
            r.ref_permute_btree_inv_u64(a.ctypes.data_as(ol.VP), n, b, 1)
            exp = ol.permute_involutions("btree", np.arange(n, dtype=np.uint64), b)
            assert np.array_equal(a, exp), (b, n)
    for k, d in [(2, 8), (3, 5), (17, 3)]:
        n = k ** d - 1
        a = np.arange(n, dtype=np.uint64)
        r.ref_unshuffle_u64(a.ctypes.data_as(ol.VP), k, n, d)
        assert np.array_equal(a, _prim("unshuffle", np.arange(n, dtype=np.uint64), k, n, d))
    for k, n in [(2, 10), (3, 300), (17, 17 * 99), (16, 4096)]:
        a = np.arange(n, dtype=np.uint64)
        r.ref_shuffle_dk_u64(a.ctypes.data_as(ol.VP), k, n)
        assert np.array_equal(a, _prim("shuffle_dk", np.arange(n, dtype=np.uint64), k, n))
        a = np.arange(n, dtype=np.uint64)
        r.ref_unshuffle_dk_u64(a.ctypes.data_as(ol.VP), k, n)
        assert np.array_equal(a, _prim("unshuffle_dk", np.arange(n, dtype=np.uint64), k, n))
