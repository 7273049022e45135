"""Writes the golden fixtures of this directory from the compiled reference
CPU code (oracle/_ref/libref.so, built by `make -C oracle ref` from
/root/reference; it exists only in the build container).

Fixtures are data only: rank layouts (the slot -> rank map of permuting
0..n-1), order-sensitive digests of larger rank layouts, and query answers.
Layouts do not depend on key values, so a rank layout fixes the layout of any
sorted key set of that size.

Run:  python tests/golden/make_golden.py           (every fixture)
      python tests/golden/make_golden.py binary    (binary_search.npz and
                                                    veb_classic_digests.json only)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

P = 8  # OpenMP threads of the reference


def ref_layout(layout, n, b, dtype=np.uint64):
    r = ol.ref()
    a = np.arange(n, dtype=dtype)
    w = "u32" if dtype == np.uint32 else "u64"
    if layout == "bst":
        getattr(r, f"ref_permute_bst_{w}")(a.ctypes.data_as(ol.VP), n, P)
    elif layout == "btree":
        getattr(r, f"ref_permute_btree_{w}")(a.ctypes.data_as(ol.VP), n, b, P)
    elif layout == "veb_classic":
        r.ref_permute_veb_u64(a.ctypes.data_as(ol.VP), n, P)
    return a.astype(np.uint64)


def edge_sizes(b, limit):
    k1 = b + 1
    out = set()
    p = k1
    while p <= limit:
        for d in (-1, 0, 1):
            if p + d >= max(b, 1):
                out.add(p + d)
        p *= k1
    return sorted(out)


# the reference computes the tree height as log10(n)/log10(b+1) in double;
# at these sizes it rounds down and the reference then crashes
REF_CRASH = {(32, 33 ** 3)}


def binary_search_fixtures():
    """Answers of the reference's CPU binary search (cpu/src/binary-search.h)
    on sorted arrays: hits, and misses above the smallest key (below it the
    reference's loop does not end, so those are not asked)."""
    r = ol.ref()
    rng = np.random.default_rng(20240602)
    out = {}
    for n in (1, 2, 3, 7, 8, 1000, 4096, 5001):
        keys = np.sort(rng.choice(np.arange(1, 10 * n + 10, dtype=np.uint64), size=n,
                                  replace=False))
        hits = keys[rng.integers(0, n, size=1000)]
        miss = rng.integers(int(keys[0]), 10 * n + 20, size=1000).astype(np.uint64)
        q = np.concatenate([hits, miss])
        ans = np.zeros(q.size, dtype=np.uint64)
        r.ref_search_binary_u64(keys.ctypes.data_as(ol.VP), n, q.ctypes.data_as(ol.VP), q.size,
                                ans.ctypes.data_as(ol.VP), 1)
        out[f"binary_n{n}_A"] = keys
        out[f"binary_n{n}_q"] = q
        out[f"binary_n{n}_ans"] = ans
    np.savez_compressed(os.path.join(HERE, "binary_search.npz"), **out)
    print("wrote", len(out) // 3, "binary-search cases")


def classic_veb_fixtures():
    """Digests of the reference's CPU classic vEB permutation
    (cpu/src/veb-cycles.h:381) of 0..n-1, perfect and non-perfect."""
    out = {}
    sizes = set(range(1, 4097))
    for k in range(2, 25):
        sizes.update({(1 << k) - 2, (1 << k) - 1, 1 << k, (1 << k) + 1})
    sizes.update({10 ** 6, 3 * 10 ** 6 + 7, 12345678})
    for n in sorted(sizes):
        out[f"veb_n{n}"] = ol.digest(ref_layout("veb_classic", n, 0))
    with open(os.path.join(HERE, "veb_classic_digests.json"), "w") as f:
        json.dump(out, f)
    print("wrote", len(out), "classic vEB digests")


def main():
    if ol.ref() is None:
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    binary_search_fixtures()
    classic_veb_fixtures()
    if sys.argv[1:] in (["binary"], ["veb"]):
        return
    small = {}
    for layout, b in [("bst", 1), ("btree", 2), ("btree", 3), ("btree", 8), ("btree", 16),
                      ("btree", 32)]:
        for n in range(max(b, 1), 65):
            small[f"{layout}_b{b}_n{n}"] = ref_layout(layout, n, b).astype(np.uint16)
    np.savez_compressed(os.path.join(HERE, "layouts_small.npz"), **small)

    digests = {}
    for layout, b in [("bst", 1), ("btree", 2), ("btree", 8), ("btree", 16), ("btree", 32)]:
        for n in range(max(b, 1), 4097):
            digests[f"{layout}_b{b}_n{n}"] = ol.digest(ref_layout(layout, n, b))
        for n in edge_sizes(b, 1 << 22):
            if (b, n) in REF_CRASH:
                continue
            digests[f"{layout}_b{b}_n{n}"] = ol.digest(ref_layout(layout, n, b))
    for n in (127, 32767, 65535):
        digests[f"veb_classic_n{n}"] = ol.digest(ref_layout("veb_classic", n, 0))
    # the sizes of the benchmark configurations (uint32 for the BST one)
    digests["bst_b1_n1048575_u32"] = ol.digest(ref_layout("bst", (1 << 20) - 1, 1, np.uint32))
    digests["btree_b16_n100000000"] = ol.digest(ref_layout("btree", 10 ** 8, 16))
    if os.environ.get("GOLDEN_BIG", "1") == "1":
        digests["btree_b16_n1073741824"] = ol.digest(ref_layout("btree", 1 << 30, 16))
        digests["bst_b1_n1073741824"] = ol.digest(ref_layout("bst", 1 << 30, 1))
    with open(os.path.join(HERE, "layout_digests.json"), "w") as f:
        json.dump(digests, f, indent=0, sort_keys=True)

    # query answers of the reference's CPU search loops on seeded batches:
    # hits and misses on perfect trees (the reference reads past n for misses
    # that end in a partial last node), hits only on non-perfect trees
    r = ol.ref()
    rng = np.random.default_rng(20240601)
    q_out = {}
    for layout, b, n, misses in [("bst", 1, 4095, True), ("bst", 1, 5000, True),
                                 ("btree", 16, 17 ** 3 - 1, True), ("btree", 16, 6000, False),
                                 ("btree", 8, 9 ** 4 - 1, True), ("btree", 8, 7777, False)]:
        keys = np.sort(rng.choice(np.arange(1, 10 * n, dtype=np.uint64), size=n, replace=False))
        A = keys[ref_layout(layout, n, b).astype(np.int64)]
        q = keys[rng.integers(0, n, size=2000)]
        if misses:
            miss = rng.integers(0, 10 * n + 5, size=2000).astype(np.uint64)
            q = np.concatenate([q, miss])
        ans = np.zeros(q.size, dtype=np.uint64)
        if layout == "bst":
            r.ref_search_bst_u64(A.ctypes.data_as(ol.VP), n, q.ctypes.data_as(ol.VP), q.size,
                                 ans.ctypes.data_as(ol.VP), 1)
        else:
            r.ref_search_btree_u64(A.ctypes.data_as(ol.VP), n, b, q.ctypes.data_as(ol.VP), q.size,
                                   ans.ctypes.data_as(ol.VP), 1)
        tag = f"{layout}_b{b}_n{n}"
        q_out[f"{tag}_A"] = A
        q_out[f"{tag}_q"] = q
        q_out[f"{tag}_ans"] = ans
    np.savez_compressed(os.path.join(HERE, "queries.npz"), **q_out)

    # classic vEB tables of the reference for small d (equal to the mvEB
    # tables wherever the two split rules agree)
    tables = {}
    for d in range(1, 13):
        t = (ol.VebTable * d)()
        r.ref_build_table_classic(ol.ctypes.cast(t, ol.VP), (1 << d) - 1, d)
        tables[str(d)] = [[int(e.L), int(e.R), int(e.D)] for e in t]
    with open(os.path.join(HERE, "veb_tables_classic.json"), "w") as f:
        json.dump(tables, f)
    print("wrote", len(small), "small layouts,", len(digests), "digests")


if __name__ == "__main__":
    main()
