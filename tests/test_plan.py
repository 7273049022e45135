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
    # the allocation-free permutation rejects sizes past its limit before it
    # touches the device (the pointer is never dereferenced)
    L = tl_amd.lib()
    st = L.tl_permute_btree_noalloc(1 << 20, tl_amd.TL_NOALLOC_MAX_KEYS + 1, 16,
                                    tl_amd.DTYPE["u64"], None)
    assert st == tl_amd.TL_EINVAL
    assert "TL_NOALLOC_MAX_KEYS" in L.tl_last_error().decode()


@pytest.mark.parametrize("layout,bs", [("btree", (16, 8, 2, 3, 32, 1)), ("bst", (1,)),
                                       ("mveb", (0,)), ("veb", (0,))])
@pytest.mark.parametrize("lds", [256, 2048])
@pytest.mark.parametrize("row_stash", [None, "0", "40"])
def test_plans_replay_to_oracle(layout, bs, lds, row_stash, monkeypatch):
    """row_stash: the library's default stash budget for row slides, none (the
    hierarchical LDS + unit-merge gathers), or 40 separators (both kinds in
    one plan)"""
    if row_stash is not None:
        monkeypatch.setenv("TL_ROW_STASH_MAX", row_stash)
    sizes = [1, 2, 3, 7, 16, 17, 100, 255, 256, 1000, 1023, 4095, 4096, 5000, 8192,
             17 ** 3 - 1, 17 ** 3, 33333, 65535, 65536]
    for b in bs:
        for n in sizes:
            buf = tl_amd.plan_export(layout, n, b, "u64", lds)
            A = np.arange(n, dtype=np.int64)
            pe.replay(buf, A)
            assert np.array_equal(A, ol.layout_ranks(layout, n, b).astype(np.int64)), \
                (layout, b, lds, n)


@pytest.mark.parametrize("layout", ["mveb", "veb"])
def test_fused_sparse_slide_plans_replay(layout):
    """Trees whose top-level gather is a chained sparse slide carry their leaf
    blocks' own gathers in it (perfect and non-perfect trees): the plans hold
    such a fused slide, replay to the oracle's layout, their inverses replay
    to sorted order, and forward then inverse leaves arbitrary data as it was"""
    rng = np.random.default_rng(41)
    for n in [300007, 1048575, 1048576, 3_000_017]:
        buf = tl_amd.plan_export(layout, n, 0, "u64", 0)
        ops, _ = pe.decode(buf)
        assert any(int(o[0]) == pe.OP_CHAINED and int(o[20]) for o in ops), (layout, n)
        A = np.arange(n, dtype=np.int64)
        pe.replay(buf, A)
        assert np.array_equal(A, ol.layout_ranks(layout, n, 0).astype(np.int64)), (layout, n)
        inv = tl_amd.plan_export(layout, n, 0, "u64", 0, inverse=True)
        pe.replay(inv, A)
        assert np.array_equal(A, np.arange(n)), (layout, n)
        X = rng.integers(-(1 << 40), 1 << 40, size=n)
        Y = X.copy()
        pe.replay(buf, Y)
        pe.replay(inv, Y)
        assert np.array_equal(X, Y), (layout, n)


def test_full_size_plans_are_in_bounds():
    for layout, b, n in [("btree", 16, 1 << 30), ("btree", 16, 10 ** 8), ("bst", 1, 1 << 30),
                         ("bst", 1, 3 << 28), ("mveb", 0, (1 << 29) - 1),
                         ("mveb", 0, 3 * 10 ** 8), ("mveb", 0, 1 << 30), ("btree", 8, 1 << 22),
                         ("veb", 0, (1 << 29) - 1), ("veb", 0, 3 * 10 ** 8), ("veb", 0, 1 << 30)]:
        for kind in ("u32", "u64"):
            buf = tl_amd.plan_export(layout, n, b, kind)
            assert int(buf[0]) > 0


@pytest.mark.parametrize("layout,bs", [("btree", (16, 8, 2, 3, 32, 1)), ("bst", (1,)),
                                       ("mveb", (0,)), ("veb", (0,))])
@pytest.mark.parametrize("lds", [256, 2048])
@pytest.mark.parametrize("row_stash", [None, "0", "40"])
def test_inverse_plans_replay_to_identity(layout, bs, lds, row_stash, monkeypatch):
    """The inverse plan takes every layout back to sorted order, and the
    forward then inverse plan leaves arbitrary data unchanged."""
    if row_stash is not None:
        monkeypatch.setenv("TL_ROW_STASH_MAX", row_stash)
    rng = np.random.default_rng(31)
    sizes = [1, 2, 3, 7, 16, 17, 100, 255, 256, 1000, 1023, 4095, 4096, 5000, 8192,
             17 ** 3 - 1, 17 ** 3, 33333, 65535, 65536]
    for b in bs:
        for n in sizes:
            inv = tl_amd.plan_export(layout, n, b, "u64", lds, inverse=True)
            A = ol.layout_ranks(layout, n, b).astype(np.int64)
            pe.replay(inv, A)
            assert np.array_equal(A, np.arange(n)), (layout, b, lds, n)
            X = rng.integers(-(1 << 40), 1 << 40, size=n)
            Y = X.copy()
            pe.replay(tl_amd.plan_export(layout, n, b, "u64", lds), Y)
            pe.replay(inv, Y)
            assert np.array_equal(X, Y), (layout, b, lds, n)


class _Move:
    def __init__(self, op):
        (self.kind, self.P, self.L, self.k, self.r, self.l, self.CH,
         self.nchunk) = (int(v) for v in op[11:19])
        self.fuse, self.rp, self.lp, self.soff = (int(v) for v in op[20:24])
        self.nmoves = int(op[3])

    def fused(self, t):
        """a fused leaf-block move of a sparse slide or spread"""
        if not self.fuse:
            return False
        if self.kind == pe.MV_SPARSE_SPREAD:
            return t < self.r * self.nchunk
        return t >= (self.L + self.CH - 1) // self.CH


def _wait_len(m, t, ln):
    """the destination length a move waits on (a fused slide move writes its
    leaf keys only)"""
    if m.kind == pe.MV_SPARSE_SLIDE and m.fused(t):
        return ln - ln // (m.lp + 1)
    return ln


def _written(m, dst, ln):
    """exact key intervals a move stores into the array"""
    if m.kind == pe.MV_ROW_SPREAD:   # keys and the separators from the stash
        return [(dst, dst + ln // m.l * (m.l + 1))]
    if m.kind == pe.MV_ROW_SLIDE:
        return [(dst, dst + ln // (m.l + 1) * m.l)]
    return [(dst, dst + ln)]


@pytest.mark.parametrize("inverse", [False, True])
def test_chained_moves_wait_for_every_overwritten_source(inverse):
    """Every chained move of forward and inverse plans: a move that stores over
    the source of another move has the larger number and waits for it (the
    device's wait_deps, mirrored in the batch emulator), and replaying the
    moves one after another in number order gives the op's specified effect."""
    import bisect
    import batch_emulator as bem
    cases = [("btree", 16, 10 ** 6), ("btree", 16, 17 ** 4 + 12345), ("btree", 2, 300001),
             ("bst", 1, 1 << 21), ("bst", 1, 3 << 19), ("mveb", 0, (1 << 24) - 1),
             ("mveb", 0, (1 << 24) + 12345), ("mveb", 0, 5 * 10 ** 6),
             ("veb", 0, (1 << 23) - 1), ("veb", 0, (1 << 24) + 12345)]
    kinds_seen = set()
    for layout, b, n in cases:
        buf = tl_amd.plan_export(layout, n, b, "u64", 0, inverse=inverse)
        ops, blob = pe.decode(buf)
        for op in ops:
            if int(op[0]) != pe.OP_CHAINED:
                continue
            m = _Move(op)
            kinds_seen.add(m.kind)
            moves = [bem.move_at(m, t) for t in range(m.nmoves)]
            srcs = sorted((s, s + ln, t) for t, (s, d, ln) in enumerate(moves))
            starts = [x[0] for x in srcs]
            for t, (s, d, ln) in enumerate(moves):
                wl = _wait_len(m, t, ln)
                waited = set(bem.awaited_moves(m, t, d, wl))
                for a, z in _written(m, d, wl):
                    i = max(0, bisect.bisect_right(starts, a) - 1)
                    while i < len(srcs) and srcs[i][0] < z:
                        s0, s1, u = srcs[i]
                        if s1 > a and u != t:
                            assert u < t and u in waited, (layout, n, m.kind, t, u)
                        i += 1
            # sequential replay in move order against the specified effect
            one = np.concatenate([buf[:2], op.astype(np.uint64)]).astype(np.uint64)
            one[0], one[1] = 1, 0
            # stash contents for a row spread or a fused spread's roots
            sep = -1 - np.arange(max(m.r, m.soff + m.r * m.rp), dtype=np.int64)
            spec = np.arange(n, dtype=np.int64)
            pe.replay(one, spec, stash=sep.copy())
            seq = np.arange(n, dtype=np.int64)
            pos = None
            if m.fuse:
                pos = pe.lds_perm(pe.LDS_MVEB if m.fuse == 1 else pe.LDS_VEB, m.lp, 0, 0)
            for t, (s, d, ln) in enumerate(moves):
                v = seq[s:s + ln].copy()
                if m.fused(t):
                    L1 = m.lp + 1
                    idx = np.arange(ln)
                    a, c = idx // L1, idx % L1
                    if m.kind == pe.MV_SPARSE_SLIDE:   # leaves to their slots
                        keep = c < m.lp
                        seq[d + a[keep] * m.lp + pos[c[keep]]] = v[keep]
                    else:   # rows back, the roots from the stash
                        rel = d - m.P
                        i = rel // (m.l + 1)
                        a0 = (rel - i * (m.l + 1)) // L1
                        leaf = c < m.lp
                        out = np.empty(ln, dtype=np.int64)
                        out[leaf] = seq[s + a[leaf] * m.lp + pos[c[leaf]]]
                        out[~leaf] = sep[m.soff + i * m.rp + a0 + a[~leaf]]
                        seq[d:d + ln] = out
                    continue
                if m.kind == pe.MV_ROW_SLIDE:
                    rows = v.reshape(-1, m.l + 1)[:, :m.l].reshape(-1)
                    seq[d:d + rows.size] = rows
                elif m.kind == pe.MV_ROW_SPREAD:
                    rows = seq[d:d + ln // m.l * (m.l + 1)].reshape(-1, m.l + 1)
                    rows[:, :m.l] = v.reshape(-1, m.l)
                    j0 = (s - m.P - m.r) // m.l
                    rows[:, m.l] = sep[j0:j0 + ln // m.l]
                else:
                    seq[d:d + ln] = v
            assert np.array_equal(seq, spec), (layout, n, m.kind)
    if inverse:
        assert {pe.MV_ROW_SPREAD, pe.MV_SPARSE_SPREAD, pe.MV_SHIFT_LEFT} <= kinds_seen
    else:
        assert {pe.MV_ROW_SLIDE, pe.MV_SPARSE_SLIDE, pe.MV_SHIFT_RIGHT} <= kinds_seen


def test_host_table_builds_match_oracle():
    """tl_veb_build_table (mvEB rule) and tl_veb_classic_build_table (classic
    rule, checked against the compiled reference's tables) on the host."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "veb_tables_classic.json")) as f:
        want = json.load(f)
    for d, exp in want.items():
        d = int(d)
        got = tl_amd.veb_build_table((1 << d) - 1, d, classic=True)
        assert [[int(a), int(b), int(c)] for a, b, c in got] == exp, d
    for d in range(1, 40):
        t = (ol.VebTable * d)()
        ol.oracle().or_build_table(t, (1 << d) - 1, d, 0)
        got = tl_amd.veb_build_table((1 << d) - 1, d)
        assert got == [(e.L, e.R, e.D) for e in t], d


def _walk_model(L, c, w, K=8):
    """The lane walk of the one-pass rotation (k_rotate_walk), lane by lane:
    returns the rotated array and how many times each position was written,
    and checks that no lane reads a position another lane already wrote."""
    A = list(range(L))
    orig = list(A)
    stash = orig[:w]
    writes = [0] * L
    owner = [None] * L
    for j in range(w):
        dst = j
        while True:
            s, stop, d = [], False, dst
            for _ in range(K):
                src = d - c if d >= c else d + L - c
                s.append(src)
                if src < w:
                    stop = True
                    break
                d = src
            vals = []
            for k, src in enumerate(s):
                if stop and k == len(s) - 1:
                    vals.append(stash[src])
                else:
                    assert owner[src] is None, ("read after another lane wrote", L, c, w, j, src)
                    vals.append(A[src])
            for k, v in enumerate(vals):
                pos = dst if k == 0 else s[k - 1]
                A[pos] = v
                writes[pos] += 1
                owner[pos] = j
            if stop:
                break
            dst = s[-1]
    return A, writes


def test_rotation_lane_walk_model():
    """Every (L, c) up to L = 40 and every lane count w >= gcd(L, c): the walk
    writes each position once and yields the rotation (numpy.roll)."""
    import math
    for L in range(2, 41):
        for c in range(1, L):
            g = math.gcd(L, c)
            for w in sorted({g, g + 1, min(L, 2 * g + 3), L - 1, L}):
                if w < g or w < 1 or w > L:
                    continue
                for K in (1, 3, 8):
                    A, writes = _walk_model(L, c, w, K)
                    assert writes == [1] * L, (L, c, w, K)
                    assert A == list(np.roll(np.arange(L), c)), (L, c, w, K)


def test_layout_file_header_errors_without_device(tmp_path):
    """tl_layout_read_info checks the header on the host: a missing file, a
    file that is not a layout file, and a truncated header are TL_EIO"""
    L = tl_amd.lib()
    info = tl_amd.LayoutInfo()
    p = os.path.join(str(tmp_path), "x.tl")
    assert L.tl_layout_read_info(p.encode(), info) == tl_amd.TL_EIO
    open(p, "wb").write(b"NOTALAYOUT" + bytes(300))
    assert L.tl_layout_read_info(p.encode(), info) == tl_amd.TL_EIO
    assert "not a layout file" in L.tl_last_error().decode()
    open(p, "wb").write(b"TLLAYOUT" + bytes(20))
    assert L.tl_layout_read_info(p.encode(), info) == tl_amd.TL_EIO
