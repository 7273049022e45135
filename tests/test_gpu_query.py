"""GPU parity of the batched queries against the oracle (hits and misses),
and of the device-built mvEB table set against the oracle's."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu

NP = {"u32": np.uint32, "u64": np.uint64, "i32": np.int32, "i64": np.int64}
STORE = {"u32": (np.int32, torch.int32), "u64": (np.int64, torch.int64),
         "i32": (np.int32, torch.int32), "i64": (np.int64, torch.int64)}
SIZES = [1, 2, 3, 7, 15, 16, 17, 100, 255, 1000, 4095, 4096, 65535, 65536, 100003, 1048575,
         1000003]


def make_keys(kind, n, rng):
    info = np.iinfo(NP[kind])
    lo = max(int(info.min), -(1 << 62))
    hi = min(int(info.max), (1 << 62))
    keys = np.unique(rng.integers(lo, hi, size=int(n * 1.2) + 16, dtype=np.int64))
    while keys.size < n:
        keys = np.unique(np.concatenate([keys, rng.integers(lo, hi, size=n, dtype=np.int64)]))
    keys = np.sort(rng.choice(keys, size=n, replace=False)).astype(NP[kind])
    return keys


def make_queries(kind, keys, nq, rng):
    hits = keys[rng.integers(0, keys.size, size=nq // 2)]
    info = np.iinfo(NP[kind])
    lo = max(int(info.min), -(1 << 62))
    hi = min(int(info.max), (1 << 62))
    cand = rng.integers(lo, hi, size=nq - nq // 2, dtype=np.int64).astype(NP[kind])
    edge = np.array([keys[0], keys[-1]], dtype=NP[kind])
    q = np.concatenate([hits, cand, edge])
    rng.shuffle(q)
    return q


def to_dev(a, kind, dev):
    st_np, st_t = STORE[kind]
    return torch.from_numpy(np.ascontiguousarray(a).view(st_np).copy()).to(dev)


def gpu_answers(layout, b, kind, A_np, Q_np, dev):
    A = to_dev(A_np, kind, dev)
    Q = to_dev(Q_np, kind, dev)
    ans = torch.zeros(Q_np.size, dtype=torch.int64, device=dev)
    n = A_np.size
    if layout == "bst":
        tl_amd.query_bst(A, Q, ans, kind=kind)
    elif layout == "btree":
        tl_amd.query_btree(A, b, Q, ans, kind=kind)
    else:
        classic = layout == "veb"
        tables, idx, tidx, meta = tl_amd.veb_tableset_device(n, device=dev, classic=classic)
        d, nt, ts = (int(v) for v in meta[:3].cpu())
        if nt == 1 and n == (1 << d) - 1:
            (tl_amd.query_veb if classic else tl_amd.query_mveb)(A, tables, d, Q, ans, kind=kind)
        elif classic:
            tl_amd.query_veb_nonperfect(A, tables, idx, tidx, d, nt, Q, ans, kind=kind)
        else:
            tl_amd.query_mveb_nonperfect(A, tables, idx, tidx, d, nt, Q, ans, kind=kind)
    torch.cuda.synchronize()
    return ans.cpu().numpy().view(np.uint64)


QCASES = [("bst", 1), ("btree", 16), ("btree", 8), ("btree", 2), ("btree", 5), ("btree", 32),
          ("btree", 64), ("btree", 1), ("btree", 100), ("mveb", 0), ("veb", 0)]


@pytest.mark.parametrize("layout,b", QCASES)
@pytest.mark.parametrize("kind", ["u32", "u64", "i32", "i64"])
def test_query_matches_oracle(layout, b, kind, dev):
    rng = np.random.default_rng(1234 + b)
    for n in SIZES:
        keys = make_keys(kind, n, rng)
        A = ol.permute(layout, keys.copy(), b)
        Q = make_queries(kind, keys, 4096, rng)
        got = gpu_answers(layout, b, kind, A, Q, dev)
        exp = ol.search(layout, A, Q, kind, b)
        assert np.array_equal(got, exp), (layout, b, kind, n)
        hit = got != n
        assert np.array_equal(A[got[hit].astype(np.int64)], Q[hit])
        assert not np.isin(Q[~hit], keys).any()


BTREE_FORMS = [  # (TL_BTREE_WIDE, TL_BTREE_IDX64, TL_BTREE_QPG, TL_LEARNED_FUSED)
    ("0", "0", "2", "1"), ("1", "0", "2", "1"), ("1", "0", "2", "0"), ("1", "0", "4", "0"),
    ("1", "1", "2", "1"), ("1", "1", "4", "1")]


@pytest.mark.parametrize("wide,idx64,qpg,fused", BTREE_FORMS)
@pytest.mark.parametrize("b", [16, 8, 2, 32, 64])
@pytest.mark.parametrize("kind", ["u32", "u64", "i64"])
def test_btree_descent_forms(wide, idx64, qpg, fused, b, kind, dev, monkeypatch):
    """The one-key-per-lane group descent and the 16-byte-per-lane wide
    descent (32- and 64-bit node numbers, 2 and 4 queries per group) give the
    oracle's answers, plain and from learned start nodes (two passes, or the
    one-pass learned kernel)."""
    monkeypatch.setenv("TL_BTREE_WIDE", wide)
    monkeypatch.setenv("TL_BTREE_IDX64", idx64)
    monkeypatch.setenv("TL_BTREE_QPG", qpg)
    monkeypatch.setenv("TL_LEARNED_FUSED", fused)
    rng = np.random.default_rng(77 + b)
    for n in [1, 5, 16, 17, 1000, 65535, 100003]:
        keys = make_keys(kind, n, rng)
        A = ol.permute("btree", keys.copy(), b)
        Q = make_queries(kind, keys, 8192, rng)
        got = gpu_answers("btree", b, kind, A, Q, dev)
        assert np.array_equal(got, ol.search("btree", A, Q, kind, b)), (b, kind, n)
        if n < 2:   # the model needs two keys
            continue
        K = to_dev(keys, kind, dev)
        model, _ = tl_amd.learned_fit(K, 64, kind=kind)
        Ad, Qd = to_dev(A, kind, dev), to_dev(Q, kind, dev)
        lrn = torch.empty(Q.size, dtype=torch.int64, device=dev)
        tl_amd.query_btree_learned(Ad, b, model, Qd, lrn, None, kind=kind)
        torch.cuda.synchronize()
        assert np.array_equal(lrn.cpu().numpy().view(np.uint64), got), (b, kind, n)


@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_btree_query_unaligned_view(kind, dev):
    """An array that starts 8 (or 4) bytes past a 16-byte boundary is served
    by the group descent; answers are slots of the view."""
    rng = np.random.default_rng(9)
    n, b = 100003, 16
    keys = make_keys(kind, n, rng)
    A = ol.permute("btree", keys.copy(), b)
    Q = make_queries(kind, keys, 8192, rng)
    st_np, st_t = STORE[kind]
    big = torch.zeros(n + 1, dtype=st_t, device=dev)
    big[1:] = torch.from_numpy(A.view(st_np).copy()).to(dev)
    view = big[1:]
    assert view.data_ptr() % 16 != 0
    ans = torch.empty(Q.size, dtype=torch.int64, device=dev)
    tl_amd.query_btree(view, b, to_dev(Q, kind, dev), ans, kind=kind)
    torch.cuda.synchronize()
    assert np.array_equal(ans.cpu().numpy().view(np.uint64), ol.search("btree", A, Q, kind, b))


@pytest.mark.parametrize("wide", ["0", "1"])
@pytest.mark.parametrize("b", [16, 8])
def test_btree_query_duplicate_keys(wide, b, dev, monkeypatch):
    """Runs of equal keys (inside and across nodes): the answer is the
    reference's first key not smaller than q on the descent path."""
    monkeypatch.setenv("TL_BTREE_WIDE", wide)
    rng = np.random.default_rng(31 + b)
    for n in [17, 1000, 100003]:
        keys = np.sort(rng.integers(0, max(2, n // 7), size=n)).astype(np.uint64)
        A = ol.permute("btree", keys.copy(), b)
        Q = np.concatenate([keys[rng.integers(0, n, 4000)],
                            rng.integers(0, n // 7 + 3, 1000).astype(np.uint64)])
        got = gpu_answers("btree", b, "u64", A, Q, dev)
        assert np.array_equal(got, ol.search("btree", A, Q, "u64", b)), (b, n)


@pytest.mark.parametrize("top", ["0", "1"])
@pytest.mark.parametrize("qpl", ["1", "2", "4"])
@pytest.mark.parametrize("layout", ["mveb", "veb"])
@pytest.mark.parametrize("kind", ["u32", "i64"])
def test_veb_query_shapes(top, qpl, layout, kind, dev, monkeypatch):
    """The vEB query with and without its BFS-staged top levels and with 1,
    2 or 4 queries per lane gives the oracle's answers (perfect and
    non-perfect sizes, every table-chain shape in the list)"""
    monkeypatch.setenv("TL_VEB_TOP", top)
    monkeypatch.setenv("TL_VEB_QPL", qpl)
    rng = np.random.default_rng(21)
    for n in [1, 2, 3, 7, 100, 1023, 1024, 1025, 4095, 65535, 100003, 1048575, 1000003]:
        keys = make_keys(kind, n, rng)
        A = ol.permute(layout, keys.copy(), 0)
        Q = make_queries(kind, keys, 8192, rng)
        got = gpu_answers(layout, 0, kind, A, Q, dev)
        assert np.array_equal(got, ol.search(layout, A, Q, kind, 0)), (layout, kind, n)


@pytest.mark.parametrize("kind", ["u32", "u64", "i32", "i64"])
def test_binary_search_matches_oracle(kind, dev):
    """Sorted array: hit -> index, miss -> predecessor's index, below the
    smallest key -> n; sizes around the 2047-key LDS top and its levels."""
    rng = np.random.default_rng(4321)
    for n in [0] + SIZES + [2047, 2048, 2049, 4094]:
        keys = make_keys(kind, n, rng) if n else np.zeros(0, dtype=NP[kind])
        if n:
            Q = make_queries(kind, keys, 4096, rng)
            info = np.iinfo(NP[kind])
            Q = np.concatenate([Q, np.array([info.min, info.max], dtype=NP[kind]), keys[:1] - 1
                                if keys[0] > info.min else keys[:1]]).astype(NP[kind])
        else:
            Q = rng.integers(0, 1000, size=100).astype(NP[kind])
        A = to_dev(keys, kind, dev) if n else torch.zeros(1, dtype=STORE[kind][1], device=dev)
        Qd = to_dev(Q, kind, dev)
        ans = torch.full((Q.size,), -7, dtype=torch.int64, device=dev)
        tl_amd.query_binary(A[:n], Qd, ans, kind=kind)
        torch.cuda.synchronize()
        got = ans.cpu().numpy().view(np.uint64)
        exp = ol.search("binary", keys, Q, kind)
        assert np.array_equal(got, exp), (kind, n)


@pytest.mark.parametrize("classic", [False, True])
def test_tableset_device_matches_oracle(classic, dev):
    for n in [1, 2, 3, 6, 7, 100, 127, 1000, 65535, 65536, 100000, 300_000_000, (1 << 29) - 1,
              3 * 10 ** 8, (1 << 30)]:
        tables, idx, tidx, meta = tl_amd.veb_tableset_device(n, device=dev, classic=classic)
        d, nt, ts = (int(v) for v in meta[:3].cpu())
        od, ont, ots, otab, oidx, otidx = ol.tableset(n, classic=classic)
        assert (d, nt, ts) == (od, ont, ots), n
        raw = tables.cpu().numpy().view(np.uint64).reshape(-1, 3)[:ts]
        got = [(int(r[0]), int(r[1]), int(r[2]) & 0xFFFFFFFF) for r in raw]
        assert got == otab, n
        assert [int(v) for v in idx[:nt].cpu()] == oidx
        assert [int(v) for v in tidx[:nt].cpu()] == otidx


def test_config3_nonperfect_tableset_shape(dev):
    """n = 3e8 mvEB: 4 tables of depths 29/16/7/3 (table_size 55)."""
    tables, idx, tidx, meta = tl_amd.veb_tableset_device(3 * 10 ** 8, device=dev)
    d, nt, ts = (int(v) for v in meta[:3].cpu())
    assert (d, nt, ts) == (29, 4, 55)
    n = 3 * 10 ** 8
    assert [int(v) for v in idx[:4].cpu()] == [n - 38114, n - 119, n - 7, n]
