"""The step before the layouts and the one after: raw keys sorted and made
distinct on the device (tl_sort_dedupe, checked against numpy.unique, which
is exactly "sorted distinct keys" for every key type), and layout files
(tl_layout_save / tl_layout_load: header, vEB table set, checksum)."""
import os
import zlib

import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu

NP = {"u32": np.uint32, "u64": np.uint64, "i32": np.int32, "i64": np.int64}
TDT = {"u32": torch.int32, "u64": torch.int64, "i32": torch.int32, "i64": torch.int64}
STORE = {"u32": np.int32, "u64": np.int64, "i32": np.int32, "i64": np.int64}


def _raw(kind, n, rng, spread):
    info = np.iinfo(NP[kind])
    if spread == "full":
        x = rng.integers(info.min, info.max, size=n, dtype=NP[kind], endpoint=True)
    elif spread == "narrow":       # many repeats, top bytes constant
        x = (rng.integers(0, max(1, n // 3), size=n) + (info.max // 2 if info.min == 0 else -7))
        x = x.astype(NP[kind])
    elif spread == "equal":
        x = np.full(n, 12345, dtype=NP[kind])
    elif spread == "sorted":
        x = np.sort(rng.integers(info.min, info.max, size=n, dtype=NP[kind], endpoint=True))
    else:                          # reverse sorted
        x = np.sort(rng.integers(info.min, info.max, size=n, dtype=NP[kind], endpoint=True))[::-1]
    return np.ascontiguousarray(x)


@pytest.mark.parametrize("kind", ["u32", "u64", "i32", "i64"])
@pytest.mark.parametrize("spread", ["full", "narrow", "equal", "sorted", "reverse"])
def test_sort_dedupe_equals_numpy_unique(kind, spread, dev):
    rng = np.random.default_rng(zlib.crc32(f"{kind}-{spread}".encode()))
    for n in [0, 1, 2, 3, 100, 8191, 8192, 8193, 65537, 1_000_003]:
        x = _raw(kind, n, rng, spread)
        A = torch.from_numpy(x.view(STORE[kind]).copy()).to(dev)
        m, _ = tl_amd.sort_dedupe(A, kind=kind)
        exp = np.unique(x)
        assert m == exp.size, (kind, spread, n, m, exp.size)
        got = A[:m].cpu().numpy().view(NP[kind])
        assert np.array_equal(got, exp), (kind, spread, n)


@pytest.mark.parametrize("kind", ["u64", "i32"])
def test_sort_dedupe_large(kind, dev):
    n = 40_000_000
    print(f"[test] sort_dedupe {kind} n={n}", flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    if kind == "u64":
        A = torch.randint(-(1 << 62), 1 << 62, (n,), generator=g, device=dev, dtype=torch.int64)
        k = A[::11].numel()
        A[: 7 * k : 7] = A[::11].clone()       # repeats
    else:
        A = torch.randint(-(1 << 30), 1 << 30, (n,), generator=g, device=dev, dtype=torch.int32)
    host = A.cpu().numpy().view(NP[kind])
    m, _ = tl_amd.sort_dedupe(A, kind=kind)
    exp = np.unique(host)
    assert m == exp.size
    assert np.array_equal(A[:m].cpu().numpy().view(NP[kind]), exp)


def _distinct_sorted(kind, n, rng):
    raw = _raw(kind, int(n * 1.3), rng, "full")
    return np.unique(raw)[:n]


@pytest.mark.parametrize("layout,b", [("btree", 16), ("bst", 1), ("mveb", 0), ("veb", 0)])
@pytest.mark.parametrize("kind", ["u64", "i32"])
def test_raw_keys_to_layout_file_and_back(layout, b, kind, dev, tmp_path):
    """raw keys with repeats -> sort_dedupe -> permute -> save -> load into a
    fresh buffer -> queries on the loaded layout equal the oracle's"""
    rng = np.random.default_rng(21)
    raw = _raw(kind, 3_000_017, rng, "full")
    k = raw[1::5].size
    raw[0:5 * k:5] = raw[1::5]                  # repeats
    A = torch.from_numpy(raw.view(STORE[kind]).copy()).to(dev)
    m, _ = tl_amd.sort_dedupe(A, kind=kind)
    keys = np.unique(raw)
    assert m == keys.size
    L = A[:m].clone()
    tl_amd.permute(layout, L, b=b, kind=kind)
    path = os.path.join(tmp_path, f"{layout}.tl")
    tl_amd.layout_save(path, L, layout, b=b, kind=kind)
    info = tl_amd.layout_read_info(path)
    assert info["n"] == m and info["layout"] == tl_amd.LAYOUT[layout]
    assert info["dtype"] == tl_amd.DTYPE[kind]
    B = torch.full((m + 10,), -1, dtype=TDT[kind], device=dev)
    tables = torch.zeros(64 * 64 * 3, dtype=torch.int64, device=dev)
    idx = torch.zeros(64, dtype=torch.int64, device=dev)
    tidx = torch.zeros(64, dtype=torch.int32, device=dev)
    got = tl_amd.layout_load(path, B, tables, idx, tidx)
    assert got == info
    assert torch.equal(B[:m], L)
    exp_layout = ol.permute(layout, keys.copy(), b)
    assert np.array_equal(B[:m].cpu().numpy().view(NP[kind]), exp_layout)
    q = np.concatenate([keys[rng.integers(0, m, 20000)], _raw(kind, 5000, rng, "full")])
    Q = torch.from_numpy(q.view(STORE[kind]).copy()).to(dev)
    ans = torch.empty(q.size, dtype=torch.int64, device=dev)
    Bm = B[:m]
    if layout == "btree":
        tl_amd.query_btree(Bm, b, Q, ans, kind=kind)
    elif layout == "bst":
        tl_amd.query_bst(Bm, Q, ans, kind=kind)
    else:
        # the table set comes from the file, and is the oracle's
        d, nt, ts, otables, oidx, otidx = ol.tableset(m, classic=(layout == "veb"))
        assert (info["d"], info["num_tables"], info["table_size"]) == (d, nt, ts)
        t = tables[: 3 * ts].cpu().numpy().reshape(ts, 3)
        assert [(int(a), int(r), int(x) & 0xFFFFFFFF) for a, r, x in t] == otables
        assert idx[:nt].cpu().tolist() == oidx and tidx[:nt].cpu().tolist() == otidx
        fn = tl_amd.query_mveb_nonperfect if layout == "mveb" else tl_amd.query_veb_nonperfect
        fn(Bm, tables, idx, tidx, info["d"], info["num_tables"], Q, ans, kind=kind)
    exp = ol.search(layout, exp_layout, q, kind, b)
    assert np.array_equal(ans.cpu().numpy().view(np.uint64), exp), layout


def test_layout_file_errors(dev, tmp_path):
    n = 100_003
    A = torch.arange(n, dtype=torch.int64, device=dev)
    tl_amd.permute_btree(A, 16, kind="u64")
    path = os.path.join(tmp_path, "a.tl")
    tl_amd.layout_save(path, A, "btree", b=16, kind="u64")
    # too small a buffer: rejected, nothing written
    small = torch.zeros(n - 1, dtype=torch.int64, device=dev)
    with pytest.raises(tl_amd.TLError) as e:
        tl_amd.layout_load(path, small)
    assert e.value.status == tl_amd.TL_EINVAL
    # one flipped byte in the keys: the checksum catches it
    info = tl_amd.layout_read_info(path)
    with open(path, "r+b") as f:
        f.seek(os.path.getsize(path) - 8 * 1000)
        byte = f.read(1)
        f.seek(-1, 1)
        f.write(bytes([byte[0] ^ 0x40]))
    B = torch.zeros(n, dtype=torch.int64, device=dev)
    with pytest.raises(tl_amd.TLError) as e:
        tl_amd.layout_load(path, B)
    assert e.value.status == tl_amd.TL_EIO
    # two keys swapped (same multiset): the position-sensitive checksum catches it
    tl_amd.layout_save(path, A, "btree", b=16, kind="u64")
    raw = bytearray(open(path, "rb").read())
    off = len(raw) - 8 * n
    raw[off:off + 8], raw[off + 8:off + 16] = raw[off + 8:off + 16], raw[off:off + 8]
    open(path, "wb").write(bytes(raw))
    with pytest.raises(tl_amd.TLError) as e:
        tl_amd.layout_load(path, B)
    assert e.value.status == tl_amd.TL_EIO
    assert info["checksum"] != 0
