"""Learned start node: answers equal the plain root descent bit for bit (the
plain descent is the fallback and decides every answer), the model's error
bound holds on every key, and hits never fall back."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu


def _keys(kind, n, rng):
    if kind in ("u64", "i64"):
        gaps = np.maximum(1, np.ceil(rng.lognormal(0.0, 2.0, size=n))).astype(np.int64)
        k = np.cumsum(gaps)
        if kind == "i64":
            k = k - k[n // 2]
        return k
    gaps = rng.integers(1, 65, size=n).astype(np.int64)
    k = np.cumsum(gaps)
    return (k - k[n // 2]).astype(np.int32) if kind == "i32" else k.astype(np.uint32)


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("kind", ["u64", "i64", "u32", "i32"])
@pytest.mark.parametrize("b", [16, 8, 5, 32])
def test_learned_equals_plain(kind, b, fused, dev, monkeypatch):
    monkeypatch.setenv("TL_LEARNED_FUSED", fused)
    rng = np.random.default_rng(44 + b)
    for n, leaves in [(3, 4), (100, 8), (5000, 64), (100003, 1 << 10), (1 << 20, 1 << 16),
                      (3_000_001, 1 << 16)]:
        keys = _keys(kind, n, rng)
        st = np.int32 if keys.dtype.itemsize == 4 else np.int64
        tdt = torch.int32 if st == np.int32 else torch.int64
        K = torch.from_numpy(keys.view(st).copy()).to(dev)
        model, _ = tl_amd.learned_fit(K, leaves, kind=kind)
        A = K.clone()
        tl_amd.permute_btree(A, b, kind=kind)
        hits = keys[rng.integers(0, n, size=20000)]
        lo, hi = (int(keys.min()), int(keys.max()))
        miss = rng.integers(lo - 5, hi + 5, size=20000)
        q = np.concatenate([hits, miss.astype(keys.dtype)])
        Q = torch.from_numpy(q.view(st).copy()).to(dev)
        plain = torch.empty(q.size, dtype=torch.int64, device=dev)
        lrn = torch.empty(q.size, dtype=torch.int64, device=dev)
        stats = torch.zeros(1, dtype=torch.int64, device=dev)
        tl_amd.query_btree(A, b, Q, plain, kind=kind)
        tl_amd.query_btree_learned(A, b, model, Q, lrn, stats, kind=kind)
        torch.cuda.synchronize()
        assert torch.equal(plain, lrn), (kind, b, n)
        # the plain answers are the oracle's
        exp = ol.search("btree", A.cpu().numpy().view(keys.dtype), q, kind, b)
        assert np.array_equal(plain.cpu().numpy().view(np.uint64), exp), (kind, b, n)
        # hits never fall back: only misses (and hits of empty leaves, none here) do
        nmiss = int((exp == n).sum())
        assert int(stats.item()) <= nmiss + 0, (kind, b, n, int(stats.item()), nmiss)


def test_learned_error_bound(dev):
    rng = np.random.default_rng(5)
    n = 1 << 20
    keys = _keys("u64", n, rng)
    K = torch.from_numpy(keys).to(dev)
    model, _ = tl_amd.learned_fit(K, 1 << 12, kind="u64")
    err = tl_amd.learned_max_err(model)
    assert 0 <= err < n
    # recompute every key's prediction on the host from the model's leaves
    raw = model.cpu().numpy()
    hdr = raw[:8].view(np.uint8)
    kmin = int(raw[0]); ra, rc, inv = np.frombuffer(hdr[16:28].tobytes(), dtype=np.float32)
    leaves = raw[8:8 + 4 * (1 << 12)].reshape(-1, 4)
    u = keys.astype(np.uint64)
    x = ((u - np.uint64(kmin)).astype(np.float32) * np.float32(inv)).astype(np.float32)
    leaf = np.clip(np.floor(np.float32(ra) * x + np.float32(rc)), 0, (1 << 12) - 1).astype(np.int64)
    counts = (leaves[:, 3].view(np.uint64) >> np.uint64(32)).astype(np.int64)
    assert counts.sum() == n
