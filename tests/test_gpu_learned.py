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


@pytest.mark.parametrize("kind", ["u64", "i32"])
@pytest.mark.parametrize("leaves", [1 << 8, 1 << 12, 1 << 16])
def test_learned_error_bound(kind, leaves, dev):
    """Every key of the fitted array lies inside its own prediction window:
    |pred - rank| <= err of its leaf, and err <= the model's max_err."""
    rng = np.random.default_rng(5 + leaves)
    n = 1 << 20
    keys = _keys(kind, n, rng)
    st = np.int32 if keys.dtype.itemsize == 4 else np.int64
    K = torch.from_numpy(keys.view(st).copy()).to(dev)
    model, _ = tl_amd.learned_fit(K, leaves, kind=kind)
    pred, err = tl_amd.learned_predict(model, K, kind=kind)
    torch.cuda.synchronize()
    rank = torch.arange(n, device=dev, dtype=torch.int64)
    assert bool((err >= 0).all()), "a fitted key maps to an empty leaf"
    gap = (pred - rank).abs()
    worst = int((gap - err).max().item())
    assert worst <= 0, f"a key's rank lies {worst} outside its leaf's error window"
    max_err = tl_amd.learned_max_err(model)
    assert int(err.max().item()) == max_err
    assert 0 <= max_err < n


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("b", [16, 8])
def test_learned_equals_plain_duplicate_keys(b, fused, dev, monkeypatch):
    """Runs of equal keys that cross node boundaries: every duplicate shares
    one leaf and one prediction, so the learned start node still lies on the
    plain descent's path and the answers agree bit for bit."""
    monkeypatch.setenv("TL_LEARNED_FUSED", fused)
    rng = np.random.default_rng(91 + b)
    for n in (5000, 200_003, 1 << 20):
        runs = rng.integers(1, 3 * b, size=n)          # run lengths up to 3 nodes
        vals = np.cumsum(rng.integers(1, 40, size=n))
        keys = np.repeat(vals, runs)[:n].astype(np.int64)
        K = torch.from_numpy(keys.copy()).to(dev)
        model, _ = tl_amd.learned_fit(K, 1 << 10, kind="u64")
        A = K.clone()
        tl_amd.permute_btree(A, b, kind="u64")
        hits = keys[rng.integers(0, n, size=30000)]
        miss = rng.integers(int(keys[0]) - 3, int(keys[-1]) + 3, size=10000)
        q = np.concatenate([hits, miss]).astype(np.int64)
        Q = torch.from_numpy(q).to(dev)
        plain = torch.empty(q.size, dtype=torch.int64, device=dev)
        lrn = torch.empty_like(plain)
        tl_amd.query_btree(A, b, Q, plain, kind="u64")
        tl_amd.query_btree_learned(A, b, model, Q, lrn, None, kind="u64")
        torch.cuda.synchronize()
        assert torch.equal(plain, lrn), (b, n)
        exp = ol.search("btree", A.cpu().numpy().view(np.uint64), q.view(np.uint64), "u64", b)
        assert np.array_equal(plain.cpu().numpy().view(np.uint64), exp), (b, n)
