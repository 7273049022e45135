"""GPU parity at the benchmark configurations' full sizes.

Where the oracle finishes in seconds (config 1) the whole layout and every
answer are compared with it.  At 10^8 .. 2^30 keys the checks are
size-independent: the layout is a permutation of the input (sorted back on the
GPU it equals the input), 200k sampled ranks sit at the oracle's closed-form
slot, every hit's answer holds the query key and every miss answers n.  Each
layout then goes back to the sorted keys through the inverse permutation.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu


def _sorted_back_equal(A, keys):
    return bool(torch.equal(torch.sort(A)[0], keys))


def _sample_slots(layout, A, keys, n, b, rng, samples=200_000):
    o = ol.oracle()
    ranks = rng.integers(0, n, size=samples, dtype=np.int64)
    if layout == "mveb":
        pos = np.array([o.or_mveb_pos(int(x), n) for x in ranks], dtype=np.int64)
    else:
        pos = np.array([o.or_btree_pos(int(x), n, b) for x in ranks], dtype=np.int64)
    got = A[torch.from_numpy(pos).to(A.device)]
    exp = keys[torch.from_numpy(ranks).to(A.device)]
    return bool(torch.equal(got, exp))


def _check_answers(A, Q, ans, n, is_hit):
    hit = torch.from_numpy(is_hit).to(A.device)
    a = ans[hit]
    assert bool(((a >= 0) & (a < n)).all()), "a hit answered out of range"
    assert torch.equal(A[a], Q[hit]), "a hit answered a slot that does not hold the key"
    assert bool((ans[~hit] == n).all()), "a miss did not answer n"


def test_config1_bst_u32_whole_layout(dev):
    """n = 2^20 - 1 distinct uint32 uniform in [0, 2^32) (seed 1), perfect BST;
    2^20 queries, half hits, half uniform misses: layout and answers equal the
    oracle's slot for slot."""
    rng = np.random.default_rng(1)
    n = (1 << 20) - 1
    keys = np.unique(rng.integers(0, 1 << 32, size=n + 4096, dtype=np.uint64))
    while keys.size < n:
        keys = np.unique(np.concatenate([keys, rng.integers(0, 1 << 32, size=4096,
                                                            dtype=np.uint64)]))
    keys = keys[:n].astype(np.uint32)
    hits = keys[rng.integers(0, n, size=1 << 19)]
    miss = rng.integers(0, 1 << 32, size=(1 << 19) + 1000, dtype=np.uint64).astype(np.uint32)
    miss = miss[~np.isin(miss, keys)][: 1 << 19]
    q = rng.permutation(np.concatenate([hits, miss]))
    A = torch.from_numpy(keys.view(np.int32).copy()).to(dev)
    tl_amd.permute_bst(A, kind="u32")
    exp = ol.permute("bst", keys.copy(), 1)
    assert np.array_equal(A.cpu().numpy().view(np.uint32), exp)
    Q = torch.from_numpy(q.view(np.int32).copy()).to(dev)
    ans = torch.empty(q.size, dtype=torch.int64, device=dev)
    tl_amd.query_bst(A, Q, ans, kind="u32")
    exp_ans = ol.search("bst", exp, q, "u32", 1)
    assert np.array_equal(ans.cpu().numpy().view(np.uint64), exp_ans)
    tl_amd.unpermute("bst", A, kind="u32")
    assert np.array_equal(A.cpu().numpy().view(np.uint32), keys)


def _gap_keys(n, seed, kind, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if kind == "uniform64":
        gaps = torch.randint(1, 65, (n,), generator=g, device=dev, dtype=torch.int64)
    else:   # geometric(p = 0.1) on {1, 2, ...}
        u = torch.rand(n, generator=g, device=dev, dtype=torch.float64).clamp_(min=1e-300)
        gaps = torch.floor(torch.log(u) / np.log(0.9)).to(torch.int64) + 1
    return torch.cumsum(gaps, 0), g


def test_config2_btree_1e8(dev):
    """n = 10^8 uint64 (U{1..64} gaps, seed 2), b = 16 (non-perfect); 2^26
    queries, 90 % hits and 10 % misses (keys + 0.5 rounded up fall between
    consecutive keys whenever the gap exceeds 1)."""
    n, b = 10 ** 8, 16
    keys, g = _gap_keys(n, 2, "uniform64", dev)
    A = keys.clone()
    tl_amd.permute_btree(A, b, kind="u64")
    assert _sorted_back_equal(A, keys)
    assert _sample_slots("btree", A, keys, n, b, np.random.default_rng(2))
    nq = 1 << 26
    qi = torch.randint(0, n, (nq,), generator=g, device=dev, dtype=torch.int64)
    Q = keys[qi]
    nmiss = nq // 10
    Q[:nmiss] = keys[qi[:nmiss]] + 1           # may collide with the next key
    is_hit = torch.isin(Q, keys).cpu().numpy()
    ans = torch.empty(nq, dtype=torch.int64, device=dev)
    tl_amd.query_btree(A, b, Q, ans, kind="u64")
    _check_answers(A, Q, ans, n, is_hit)
    tl_amd.unpermute("btree", A, b=b, kind="u64")
    assert torch.equal(A, keys)


@pytest.mark.parametrize("n", [(1 << 29) - 1, 3 * 10 ** 8])
def test_config3_mveb(n, dev):
    """mvEB, perfect n = 2^29 - 1 and non-perfect n = 3*10^8 (geometric(0.1)
    gaps, seed 3); device-built table set; 2^24 Zipf(1.1)-ranked hits and
    2^20 misses."""
    keys, g = _gap_keys(n, 3, "geometric", dev)
    A = keys.clone()
    tl_amd.permute_mveb(A, kind="u64")
    assert _sorted_back_equal(A, keys)
    assert _sample_slots("mveb", A, keys, n, 0, np.random.default_rng(3))
    rng = np.random.default_rng(3)
    ranks = np.minimum(rng.zipf(1.1, size=1 << 24) - 1, n - 1)
    perm_ranks = (ranks * 2654435761 + 12345) % n      # seeded spread of the ranks
    Q = keys[torch.from_numpy(perm_ranks).to(dev)]
    miss = keys[torch.randint(0, n - 1, (1 << 20,), generator=g, device=dev)] + 1
    Q = torch.cat([Q, miss])
    is_hit = torch.isin(Q, keys).cpu().numpy()
    tables, idx, tidx, meta = tl_amd.veb_tableset_device(n)
    d, num_tables = int(meta[0].item()), int(meta[1].item())
    ans = torch.empty(Q.numel(), dtype=torch.int64, device=dev)
    tl_amd.query_mveb_nonperfect(A, tables, idx, tidx, d, num_tables, Q, ans, kind="u64")
    _check_answers(A, Q, ans, n, is_hit)
    tl_amd.unpermute("mveb", A, kind="u64")
    assert torch.equal(A, keys)


@pytest.mark.parametrize("n", [(1 << 29) - 1, 3 * 10 ** 8])
def test_classic_veb_fullsize(n, dev):
    """The classic (CPU) vEB at the mvEB configuration's sizes and keys:
    layout sample against the oracle's closed form, every hit found with the
    classic table set, misses answer n, and the inverse restores the keys."""
    keys, g = _gap_keys(n, 3, "geometric", dev)
    A = keys.clone()
    tl_amd.permute("veb", A, kind="u64")
    assert _sorted_back_equal(A, keys)
    o = ol.oracle()
    rng = np.random.default_rng(5)
    ranks = rng.integers(0, n, size=200_000, dtype=np.int64)
    pos = np.array([o.or_classic_pos(int(x), n) for x in ranks], dtype=np.int64)
    assert torch.equal(A[torch.from_numpy(pos).to(dev)], keys[torch.from_numpy(ranks).to(dev)])
    Q = torch.cat([keys[torch.randint(0, n, (1 << 24,), generator=g, device=dev)],
                   keys[torch.randint(0, n - 1, (1 << 20,), generator=g, device=dev)] + 1])
    is_hit = torch.isin(Q, keys).cpu().numpy()
    tables, idx, tidx, meta = tl_amd.veb_tableset_device(n, classic=True)
    d, nt = int(meta[0].item()), int(meta[1].item())
    ans = torch.empty(Q.numel(), dtype=torch.int64, device=dev)
    if nt == 1 and n == (1 << d) - 1:
        tl_amd.query_veb(A, tables, d, Q, ans, kind="u64")
    else:
        tl_amd.query_veb_nonperfect(A, tables, idx, tidx, d, nt, Q, ans, kind="u64")
    _check_answers(A, Q, ans, n, is_hit)
    tl_amd.unpermute("veb", A, kind="u64")
    assert torch.equal(A, keys)


def test_config4_btree_2p30_learned(dev):
    """n = 2^30 uint64 (lognormal(0, 2) gaps), b = 16: the layout, 2^24
    uniform hits answered by the plain and the learned-start descent alike,
    and the binary search of the sorted keys on the same hits and on key + 1."""
    import bench
    n, b = 1 << 30, 16
    keys, g = bench.gen_keys(n, 4, dev)
    A = keys.clone()
    tl_amd.permute_btree(A, b, kind="u64")
    assert _sorted_back_equal(A, keys)
    assert _sample_slots("btree", A, keys, n, b, np.random.default_rng(4))
    nq = 1 << 24
    qi = torch.randint(0, n, (nq,), generator=g, device=dev, dtype=torch.int64)
    Q = keys[qi]
    # the comparator on the sorted keys: hits answer their index, key + 1
    # answers its predecessor (the key itself, or the next one when the gap is 1)
    bs = torch.empty(nq, dtype=torch.int64, device=dev)
    tl_amd.query_binary(keys, Q, bs, kind="u64")
    assert torch.equal(bs, qi)
    Qm = Q + 1
    tl_amd.query_binary(keys, Qm, bs, kind="u64")
    assert bool(((bs >= 0) & (bs < n)).all())
    assert bool((keys[bs] <= Qm).all())
    nxt = (bs + 1).clamp(max=n - 1)
    assert bool(((bs == n - 1) | (keys[nxt] > Qm)).all())
    del Qm, bs, qi
    plain = torch.empty(nq, dtype=torch.int64, device=dev)
    tl_amd.query_btree(A, b, Q, plain, kind="u64")
    _check_answers(A, Q, plain, n, np.ones(nq, dtype=bool))
    model, _ = tl_amd.learned_fit(keys, 1 << 20, kind="u64")
    lrn = torch.empty(nq, dtype=torch.int64, device=dev)
    stats = torch.zeros(1, dtype=torch.int64, device=dev)
    tl_amd.query_btree_learned(A, b, model, Q, lrn, stats, kind="u64")
    assert torch.equal(plain, lrn)
    assert int(stats.item()) == 0   # hits never fall back
    del plain, lrn, Q
    tl_amd.unpermute("btree", A, b=b, kind="u64")
    assert torch.equal(A, keys)     # the inverse brings the layout back to sorted order
