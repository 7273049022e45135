"""The allocation-free B-tree permutation (tl_permute_btree_noalloc), called
through the C ABI, against tl_permute_btree on the same input, slot for slot.

Each case is two calls in a row on one device, the second on a smaller
array, so the second call runs on the static scratch the first one left
behind (its flags must be cleared again, its tables replaced).  The large
first sizes are ones whose plan differs from tl_permute_btree's: their
stashes are capped by the static array (a capped one-pass rotation, or
hierarchical gathers with unit merges and three-reversal rotations instead
of row slides), so the check compares two different plans of one layout.
Small sizes are also compared with the oracle in full; every size is
sample-checked against the oracle's closed-form slot of a rank.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu

TDT = {"u32": torch.int32, "u64": torch.int64}


def _keys(n, kind, dev, seed):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    gaps = torch.randint(1, 4 if kind == "u32" else 64, (n,), generator=g, device=dev,
                         dtype=torch.int64)
    k = torch.cumsum(gaps, 0)
    return k.to(TDT[kind])


def _noalloc(A, b, kind):
    L = tl_amd.lib()
    st = L.tl_permute_btree_noalloc(ctypes.c_void_p(A.data_ptr()), A.numel(), b,
                                    tl_amd.DTYPE[kind],
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == tl_amd.TL_OK, L.tl_last_error().decode()
    tl_amd.sync()   # no in-kernel wait gave up


def _sample_check(A, keys, n, b, rng, samples=100_000):
    o = ol.oracle()
    ranks = rng.integers(0, n, size=samples, dtype=np.int64)
    pos = np.array([o.or_btree_pos(int(x), n, b) for x in ranks], dtype=np.int64)
    got = A[torch.from_numpy(pos).to(A.device)].cpu()
    exp = keys[torch.from_numpy(ranks).to(A.device)].cpu()
    assert torch.equal(got, exp)


# (first n, second smaller n, b, key type)
CASES = [
    (300_000_001, 5_000_011, 16, "u64"),    # first: hierarchical gathers, reversals
    (120_000_007, 777_777, 16, "u64"),      # first: capped one-pass rotation
    (50_000_017, 4_782_970, 8, "u32"),
    (2_000_003, 100_003, 2, "u64"),
]


@pytest.mark.parametrize("n1,n2,b,kind", CASES)
def test_noalloc_equals_permute_btree(n1, n2, b, kind, dev):
    rng = np.random.default_rng(7)
    for call, n in enumerate((n1, n2)):
        print(f"[test] noalloc b={b} {kind} n={n} (call {call + 1})", flush=True)
        keys = _keys(n, kind, dev, seed=n)
        ref = keys.clone()
        tl_amd.permute_btree(ref, b, kind=kind)
        A = keys.clone()
        _noalloc(A, b, kind)
        torch.cuda.synchronize()
        assert torch.equal(A, ref), (n, b, kind)
        if n <= 5_000_011:
            exp = keys.cpu().numpy()[ol.layout_ranks("btree", n, b).astype(np.int64)]
            assert np.array_equal(A.cpu().numpy(), exp), (n, b, kind)
        _sample_check(A, keys, n, b, rng)
        del A, ref, keys
        torch.cuda.empty_cache()


def test_noalloc_rejects_sizes_beyond_its_limit(dev):
    A = torch.zeros(16, dtype=torch.int64, device=dev)
    L = tl_amd.lib()
    # rejected before any launch: the array is never touched
    st = L.tl_permute_btree_noalloc(ctypes.c_void_p(A.data_ptr()), tl_amd.TL_NOALLOC_MAX_KEYS + 1,
                                    16, tl_amd.DTYPE["u64"], None)
    assert st == tl_amd.TL_EINVAL
    assert "TL_NOALLOC_MAX_KEYS" in L.tl_last_error().decode()
    assert torch.count_nonzero(A).item() == 0
