"""The uneven batch at its benchmark size, through tl_batch_run replayed from
its HIP graph: 4096 arrays, sizes floor(2^U) with U ~ U[8, 22] (seed 5),
int32 / uint64 keys 50/50, BST / B-tree b=8 / mvEB round robin, 4 * size
uniform-hit queries per array (about 1.76e9 keys and 7.0e9 queries; the
generator is bench.py's).

Checks, over every array of the batch:
  * the layout holds exactly the array's keys (sorted layout == sorted input);
  * every query is answered with a slot that holds its key (all are hits);
  * a second graph replay, after the arrays are restored, gives the same
    layout and answers bit for bit;
  * the per-array inverse permutation gives the sorted input back.
Over a sample (every 16th array, plus the largest array of each layout and
key width): the layout equals the oracle's slot for slot, and the first
65,536 answers equal the oracle's search of that layout.
"""
import gc

import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd
from bench import config5_arrays

pytestmark = pytest.mark.gpu


def _host(t, kind):
    a = t.cpu().numpy()
    return a.view(np.uint64) if kind == "u64" else a


def test_config5_batch_full_size(dev):
    # about 190 GB of tensors: hand the memory back to the device afterwards,
    # so that later tests (and the library's own allocations) find it free
    try:
        _check_config5(dev)
    finally:
        gc.collect()
        torch.cuda.empty_cache()


def _check_config5(dev):
    print("[test] config-5 batch: generating 4096 arrays", flush=True)
    arrays, pristine, _, _ = config5_arrays(dev, 4096)
    total = sum(p.numel() for p in pristine)
    assert 1.5e9 < total < 2.0e9, total
    B = tl_amd.Batch(arrays)
    print(f"[test] config-5 batch: {total} keys, first graph replay", flush=True)
    B.run(graph=True)
    torch.cuda.synchronize()

    # whole batch: multiset of keys, hit answers
    for i, (a, p) in enumerate(zip(arrays, pristine)):
        A, n = a["A"], p.numel()
        assert torch.equal(torch.sort(A).values, p), (i, a["layout"], a["kind"], n)
        ans = a["ans"]
        assert bool((ans >= 0).all()) and bool((ans < n).all()), (i, a["layout"], n)
        assert torch.equal(A[ans], a["Q"]), (i, a["layout"], a["kind"], n)
    print("[test] config-5 batch: every array a permutation, every answer holds its key",
          flush=True)

    # sample: slot-for-slot against the oracle
    sizes = [p.numel() for p in pristine]
    sample = set(range(0, 4096, 16))
    for lay in ("bst", "btree", "mveb"):
        for kind in ("i32", "u64"):
            idx = [i for i, a in enumerate(arrays) if a["layout"] == lay and a["kind"] == kind]
            sample.add(max(idx, key=lambda i: sizes[i]))
    for i in sorted(sample):
        a = arrays[i]
        keys = _host(pristine[i], a["kind"])
        exp = ol.permute(a["layout"], keys.copy(), 8)
        got = _host(a["A"], a["kind"])
        assert np.array_equal(got, exp), (i, a["layout"], a["kind"], keys.size)
        q = _host(a["Q"][:65536], a["kind"])
        exp_ans = ol.search(a["layout"], exp, q, a["kind"], 8)
        assert np.array_equal(a["ans"][:65536].cpu().numpy().view(np.uint64), exp_ans), \
            (i, a["layout"], a["kind"], keys.size)
    print(f"[test] config-5 batch: {len(sample)} sampled arrays equal the oracle", flush=True)

    # a second replay from restored input gives the same bits
    first = [(a["A"].clone(), a["ans"].clone()) for a in arrays]
    for a, p in zip(arrays, pristine):
        a["A"].copy_(p)
        a["ans"].fill_(-1)
    torch.cuda.synchronize()
    B.run(graph=True)
    torch.cuda.synchronize()
    for i, (a, (A1, ans1)) in enumerate(zip(arrays, first)):
        assert torch.equal(a["A"], A1) and torch.equal(a["ans"], ans1), (i, a["layout"])
    del first
    B.close()
    print("[test] config-5 batch: second replay identical", flush=True)

    # round trip through the inverse permutations
    for i, (a, p) in enumerate(zip(arrays, pristine)):
        kind = "i32" if a["kind"] == "i32" else "u64"
        tl_amd.unpermute(a["layout"], a["A"], b=8, kind=kind, timed=False)
    torch.cuda.synchronize()
    tl_amd.sync()
    for i, (a, p) in enumerate(zip(arrays, pristine)):
        assert torch.equal(a["A"], p), (i, a["layout"], a["kind"], p.numel())
    print("[test] config-5 batch: inverse permutations restore every array", flush=True)
