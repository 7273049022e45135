"""Host-side checks of the uneven-batch engine (no GPU): the sub-task list of
a mixed batch, replayed one sub-task at a time in ticket order with each
executor's index arithmetic, reproduces the oracle's layouts and answers, and
every wait of the persistent launch targets a smaller ticket."""
import ctypes

import numpy as np
import pytest

import batch_emulator as be
import oracle_lib as ol
import tl_amd

LAYOUTS = ["bst", "btree", "mveb"]
KIND_OF = {0: "u32", 1: "u64", 2: "i32", 3: "i64"}


def synthetic_batch(count, lo, hi, seed, b=8):
    """items with fake, disjoint device addresses, and the host arrays behind them"""
    rng = np.random.default_rng(seed)
    items = (tl_amd.BatchItem * count)()
    arrays, queries, answers, host = {}, {}, {}, []
    for i in range(count):
        n = int(2 ** rng.uniform(lo, hi))
        kind = ["i32", "u64", "u32", "i64"][i % 4]
        gaps = rng.integers(1, 65, size=n)
        keys = np.cumsum(gaps).astype(ol.KIND_NP[kind])
        if kind[0] == "i":
            keys = (keys.astype(np.int64) - int(keys[n // 2])).astype(ol.KIND_NP[kind])
        q = keys[rng.integers(0, n, size=3 * n)]
        miss = rng.integers(int(keys.min()) - 3, int(keys.max()) + 3, size=n // 2 + 1)
        q = np.concatenate([q, miss.astype(keys.dtype)])
        a_A, a_Q, a_ans = (i + 1) << 40, ((i + 1) << 40) | (1 << 36), ((i + 1) << 40) | (1 << 37)
        arrays[a_A] = keys.copy()
        queries[a_Q] = q
        answers[a_ans] = np.full(q.size, 2 ** 64 - 1, dtype=np.uint64)
        items[i].A, items[i].n, items[i].dtype = a_A, n, tl_amd.DTYPE[kind]
        items[i].layout, items[i].b = i % 3, b
        items[i].Q, items[i].nq, items[i].ans = a_Q, q.size, a_ans
        host.append((a_A, a_Q, a_ans, LAYOUTS[i % 3], kind, keys))
    return items, arrays, queries, answers, host


def search(op, A, Q):
    return ol.search(LAYOUTS[op.layout], A, Q, KIND_OF[op.dtype], int(op.b))


@pytest.mark.parametrize("lds", [0, 32768, 4096])
def test_batch_subtasks_replay_to_oracle(lds):
    items, arrays, queries, answers, host = synthetic_batch(90, 0, 16, 11)
    ops, tasks, blob, smem = be.export(tl_amd.lib(), items, len(items), lds)
    assert len(tasks) > len(ops) > 90 and smem <= 163840
    R = be.Replay(ops, tasks, blob, arrays, queries, answers, search)
    R.smem = smem
    R.run()
    for a_A, a_Q, a_ans, layout, kind, keys in host:
        exp = ol.permute(layout, keys.copy(), 8)
        assert np.array_equal(arrays[a_A], exp), (layout, kind, keys.size)
        exp_ans = ol.search(layout, exp, queries[a_Q], kind, 8)
        assert np.array_equal(answers[a_ans], exp_ans), (layout, kind, keys.size)


def test_batch_export_rejects_bad_items():
    items, *_ = synthetic_batch(3, 4, 6, 12)
    items[1].b = 0
    sizes = (ctypes.c_uint64 * 5)()
    cfg = tl_amd.BatchConfig(0, 0, 0)
    st = tl_amd.lib().tl_batch_export(items, 3, ctypes.byref(cfg), 0, 0, None, 0, None, 0, None,
                                      0, sizes)
    assert st == tl_amd.TL_EINVAL
