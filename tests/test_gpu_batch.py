"""Uneven batch: every array of a mixed batch (sizes spread over four orders
of magnitude, int32/uint64 keys, BST / B-tree b=8 / mvEB) comes out of the
batch engine with the oracle's layout and the oracle's answers: the planned
path (tl_batch_run, eager and replayed from its HIP graph) and the per-array
calls (tl_batch_run_separate).  The single persistent launch is disabled; its
tests are skipped with the reason."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
import tl_amd

pytestmark = pytest.mark.gpu

LAYOUTS = ["bst", "btree", "mveb"]


def make_batch(count, lo, hi, seed, dev, qmult=4):
    rng = np.random.default_rng(seed)
    arrays, host = [], []
    for i in range(count):
        n = int(2 ** rng.uniform(lo, hi))
        kind = "i32" if rng.random() < 0.5 else "u64"
        layout = LAYOUTS[i % 3]
        gaps = rng.integers(1, 65, size=n)
        keys = np.cumsum(gaps).astype(np.int32 if kind == "i32" else np.uint64)
        if kind == "i32":
            keys = (keys.astype(np.int64) - int(keys[n // 2])).astype(np.int32)
        nq = qmult * n
        q = keys[rng.integers(0, n, size=nq)]
        miss = rng.integers(int(keys.min()) - 3, int(keys.max()) + 3, size=nq // 4)
        q = np.concatenate([q, miss.astype(keys.dtype)])
        st = np.int32 if kind == "i32" else np.int64
        A = torch.from_numpy(keys.view(st).copy()).to(dev)
        Q = torch.from_numpy(q.view(st).copy()).to(dev)
        ans = torch.full((q.size,), -1, dtype=torch.int64, device=dev)
        arrays.append({"A": A, "kind": kind, "layout": layout, "b": 8, "Q": Q, "ans": ans})
        host.append((keys, q))
    return arrays, host


def check(arrays, host):
    for a, (keys, q) in zip(arrays, host):
        exp_layout = ol.permute(a["layout"], keys.copy(), 8)
        got = a["A"].cpu().numpy().view(keys.dtype)
        assert np.array_equal(got, exp_layout), (a["layout"], a["kind"], keys.size)
        exp_ans = ol.search(a["layout"], exp_layout, q, a["kind"], 8)
        assert np.array_equal(a["ans"].cpu().numpy().view(np.uint64), exp_ans), \
            (a["layout"], a["kind"], keys.size)


def restore(arrays, host, dev):
    for a, (keys, _) in zip(arrays, host):
        st = np.int32 if keys.dtype.itemsize == 4 else np.int64
        a["A"].copy_(torch.from_numpy(keys.view(st).copy()).to(dev))
        a["ans"].fill_(-1)


PERSISTENT_OFF = ("the single persistent launch stalled on gfx950 (twice, the second time with "
                  "wall-clock-bounded waits) and is disabled until its cause is shown; the "
                  "sub-task list is checked by the host replay in test_batch_plan.py and the "
                  "executors by the CPU emulation in test_batch_emulated.py")


@pytest.mark.skip(reason=PERSISTENT_OFF)
@pytest.mark.parametrize("lds", [0, 32768])
def test_batch_matches_oracle(lds, dev):
    arrays, host = make_batch(240, 0, 18, 5, dev)
    B = tl_amd.Batch(arrays, lds_bytes=lds)
    st = B.stats()
    assert st["tasks"] > st["ops"] > 240
    B.run(graph=False)
    torch.cuda.synchronize()
    check(arrays, host)
    # graph capture and two replays, with the arrays restored in between
    for _ in range(2):
        restore(arrays, host, dev)
        torch.cuda.synchronize()
        B.run(graph=True)
        torch.cuda.synchronize()
        check(arrays, host)
    B.close()


def test_batch_separate_path_matches(dev):
    arrays, host = make_batch(60, 0, 16, 6, dev)
    B = tl_amd.Batch(arrays)
    B.run_separate()
    torch.cuda.synchronize()
    check(arrays, host)
    B.close()


@pytest.mark.parametrize("waves", ["1", "0"])
@pytest.mark.parametrize("lds", [0, 32768])
def test_batch_planned_path_matches_oracle(lds, waves, dev, monkeypatch):
    """tl_batch_run (the planned path: small arrays in batched LDS launches,
    the others by waves of multi-array launches, or each on its own plan with
    TL_BATCH_WAVES=0; queries per array) matches the oracle eagerly and on
    three replays of its HIP graph, with the arrays restored and the
    program's own work and synchronizations in between."""
    monkeypatch.setenv("TL_BATCH_WAVES", waves)
    arrays, host = make_batch(240, 0, 18, 5, dev)
    B = tl_amd.Batch(arrays, lds_bytes=lds)
    B.run(graph=False)
    torch.cuda.synchronize()
    check(arrays, host)
    for _ in range(3):
        restore(arrays, host, dev)
        torch.cuda.synchronize()
        B.run(graph=True)
        torch.cuda.synchronize()
        check(arrays, host)
    B.close()


@pytest.mark.parametrize("knobs", [{"TL_ROW_STASH_MAX": "0"}, {"TL_WALK_LANES": "0"},
                                   {"TL_ROW_STASH_MAX": "0", "TL_WALK_LANES": "0"}])
def test_batch_schedule_with_single_ops(knobs, dev, monkeypatch):
    """Plans with unit merges (no row slides) and rotations by reversals (no
    lane walks): ops without a multi-array kernel run alone inside the
    multi-array launch schedule, which must keep every array's ops in order"""
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    monkeypatch.setenv("TL_BATCH_WAVES", "1")
    arrays, host = make_batch(120, 6, 18, 11, dev)
    B = tl_amd.Batch(arrays)
    B.run(graph=False)
    torch.cuda.synchronize()
    check(arrays, host)
    restore(arrays, host, dev)
    torch.cuda.synchronize()
    B.run(graph=True)
    torch.cuda.synchronize()
    check(arrays, host)
    B.close()


def test_batch_planned_path_autotune(dev):
    """The autotuner returns a configuration inside its search space, and the
    batch built with it matches the oracle (eagerly and from its graph)."""
    arrays, host = make_batch(90, 6, 16, 7, dev)
    cfg = tl_amd.batch_autotune(arrays, sample=30)
    assert cfg["lds_bytes"] in (0, 32768, 65536, 98304, 163840)
    assert cfg["tile_k"] in (5, 6) and cfg["stream_threads"] in (256, 512, 1024)
    assert cfg["btree_qpg"] in (2, 4) and cfg["veb_qpl"] in (1, 2, 4)
    assert cfg["query_queue"] in (0, 1)
    assert 0 < cfg["sample_ms"] < 1e29
    B = tl_amd.Batch(arrays, **{k: cfg[k] for k in tl_amd.BATCH_KNOBS})
    B.run(graph=False)
    torch.cuda.synchronize()
    check(arrays, host)
    restore(arrays, host, dev)
    torch.cuda.synchronize()
    B.run(graph=True)
    torch.cuda.synchronize()
    check(arrays, host)
    B.close()


@pytest.mark.parametrize("knobs", [{"tile_k": 6, "stream_threads": 256, "btree_qpg": 2,
                                    "veb_qpl": 1},
                                   {"tile_k": 5, "stream_threads": 512, "btree_qpg": 4,
                                    "veb_qpl": 4, "lds_bytes": 32768, "query_queue": 1}])
def test_batch_knobs_keep_results(knobs, dev):
    """Every knob the autotuner turns leaves layouts and answers unchanged."""
    arrays, host = make_batch(120, 4, 18, 13, dev)
    B = tl_amd.Batch(arrays, **knobs)
    B.run(graph=True)
    torch.cuda.synchronize()
    check(arrays, host)
    B.close()


@pytest.mark.parametrize("lds,chunk", [(0, None), (32768, None), (0, "1024"), (0, "65536")])
def test_batch_query_queue_matches_oracle(lds, chunk, dev, monkeypatch):
    """The persistent query queues (one work-queue launch per layout:
    workgroups take chunks of queries of any array from an atomic ticket)
    answer every array like the oracle, eagerly and on three graph replays;
    chunks of 1024 to 65536 queries"""
    if chunk:
        monkeypatch.setenv("TL_QQ_CHUNK", chunk)
    arrays, host = make_batch(240, 0, 18, 5, dev)
    B = tl_amd.Batch(arrays, lds_bytes=lds, query_queue=1)
    B.run(graph=False)
    torch.cuda.synchronize()
    check(arrays, host)
    for _ in range(3):
        restore(arrays, host, dev)
        torch.cuda.synchronize()
        B.run(graph=True)
        torch.cuda.synchronize()
        check(arrays, host)
    B.close()


def test_batch_query_queue_with_leftovers(dev):
    """B-tree arrays whose nodes are not whole 16-byte words (b = 3) are
    answered by their own launches beside the queues; all answers match"""
    arrays, host = make_batch(90, 4, 17, 23, dev)
    bs = []
    for i, a in enumerate(arrays):
        b = 3 if (a["layout"] == "btree" and i % 2) else 8
        a["b"] = b
        bs.append(b)
    B = tl_amd.Batch(arrays, query_queue=1)
    B.run(graph=True)
    torch.cuda.synchronize()
    for a, (keys, q), b in zip(arrays, host, bs):
        exp_layout = ol.permute(a["layout"], keys.copy(), b)
        got = a["A"].cpu().numpy().view(keys.dtype)
        assert np.array_equal(got, exp_layout), (a["layout"], b, keys.size)
        exp_ans = ol.search(a["layout"], exp_layout, q, a["kind"], b)
        assert np.array_equal(a["ans"].cpu().numpy().view(np.uint64), exp_ans), \
            (a["layout"], b, keys.size)
    B.close()

