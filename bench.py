"""Benchmark of the Tree-Layouts hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): level-order
B-tree, b = 16, n = 2^30 uint64 keys from the cumulative sum of
max(1, ceil(lognormal(0, 2))) gaps (torch CUDA generator, seed 4), in-place
permute, then 2^27 uniform-hit queries answered by plain root descent and by
the learned start-node descent.

A step is one in-place permute of the whole array.  The permutation only
depends on (n, b), so repeating it on the same buffer repeats exactly the
same data movement; the first permute (from sorted input) is the one whose
result is checked.  Queries are timed separately on a freshly built layout.

Prints one JSON line (see the keys below); --gpus N > 1 is rejected because
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tree-layouts_amd", "python"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_keys(n, seed, dev):
    import torch
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    chunk = 1 << 27
    carry = 0
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        gaps = torch.empty(m, dtype=torch.float64, device=dev).log_normal_(0.0, 2.0, generator=g)
        gaps = torch.clamp(torch.ceil(gaps), min=1.0, max=float(1 << 40)).to(torch.int64)
        part = torch.cumsum(gaps, 0) + carry
        keys[s:s + m] = part
        carry = int(part[-1].item())
        del gaps, part
    return keys, g


def sample_check_layout(A, keys_sorted, n, b, rng, samples=200_000):
    """slot of sampled ranks from the oracle's closed form; checker only"""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.oracle()
    ranks = rng.integers(0, n, size=samples, dtype=np.int64)
    pos = np.array([o.or_btree_pos(int(x), n, b) for x in ranks], dtype=np.int64)
    import torch
    got = A[torch.from_numpy(pos).to(A.device)].cpu().numpy()
    exp = keys_sorted[torch.from_numpy(ranks).to(A.device)].cpu().numpy()
    return bool(np.array_equal(got, exp))


def measured_traffic(name):
    """HBM bytes per permute call from a committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs, scaled by a copy of known size)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("Total per call:"):
                return float(line.split("=")[1].split("GB")[0]) * 1e9
    except OSError:
        pass
    return None


def cpu_baseline(keys_np, b, nq_sample, rng):
    """Oracle OpenMP restatement of the reference CPU path (btree-cycles +
    query-btree), on a bounded sample, on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    cores = len(os.sched_getaffinity(0))
    cores = max(1, min(cores, 16))
    n = keys_np.size
    times = []
    for _ in range(3):
        a = keys_np.copy()
        t0 = time.perf_counter()
        ol.permute("btree", a, b, p=cores)
        times.append(time.perf_counter() - t0)
    q = keys_np[rng.integers(0, n, size=nq_sample)]
    t0 = time.perf_counter()
    ans = ol.search("btree", a, q, "u64", b, p=cores)
    tq = time.perf_counter() - t0
    assert (ans < n).all()
    return {"permute_gkeys_s": n / min(times) / 1e9, "query_mq_s": nq_sample / tq / 1e6,
            "cores": cores, "n": n, "nq": nq_sample}


def run_batch_permutes(args, dev):
    """The uneven batch's permutations alone (no queries), with the arrays of
    the batch workload (same sizes, dtypes and layouts) laid out as slices of
    one int32 and one int64 buffer, restored by one copy each between
    replays, so that a profiled run makes few dispatches besides the batch's
    own.  Starts with four calibration copies of 2^27 int64 keys (1 GiB read,
    1 GiB written) that scale the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    import tl_amd
    rng = np.random.default_rng(5)
    count = args.batch_arrays
    sizes = np.floor(2.0 ** rng.uniform(8, 22, size=count)).astype(np.int64)
    kinds = np.where(rng.random(count) < 0.5, "i32", "u64")
    layouts = ["bst", "btree", "mveb"]
    pad = lambda x: (x + 63) // 64 * 64
    n32 = sum(pad(int(z)) for z, k in zip(sizes, kinds) if k == "i32")
    n64 = sum(pad(int(z)) for z, k in zip(sizes, kinds) if k == "u64")
    cal_a = torch.arange(1 << 27, device=dev, dtype=torch.int64)
    cal_b = torch.empty_like(cal_a)
    for _ in range(4):
        cal_b.copy_(cal_a)
    del cal_a, cal_b
    b32 = torch.arange(n32, device=dev, dtype=torch.int32)
    b64 = torch.arange(n64, device=dev, dtype=torch.int64)
    p32, p64 = b32.clone(), b64.clone()
    arrays, o32, o64, keys_total, perm_bytes = [], 0, 0, 0, 0.0
    for i in range(count):
        n, kind = int(sizes[i]), str(kinds[i])
        if kind == "i32":
            A = b32[o32:o32 + n]
            o32 += pad(n)
        else:
            A = b64[o64:o64 + n]
            o64 += pad(n)
        keys_total += n
        perm_bytes += 2.0 * n * (4 if kind == "i32" else 8)
        arrays.append({"A": A, "kind": kind, "layout": layouts[i % 3], "b": 8, "Q": None,
                       "ans": None})
    P = tl_amd.Batch(arrays, lds_bytes=args.batch_lds)
    log(f"[bench] batch permutes: {count} arrays, {keys_total} keys")

    def restore():
        b32.copy_(p32)
        b64.copy_(p64)
        torch.cuda.synchronize()

    times = []
    for step in range(args.warmup + args.steps):
        restore()
        t = P.run(graph=True)
        if step >= args.warmup:
            times.append(t)
        log(f"[bench] batch permute replay {step}: {t:.3f} ms")
    P.close()
    ms = float(np.mean(times))
    traffic = measured_traffic("batch_permutes_traffic.txt")
    out = {"workload": "uneven batch, permutations only", "ms": round(ms, 3),
           "Gkeys_s": round(keys_total / (ms * 1e-3) / 1e9, 3),
           "algorithmic_frac": round(perm_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "traffic": None if traffic is None else round(traffic),
           "traffic_frac": None if traffic is None
           else round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    print(json.dumps(out), flush=True)


def run_batch(args, dev):
    """Uneven batch: 4096 sorted arrays, sizes floor(2^U), U ~ U[8, 22]
    (seed 5), int32 / uint64 keys 50/50, layouts BST / B-tree b=8 / mvEB
    round-robin, keys = cumsum of U{1..64} gaps seeded per array, 4 * size
    uniform-hit queries per array, one permute and one query call per
    array."""
    import torch
    import tl_amd
    rng = np.random.default_rng(5)
    count = args.batch_arrays
    sizes = np.floor(2.0 ** rng.uniform(8, 22, size=count)).astype(np.int64)
    kinds = np.where(rng.random(count) < 0.5, "i32", "u64")
    layouts = ["bst", "btree", "mveb"]
    arrays, pristine = [], []
    keys_total = 0
    perm_bytes = 0.0
    query_bytes = 0.0
    log(f"[bench] generating {count} arrays, {int(sizes.sum())} keys")
    for i in range(count):
        n, kind, lay = int(sizes[i]), str(kinds[i]), layouts[i % 3]
        g = torch.Generator(device=dev)
        g.manual_seed(5 * 1_000_003 + i)
        tdt = torch.int32 if kind == "i32" else torch.int64
        gaps = torch.randint(1, 65, (n,), generator=g, device=dev, dtype=torch.int64)
        keys = torch.cumsum(gaps, 0).to(tdt)
        qi = torch.randint(0, n, (4 * n,), generator=g, device=dev, dtype=torch.int64)
        Q = keys[qi]
        del gaps, qi
        w = 4 if kind == "i32" else 8
        arrays.append({"A": keys.clone(), "kind": kind, "layout": lay, "b": 8, "Q": Q,
                       "ans": torch.empty(4 * n, dtype=torch.int64, device=dev)})
        pristine.append(keys)
        keys_total += n
        perm_bytes += 2.0 * n * w
        if lay == "btree":
            depth = 1
            cap = 8
            while cap < n:
                cap = cap * 9 + 8
                depth += 1
            node = 8 * w
        else:
            depth = int(np.floor(np.log2(n))) + 1
            node = w
        query_bytes += 4.0 * n * (w + 8 + depth * node)
    torch.cuda.synchronize()
    # tl_batch_run: every plan built and uploaded at create time, the arrays
    # that fit in LDS in batched launches, one HIP graph replayed per step
    t0 = time.perf_counter()
    B = tl_amd.Batch(arrays, lds_bytes=args.batch_lds)
    create_s = time.perf_counter() - t0

    def restore():
        for a, p in zip(arrays, pristine):
            a["A"].copy_(p)
        torch.cuda.synchronize()

    restore()
    B.run(graph=True)
    torch.cuda.synchronize()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    ok = True
    for i in range(0, count, max(1, count // 48)):
        a = arrays[i]
        keys = pristine[i].cpu().numpy()
        if a["kind"] == "u64":
            keys = keys.view(np.uint64)
        exp = ol.permute(a["layout"], keys.copy(), 8)
        got = a["A"].cpu().numpy().view(keys.dtype)
        ok &= bool(np.array_equal(got, exp))
        q = a["Q"].cpu().numpy().view(keys.dtype)
        ok &= bool(np.array_equal(a["ans"].cpu().numpy().view(np.uint64),
                                  ol.search(a["layout"], exp, q, a["kind"], 8)))
    log(f"[bench] batch sample check: {ok}")
    times = []
    for step in range(args.warmup + args.steps):
        restore()
        t = B.run(graph=True)
        if step >= args.warmup:
            times.append(t)
        log(f"[bench] batch step {step}: {t:.2f} ms")
    ms = float(np.mean(times))
    restore()
    sep_ms = B.run_separate()
    log(f"[bench] per-array calls: {sep_ms:.2f} ms")
    # the permutes alone, through a batch without queries
    perm_only = [dict(a, Q=None) for a in arrays]
    P = tl_amd.Batch(perm_only, lds_bytes=args.batch_lds)
    restore()
    P.run(graph=True)
    ptimes = []
    for _ in range(max(1, args.steps)):
        restore()
        ptimes.append(P.run(graph=True))
    perm_ms = float(np.mean(ptimes))
    P.close()
    log(f"[bench] permutes alone: {perm_ms:.2f} ms")
    total_bytes = perm_bytes + query_bytes
    achieved = total_bytes / (ms * 1e-3) / 1e9
    # HBM bytes of one permutations-only replay, from the committed PMC summary
    perm_traffic = measured_traffic("batch_permutes_traffic.txt")
    perm_traffic = None if perm_traffic is None else round(perm_traffic)
    out = {
        "metric": "Gkeys/s permute + Mqueries/s, n=2^30 u64; learned vs plain start, exact answers",
        "value": round(keys_total / (ms * 1e-3) / 1e9, 3), "unit": "Gkeys/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i32/u64",
        "data": "synthetic: U{1..64} gaps per array (torch CUDA generator), sizes seed 5",
        "config": {"workload": f"uneven batch: {count} arrays, sizes floor(2^U) U~U[8,22], "
                               "BST / B-tree b=8 / mvEB round robin, 4*size hit queries each; "
                               "tl_batch_run replayed from its HIP graph",
                   "keys": keys_total, "queries": 4 * keys_total,
                   "lds_bytes": args.batch_lds or "device budget"},
        "per_array_calls_ms": round(sep_ms, 3), "create_s": round(create_s, 3),
        "queries": {"Mq_s": round(4 * keys_total / (ms * 1e-3) / 1e6, 1)},
        "permutes_alone": {"ms": round(perm_ms, 3),
                           "Gkeys_s": round(keys_total / (perm_ms * 1e-3) / 1e9, 3),
                           "hbm_frac": round(perm_bytes / (perm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "traffic": perm_traffic,
                           "traffic_frac": None if perm_traffic is None else round(
                               perm_traffic / (perm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "checks": {"sample_arrays": ok},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batch (all per-array permute and query launches)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    B.close()


def time_unpermute(args, A, layout, b, stream):
    """ms per in-place inverse permutation (layout -> sorted order) of A; the
    data movement does not depend on the values, so A's contents do not matter"""
    import torch
    import tl_amd
    for _ in range(max(1, args.warmup)):
        tl_amd.unpermute(layout, A, b=b, kind="u64", timed=False)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        tl_amd.unpermute(layout, A, b=b, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / args.steps


def run_layout(args, dev, layout):
    """Level-order BST or mvEB at n = 2^log2n uint64 (same key generator as the
    headline): in-place permute timing and plain query timing."""
    import torch
    import tl_amd
    n, nq = 1 << args.log2n, 1 << args.log2q
    rng = np.random.default_rng(4)
    log(f"[bench] {layout}: generating n={n} keys and {nq} queries")
    keys, g = gen_keys(n, 4, dev)
    qidx = torch.randint(0, n, (nq,), generator=g, device=dev, dtype=torch.int64)
    Q = keys[qidx]
    del qidx
    A = keys.clone()
    stream = torch.cuda.current_stream()
    tl_amd.permute(layout, A, kind="u64")
    torch.cuda.synchronize()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.oracle()
    ranks = rng.integers(0, n, size=100_000, dtype=np.int64)
    if layout == "bst":
        pos = np.array([o.or_btree_pos(int(x), n, 1) for x in ranks], dtype=np.int64)
    elif layout == "veb":
        pos = np.array([o.or_classic_pos(int(x), n) for x in ranks], dtype=np.int64)
    else:
        pos = np.array([o.or_mveb_pos(int(x), n) for x in ranks], dtype=np.int64)
    got = A[torch.from_numpy(pos).to(dev)].cpu().numpy()
    layout_ok = bool(np.array_equal(got, keys[torch.from_numpy(ranks).to(dev)].cpu().numpy()))
    log(f"[bench] layout sample check: {layout_ok}")
    for _ in range(args.warmup):
        tl_amd.permute(layout, A, kind="u64", timed=False)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        tl_amd.permute(layout, A, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    perm_ms = ev0.elapsed_time(ev1) / args.steps
    log(f"[bench] permute {perm_ms:.3f} ms/step")
    unperm_ms = time_unpermute(args, A, layout, 1, stream)
    log(f"[bench] unpermute {unperm_ms:.3f} ms/step")
    A.copy_(keys)
    tl_amd.permute(layout, A, kind="u64")
    ans = torch.empty(nq, dtype=torch.int64, device=dev)
    if layout == "bst":
        def q():
            tl_amd.query_bst(A, Q, ans, kind="u64", timed=False)
    else:
        classic = layout == "veb"
        tables, idx, tidx, meta = tl_amd.veb_tableset_device(n, classic=classic)
        d, num_tables = int(meta[0].item()), int(meta[1].item())
        search = tl_amd.query_veb_nonperfect if classic else tl_amd.query_mveb_nonperfect

        def q():
            search(A, tables, idx, tidx, d, num_tables, Q, ans, kind="u64", timed=False)
    q()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        q()
    ev1.record(stream)
    torch.cuda.synchronize()
    q_ms = ev0.elapsed_time(ev1) / args.steps
    ok = bool(torch.equal(A[ans.clamp(0, n - 1)], Q)) and bool(((ans >= 0) & (ans < n)).all())
    log(f"[bench] query {q_ms:.3f} ms ({nq / q_ms / 1e3:.1f} Mq/s) ok={ok}")
    achieved = 2.0 * n * 8 / (perm_ms * 1e-3) / 1e9
    out = {
        "metric": "Gkeys/s permute + Mqueries/s, n=2^30 u64; learned vs plain start, exact answers",
        "value": round(n / (perm_ms * 1e-3) / 1e9, 3), "unit": "Gkeys/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(perm_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: lognormal(0,2) gaps, torch CUDA generator seed 4",
        "config": {"workload": f"level-order {layout}, in-place permute of n=2^{args.log2n} uint64 "
                               f"keys; 2^{args.log2q} uniform-hit queries", "n": n, "queries": nq},
        "queries": {"plain_ms": round(q_ms, 4), "plain_Mq_s": round(nq / q_ms / 1e3, 2)},
        "unpermute": {"ms": round(unperm_ms, 4), "Gkeys_s": round(n / (unperm_ms * 1e-3) / 1e9, 3)},
        "checks": {"layout_sample": layout_ok, "answers": ok},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"whole tl_permute_{layout} call (all its launches)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=30)
    ap.add_argument("--log2q", type=int, default=27)
    ap.add_argument("--b", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--log2leaves", type=int, default=20, help="learned model leaves (log2)")
    ap.add_argument("--workload", choices=["btree", "batch", "bst", "mveb", "veb"],
                    default="btree")
    ap.add_argument("--batch-arrays", type=int, default=4096)
    ap.add_argument("--batch-perm-only", action="store_true",
                    help="time the batch's permutations alone (few other dispatches)")
    ap.add_argument("--batch-lds", type=int, default=0,
                    help="LDS budget of the batch's plans in bytes (0: the library default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(a permute would split the ranks into one contiguous slice per "
              "GPU and exchange keys by destination with one all-to-all)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import tl_amd
    dev = torch.device("cuda:0")
    if args.workload == "batch":
        if args.batch_perm_only:
            return run_batch_permutes(args, dev)
        return run_batch(args, dev)
    if args.workload in ("bst", "mveb", "veb"):
        return run_layout(args, dev, args.workload)
    n, nq, b = 1 << args.log2n, 1 << args.log2q, args.b
    rng = np.random.default_rng(4)
    log(f"[bench] generating n={n} keys and {nq} queries")
    keys, g = gen_keys(n, 4, dev)
    qidx = torch.randint(0, n, (nq,), generator=g, device=dev, dtype=torch.int64)
    Q = keys[qidx]
    del qidx
    A = keys.clone()
    stream = torch.cuda.current_stream()

    # ---- correctness of the first permute (from sorted input)
    log("[bench] first permute + layout check")
    tl_amd.permute_btree(A, b, kind="u64")
    torch.cuda.synchronize()
    layout_ok = sample_check_layout(A, keys, n, b, rng)
    log(f"[bench] layout sample check: {layout_ok}")

    # ---- timed permutes (same buffer, same data movement every step)
    for _ in range(args.warmup):
        tl_amd.permute_btree(A, b, kind="u64", timed=False)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        tl_amd.permute_btree(A, b, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_wall) * 1e3
    perm_ms = ev0.elapsed_time(ev1) / args.steps
    log(f"[bench] permute {perm_ms:.3f} ms/step (host wall {wall_ms / args.steps:.3f})")
    unperm_ms = time_unpermute(args, A, "btree", b, stream)
    log(f"[bench] unpermute {unperm_ms:.3f} ms/step")

    # ---- rebuild the true layout for the queries
    A.copy_(keys)
    tl_amd.permute_btree(A, b, kind="u64")
    ans = torch.empty(nq, dtype=torch.int64, device=dev)
    for _ in range(max(1, args.warmup)):
        tl_amd.query_btree(A, b, Q, ans, kind="u64", timed=False)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        tl_amd.query_btree(A, b, Q, ans, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    q_ms = ev0.elapsed_time(ev1) / args.steps
    plain_ok = bool(((ans >= 0) & (ans < n)).all().item()) and bool(
        torch.equal(A[ans.clamp(0, n - 1)], Q))
    log(f"[bench] plain query {q_ms:.3f} ms ({nq / q_ms / 1e3:.1f} Mq/s) ok={plain_ok}")

    # ---- the comparator: binary search of the sorted keys, same queries (its
    # own answer buffer: the plain answers are what the learned ones must equal)
    bans = torch.empty_like(ans)
    for _ in range(max(1, args.warmup)):
        tl_amd.query_binary(keys, Q, bans, kind="u64", timed=False)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        tl_amd.query_binary(keys, Q, bans, kind="u64", timed=False)
    ev1.record(stream)
    torch.cuda.synchronize()
    bs_ms = ev0.elapsed_time(ev1) / args.steps
    bs_ok = bool(((bans >= 0) & (bans < n)).all().item()) and bool(
        torch.equal(keys[bans.clamp(0, n - 1)], Q))
    del bans
    log(f"[bench] binary search {bs_ms:.3f} ms ({nq / bs_ms / 1e3:.1f} Mq/s) ok={bs_ok}")

    learned = None
    if not args.skip_learned and hasattr(tl_amd, "learned_fit"):
        learned = tl_amd.bench_learned(A, keys, Q, ans, b, args.steps, args.warmup, n,
                                       num_leaves=1 << args.log2leaves)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline sample")
        ns = 1 << 25
        keys_np = keys[:ns].cpu().numpy()
        cpu = cpu_baseline(keys_np, b, 1 << 22, rng)
        log(f"[bench] CPU {cpu}")

    algo_bytes = 2.0 * n * 8
    achieved = algo_bytes / (perm_ms * 1e-3) / 1e9
    traffic = None
    if n == 1 << 30 and b == 16:
        traffic = measured_traffic("btree_b16_n2p30_u64_permute_traffic.txt")
    out = {
        "metric": "Gkeys/s permute + Mqueries/s, n=2^30 u64; learned vs plain start, exact answers",
        "value": round(n / (perm_ms * 1e-3) / 1e9, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(perm_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: lognormal(0,2) gaps, torch CUDA generator seed 4",
        "config": {"workload": "level-order B-tree b=16, in-place permute of n=2^30 uint64 keys; "
                               "2^27 uniform-hit queries (plain and learned start)",
                   "n": n, "b": b, "queries": nq},
        "unpermute": {"ms": round(unperm_ms, 4), "Gkeys_s": round(n / (unperm_ms * 1e-3) / 1e9, 3)},
        "queries": {"plain_ms": round(q_ms, 4), "plain_Mq_s": round(nq / q_ms / 1e3, 2),
                    "learned": learned,
                    "binary_search": {"ms": round(bs_ms, 4), "Mq_s": round(nq / bs_ms / 1e3, 2),
                                      "answers_ok": bs_ok}},
        "checks": {"layout_sample": layout_ok, "plain_answers": plain_ok},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_GBs": None if traffic is None
                     else round(traffic / (perm_ms * 1e-3) / 1e9, 1),
                     "traffic_frac": None if traffic is None
                     else round(traffic / (perm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "kernel": "one tl_permute_btree call (all its launches); traffic from "
                               "profiles/btree_b16_n2p30_u64_permute_traffic.txt"},
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["permute_gkeys_s"], 4), "unit": "Gkeys/s", "cores": cpu["cores"],
            "kind": "port",
            "sample": f"oracle OpenMP restatement of btree-cycles, n=2^25 prefix of the same "
                      f"keys, best of 3; query {cpu['query_mq_s']:.2f} Mq/s on 2^22 hits"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
