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


def timing_stats(count, runs_ms, scale=1e-9):
    """printTimings-style summary of per-run times (cpu/src/common.cpp:31-58,
    gpu/src/common.cu:41-67): mean and sd of the times, the throughput of the
    mean time (count per second * scale) and the spread of the per-run
    throughputs about it, and every run."""
    t = np.asarray(runs_ms, dtype=np.float64)
    mean = float(t.mean())
    tp = count / (t * 1e-3) * scale
    tp_mean = count / (mean * 1e-3) * scale
    return {"runs": len(t), "mean_ms": round(mean, 4), "sd_ms": round(float(t.std()), 4),
            "tp_mean": round(tp_mean, 3),
            "tp_sd": round(float(np.sqrt(np.mean((tp - tp_mean) ** 2))), 3),
            "runs_ms": [round(float(x), 4) for x in t]}


def event_steps(fn, steps, stream):
    """Runs fn() `steps` times on `stream` with an event between calls;
    returns (total ms, per-step ms)."""
    import torch
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    evs[0].record(stream)
    for i in range(steps):
        fn()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
    return evs[0].elapsed_time(evs[-1]), per


def btree_query_bytes(n, b, w, cache_bytes=256 << 20):
    """Algorithmic bytes per uniform-hit query of a level-order B-tree: the
    query key, the 8-byte answer and one b*w node per visited level.  Returns
    (all levels, DRAM-resident levels only): levels are counted from the root
    while their cumulative size fits the on-die cache (the 256 MiB Infinity
    Cache), and a hit visits a level with the probability that its key lies
    at that level or below."""
    k1 = b + 1
    h, p = 0, 1                            # p = k1^h <= n < k1^(h+1)
    while p * k1 <= n:
        p *= k1
        h += 1
    if n == p * k1 - 1:                    # perfect: h + 1 full levels
        levels = [k1 ** t * b for t in range(h + 1)]
    else:                                  # perfect internal tree + leaf level
        levels = [k1 ** t * b for t in range(h)] + [n - (p - 1)]
    below = n
    all_b, dram_b, cum = 0.0, 0.0, 0
    for keys_here in levels:
        reach = below / n                  # P(a hit reaches this level)
        cum += keys_here * w
        all_b += reach * b * w
        if cum > cache_bytes:
            dram_b += reach * b * w
        below -= keys_here
    return w + 8 + all_b, w + 8 + dram_b


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


def measured_traffic_entry(name):
    """All 'key = value' lines of a committed rocprofv3 summary's 'Bench:'
    block, or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    out, on = {}, False
    for line in lines:
        if line.startswith("Bench:"):
            on = True
            continue
        if on:
            if "=" not in line:
                break
            k, v = line.split("=", 1)
            try:
                out[k.strip()] = float(v.split()[0])
            except ValueError:
                out[k.strip()] = v.strip()
    return out or None


def cpu_baseline(keys_np, b, nq_sample, rng, iters=10, budget_s=30.0):
    """The reference's CPU B-tree path (btree-cycles permute, query-btree
    search; cpu/src/btree-cycles.h:71-96, cpu/src/query-btree.h:59-64), as
    compiled from the reference sources into oracle/_ref/libref.so, timed like
    the reference's drivers: `iters` runs, the sorted input restored before
    each permute (outside the timed region), query generation excluded
    (cpu/test/build-btree-cycles.cpp, cpu/src/query-btree.h:68-80).  Without
    that library the oracle's OpenMP restatement runs instead, which is about
    15x slower than the reference (kind "port").  Runs stop early once
    budget_s of CPU time is spent.  Threads: the CPUs this process may use,
    at most 16 (the host share of one GPU)."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    n = keys_np.size
    r = ol.ref()
    kind = "reference" if r is not None else "port"
    a = np.empty_like(keys_np)
    times, spent = [], 0.0
    for _ in range(iters):
        np.copyto(a, keys_np)
        t0 = time.perf_counter()
        if r is not None:
            r.ref_permute_btree_u64(a.ctypes.data_as(ol.VP), n, b, cores)
        else:
            ol.permute("btree", a, b, p=cores)
        dt = time.perf_counter() - t0
        times.append(dt * 1e3)
        spent += dt
        if spent > budget_s:
            break
    q = keys_np[rng.integers(0, n, size=nq_sample)]
    ans = np.empty(nq_sample, dtype=np.uint64)
    qtimes = []
    for _ in range(iters):
        t0 = time.perf_counter()
        if r is not None:
            r.ref_search_btree_u64(a.ctypes.data_as(ol.VP), n, b, q.ctypes.data_as(ol.VP),
                                   nq_sample, ans.ctypes.data_as(ol.VP), cores)
        else:
            ans = ol.search("btree", a, q, "u64", b, p=cores)
        dt = time.perf_counter() - t0
        qtimes.append(dt * 1e3)
        spent += dt
        if spent > 2 * budget_s:
            break
    assert (ans < n).all() and np.array_equal(a[ans.astype(np.int64)], q)
    return {"kind": kind, "cores": cores, "n": n, "nq": nq_sample,
            "permute": timing_stats(n, times), "query": timing_stats(nq_sample, qtimes, 1e-6)}


def measure_batch_permutes(args, dev, calibrate=False):
    """The uneven batch's permutations alone (no queries): the arrays of the
    batch workload (4096 arrays, sizes floor(2^U), U ~ U[8, 22], seed 5,
    int32 / uint64 50/50, BST / B-tree b=8 / mvEB round robin) as slices of
    one int32 and one int64 buffer, restored by one copy each between graph
    replays.  With `calibrate`, four copies of 2^27 int64 keys (1 GiB read,
    1 GiB written) open the run to scale the FETCH_SIZE / WRITE_SIZE counters
    of a profiled run."""
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
    if calibrate:
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
    t0 = time.perf_counter()
    P = tl_amd.Batch(arrays, lds_bytes=args.batch_lds)
    create_s = time.perf_counter() - t0
    log(f"[bench] batch permutes: {count} arrays, {keys_total} keys, created in {create_s:.1f} s")

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
    del P, arrays, b32, b64, p32, p64
    torch.cuda.empty_cache()
    ms = float(np.mean(times))
    traffic = measured_traffic("batch_permutes_traffic.txt")
    return {"workload": f"uneven batch, permutations only: {count} arrays, {keys_total} keys, "
                        "one HIP graph replay per step",
            "ms": round(ms, 3), "Gkeys_s": round(keys_total / (ms * 1e-3) / 1e9, 3),
            "timings": timing_stats(keys_total, times),
            "algorithmic_bytes": round(perm_bytes),
            "algorithmic_frac": round(perm_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic),
            "traffic_frac": None if traffic is None
            else round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "create_s": round(create_s, 2)}


def run_batch_permutes(args, dev):
    print(json.dumps(measure_batch_permutes(args, dev, calibrate=True)), flush=True)


def config5_arrays(dev, count=4096, queries=True):
    """The uneven-batch workload: `count` sorted arrays, sizes floor(2^U),
    U ~ U[8, 22] (numpy seed 5), int32 / uint64 keys 50/50 (same generator),
    layouts BST / B-tree b=8 / mvEB round robin, keys = cumsum of U{1..64}
    gaps (torch CUDA generator seeded 5*1000003 + i), and with `queries`
    4 * size uniform-hit queries per array.  Returns (arrays, pristine
    keys, algorithmic permute bytes, algorithmic query bytes)."""
    import torch
    rng = np.random.default_rng(5)
    sizes = np.floor(2.0 ** rng.uniform(8, 22, size=count)).astype(np.int64)
    kinds = np.where(rng.random(count) < 0.5, "i32", "u64")
    layouts = ["bst", "btree", "mveb"]
    arrays, pristine = [], []
    perm_bytes = 0.0
    query_bytes = 0.0
    for i in range(count):
        n, kind, lay = int(sizes[i]), str(kinds[i]), layouts[i % 3]
        g = torch.Generator(device=dev)
        g.manual_seed(5 * 1_000_003 + i)
        tdt = torch.int32 if kind == "i32" else torch.int64
        gaps = torch.randint(1, 65, (n,), generator=g, device=dev, dtype=torch.int64)
        keys = torch.cumsum(gaps, 0).to(tdt)
        Q = ans = None
        if queries:
            qi = torch.randint(0, n, (4 * n,), generator=g, device=dev, dtype=torch.int64)
            Q = keys[qi]
            ans = torch.empty(4 * n, dtype=torch.int64, device=dev)
            del qi
        del gaps
        w = 4 if kind == "i32" else 8
        arrays.append({"A": keys.clone(), "kind": kind, "layout": lay, "b": 8, "Q": Q,
                       "ans": ans})
        pristine.append(keys)
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
        if queries:
            query_bytes += 4.0 * n * (w + 8 + depth * node)
    return arrays, pristine, perm_bytes, query_bytes


def run_batch(args, dev):
    """Uneven batch (config5_arrays): one permute and query pass over the
    whole batch per step, through tl_batch_run replayed from its HIP graph."""
    import torch
    import tl_amd
    count = args.batch_arrays
    log(f"[bench] generating {count} arrays")
    arrays, pristine, perm_bytes, query_bytes = config5_arrays(dev, count)
    keys_total = sum(p.numel() for p in pristine)
    torch.cuda.synchronize()
    # tl_batch_run: every plan built and uploaded at create time, the arrays
    # that fit in LDS in batched launches, one HIP graph replayed per step
    t0 = time.perf_counter()
    B = tl_amd.Batch(arrays, lds_bytes=args.batch_lds, query_queue=args.batch_query_queue)
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


def layout_traffic(layout, n):
    """HBM bytes per permute call of the BST or mvEB at n = 2^30 from the
    committed PMC summaries (None for other sizes and layouts)"""
    if n != 1 << 30 or layout not in ("bst", "mveb"):
        return None
    t = measured_traffic(f"{layout}_n2p30_u64_permute_traffic.txt")
    return None if t is None else round(t)


def involution_roofline(layout, n, ms):
    """Algorithmic fraction of one involution-family permute (2 * n * 8 bytes
    over the call) and, at n = 2^30, its HBM bytes per call from the committed
    PMC summary with the fraction they make of the peak"""
    out = {"algorithmic_frac": round(2.0 * n * 8 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "traffic": None, "traffic_frac": None}
    t = measured_traffic_entry("involutions_n2p30_u64_traffic.txt") if n == 1 << 30 else None
    if t and f"{layout}_read_GB" in t:
        traffic = (t[f"{layout}_read_GB"] + t[f"{layout}_write_GB"]) * 1e9
        out["traffic"] = round(traffic)
        out["traffic_frac"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def other_layouts(args, A, keys, stream, rng, samples=20_000):
    """BST, mvEB and classic vEB permutes (and inverses) of the headline's
    n = 2^log2n uint64 keys in the same buffer: a sampled slot check against
    the oracle's position functions, then ms per call over steps/2 calls"""
    import torch
    import tl_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    o = ol.oracle()
    n = A.numel()
    ranks = rng.integers(0, n, size=samples, dtype=np.int64)
    want = keys[torch.from_numpy(ranks).to(A.device)].cpu().numpy()
    steps = max(1, args.steps // 2)
    out = {}
    for layout in ("bst", "mveb", "veb"):
        A.copy_(keys)
        tl_amd.permute(layout, A, kind="u64")
        torch.cuda.synchronize()
        if layout == "bst":
            pos = [o.or_btree_pos(int(x), n, 1) for x in ranks]
        elif layout == "veb":
            pos = [o.or_classic_pos(int(x), n) for x in ranks]
        else:
            pos = [o.or_mveb_pos(int(x), n) for x in ranks]
        got = A[torch.tensor(pos, dtype=torch.int64, device=A.device)].cpu().numpy()
        ok = bool(np.array_equal(got, want))
        tl_amd.permute(layout, A, kind="u64", timed=False)
        total, per = event_steps(lambda: tl_amd.permute(layout, A, kind="u64", timed=False),
                                 steps, stream)
        ms = total / steps
        utotal, _ = event_steps(lambda: tl_amd.unpermute(layout, A, kind="u64", timed=False),
                                steps, stream)
        ums = utotal / steps
        traffic = layout_traffic(layout, n)
        out[layout] = {"ms": round(ms, 4), "Gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                       "algorithmic_frac": round(2.0 * n * 8 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "traffic": traffic,
                       "traffic_frac": None if traffic is None
                       else round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "unpermute_ms": round(ums, 4), "layout_sample": ok, "runs": steps}
        log(f"[bench] {layout} permute {ms:.3f} ms, unpermute {ums:.3f} ms, sample check {ok}")
    # the involution family: same layouts by swap passes, one timed call each
    # from the sorted keys, with the same sampled slot check
    inv = {}
    for layout, b in (("bst", 1), ("btree", 16), ("mveb", 0)):
        A.copy_(keys)
        ms = tl_amd.permute_involutions(layout, A, b=b, kind="u64")
        if layout == "mveb":
            pos = [o.or_mveb_pos(int(x), n) for x in ranks]
        else:
            pos = [o.or_btree_pos(int(x), n, b) for x in ranks]
        got = A[torch.tensor(pos, dtype=torch.int64, device=A.device)].cpu().numpy()
        ok = bool(np.array_equal(got, want))
        inv[layout] = {"ms": round(ms, 3), "Gkeys_s": round(n / (ms * 1e-3) / 1e9, 3),
                       **involution_roofline(layout, n, ms), "layout_sample": ok}
        log(f"[bench] {layout} involutions permute {ms:.3f} ms, sample check {ok}")
    out["involutions"] = inv
    return out


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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": layout_traffic(layout, n),
                     "kernel": f"whole tl_permute_{layout} call (all its launches)"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_involutions(args, dev):
    """The involution family at n = 2^log2n uint64: each permutation `steps`
    times from the sorted keys 0..n-1, its time per call and a full check of
    the last call's layout against the oracle's position functions.  Four
    copies of 2^27 int64 keys (1 GiB read, 1 GiB written) open the run to
    scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch
    import tl_amd
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    cal_a = torch.arange(1 << 27, device=dev, dtype=torch.int64)
    cal_b = torch.empty_like(cal_a)
    for _ in range(4):
        cal_b.copy_(cal_a)
    del cal_a, cal_b
    n = 1 << args.log2n
    A = torch.empty(n, dtype=torch.int64, device=dev)
    out = {}
    for layout, b in (("bst", 1), ("btree", args.b), ("mveb", 0)):
        times = []
        for _ in range(max(1, args.steps)):
            torch.arange(n, dtype=torch.int64, device=dev, out=A)
            times.append(tl_amd.permute_involutions(layout, A, b=b, kind="u64"))
        ranks = np.random.default_rng(7).integers(0, n, size=20_000, dtype=np.int64)
        o = ol.oracle()
        if layout == "mveb":
            pos = [o.or_mveb_pos(int(x), n) for x in ranks]
        else:
            pos = [o.or_btree_pos(int(x), n, b) for x in ranks]
        got = A[torch.tensor(pos, dtype=torch.int64, device=dev)].cpu().numpy()
        ms = float(np.mean(times))
        out[layout] = {"b": b, "ms": round(ms, 3), "runs_ms": [round(t, 3) for t in times],
                       **involution_roofline(layout, n, ms),
                       "layout_sample": bool(np.array_equal(got, ranks))}
        log(f"[bench] {layout} involutions {np.mean(times):.3f} ms")
    print(json.dumps({"workload": f"involution family, n=2^{args.log2n} uint64", "layouts": out}),
          flush=True)


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
    ap.add_argument("--log2leaves", type=int, default=16,
                    help="learned model leaves (log2), the headline model")
    ap.add_argument("--log2leaves-big", type=int, default=20,
                    help="a second, larger learned model reported beside it (0: none)")
    ap.add_argument("--skip-layouts", action="store_true",
                    help="do not time the BST, mvEB and classic vEB permutes beside the B-tree")
    ap.add_argument("--skip-batch", action="store_true",
                    help="leave the batch permutations out of the headline line")
    ap.add_argument("--cpu-log2n", type=int, default=28, help="CPU baseline sample (log2 keys)")
    ap.add_argument("--cpu-iters", type=int, default=10, help="CPU baseline runs")
    ap.add_argument("--workload", choices=["btree", "batch", "bst", "mveb", "veb", "involutions"],
                    default="btree")
    ap.add_argument("--batch-arrays", type=int, default=4096)
    ap.add_argument("--batch-perm-only", action="store_true",
                    help="time the batch's permutations alone (few other dispatches)")
    ap.add_argument("--batch-lds", type=int, default=0,
                    help="LDS budget of the batch's plans in bytes (0: the library default)")
    ap.add_argument("--batch-query-queue", type=int, default=0, choices=[0, 1],
                    help="answer the batch's queries through one work-queue launch per layout")
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
    if args.workload == "involutions":
        return run_involutions(args, dev)
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
    t_wall = time.perf_counter()
    total_ms, per_ms = event_steps(lambda: tl_amd.permute_btree(A, b, kind="u64", timed=False),
                                   args.steps, stream)
    wall_ms = (time.perf_counter() - t_wall) * 1e3
    tl_amd.sync()
    perm_ms = total_ms / args.steps
    perm_stats = timing_stats(n, per_ms)
    log(f"[bench] permute {perm_ms:.3f} ms/step (sd {perm_stats['sd_ms']:.3f}, host wall "
        f"{wall_ms / args.steps:.3f})")
    unperm_ms = time_unpermute(args, A, "btree", b, stream)
    log(f"[bench] unpermute {unperm_ms:.3f} ms/step")
    # the same layout from the allocation-free entry point, whose plan keeps
    # every stash within its fixed 64 MiB device array (so the scratch does not
    # grow with n); the permutation only depends on (n, b), so repeating it on
    # the same buffer repeats the same data movement
    noalloc = None
    if n <= tl_amd.TL_NOALLOC_MAX_KEYS:
        tl_amd.permute_btree_noalloc(A, b, kind="u64")
        na_total, na_per = event_steps(lambda: tl_amd.permute_btree_noalloc(A, b, kind="u64"),
                                       max(1, args.steps // 2), stream)
        tl_amd.sync()
        na_ms = na_total / max(1, args.steps // 2)
        noalloc = {"ms": round(na_ms, 4), "Gkeys_s": round(n / (na_ms * 1e-3) / 1e9, 3),
                   "scratch_bytes": tl_amd.TL_NOALLOC_SCRATCH_BYTES,
                   "algorithmic_frac": round(2.0 * n * 8 / (na_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] allocation-free permute {na_ms:.3f} ms/step")

    # ---- the other layouts' permutes on the same keys
    layouts = None if args.skip_layouts else other_layouts(args, A, keys, stream, rng)

    # ---- rebuild the true layout for the queries
    A.copy_(keys)
    tl_amd.permute_btree(A, b, kind="u64")
    ans = torch.empty(nq, dtype=torch.int64, device=dev)
    for _ in range(max(1, args.warmup)):
        tl_amd.query_btree(A, b, Q, ans, kind="u64", timed=False)
    q_total, q_per = event_steps(lambda: tl_amd.query_btree(A, b, Q, ans, kind="u64", timed=False),
                                 args.steps, stream)
    q_ms = q_total / args.steps
    plain_ok = bool(((ans >= 0) & (ans < n)).all().item()) and bool(
        torch.equal(A[ans.clamp(0, n - 1)], Q))
    log(f"[bench] plain query {q_ms:.3f} ms ({nq / q_ms / 1e3:.1f} Mq/s) ok={plain_ok}")

    # ---- the comparator: binary search of the sorted keys, same queries (its
    # own answer buffer: the plain answers are what the learned ones must equal)
    bans = torch.empty_like(ans)
    for _ in range(max(1, args.warmup)):
        tl_amd.query_binary(keys, Q, bans, kind="u64", timed=False)
    bs_total, _ = event_steps(lambda: tl_amd.query_binary(keys, Q, bans, kind="u64", timed=False),
                              args.steps, stream)
    bs_ms = bs_total / args.steps
    bs_ok = bool(((bans >= 0) & (bans < n)).all().item()) and bool(
        torch.equal(keys[bans.clamp(0, n - 1)], Q))
    del bans
    log(f"[bench] binary search {bs_ms:.3f} ms ({nq / bs_ms / 1e3:.1f} Mq/s) ok={bs_ok}")

    # ---- learned start node: the configuration's model (2^16 leaves) is the
    # headline; a larger leaf count is reported beside it
    learned, learned_big = None, None
    if not args.skip_learned:
        learned = tl_amd.bench_learned(A, keys, Q, ans, b, args.steps, args.warmup, n,
                                       num_leaves=1 << args.log2leaves)
        learned["vs_plain"] = round(q_ms / learned["ms"], 4)
        log(f"[bench] learned (2^{args.log2leaves} leaves) {learned}")
        if args.log2leaves_big and args.log2leaves_big != args.log2leaves:
            learned_big = tl_amd.bench_learned(A, keys, Q, ans, b, args.steps, args.warmup, n,
                                               num_leaves=1 << args.log2leaves_big)
            learned_big["vs_plain"] = round(q_ms / learned_big["ms"], 4)
            log(f"[bench] learned (2^{args.log2leaves_big} leaves) {learned_big}")

    # query roofline: algorithmic bytes of the DRAM-resident levels
    q_all_b, q_dram_b = btree_query_bytes(n, b, 8)

    # measured traffic per call (FETCH_SIZE + WRITE_SIZE, committed PMC summary;
    # L2 misses, which include lines the Infinity Cache serves) at this size
    q_traffic = None
    if n == 1 << 30 and b == 16 and nq == 1 << 27:
        q_traffic = measured_traffic_entry("btree_b16_n2p30_u64_query_traffic.txt")

    def qroof(ms, tag=None):
        ach = nq * q_dram_b / (ms * 1e-3) / 1e9
        out = {"bound": "hbm", "bytes_per_query": round(q_dram_b, 1),
               "bytes_per_query_all_levels": round(q_all_b, 1), "achieved": round(ach, 1),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
               "traffic": None}
        if q_traffic and tag and f"{tag}_read_GB" in q_traffic:
            t = (q_traffic[f"{tag}_read_GB"] + q_traffic[f"{tag}_write_GB"]) * 1e9
            out["traffic"] = round(t)
            out["traffic_GBs"] = round(t / (ms * 1e-3) / 1e9, 1)
            out["traffic_frac"] = round(t / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        return out

    # ---- the uneven batch's permutations (4096 arrays, one graph replay a step)
    batch = None
    if not args.skip_batch:
        del A, Q, ans
        torch.cuda.empty_cache()
        batch = measure_batch_permutes(args, dev)

    cpu = None
    if not args.no_cpu:
        ns = 1 << min(args.cpu_log2n, args.log2n)
        log(f"[bench] CPU baseline: first 2^{ns.bit_length() - 1} keys, {args.cpu_iters} runs")
        keys_np = keys[:ns].cpu().numpy().view(np.uint64)
        cpu = cpu_baseline(keys_np, b, 1 << 22, rng, iters=args.cpu_iters)
        log(f"[bench] CPU {cpu}")

    algo_bytes = 2.0 * n * 8
    achieved = algo_bytes / (perm_ms * 1e-3) / 1e9
    traffic = None
    if n == 1 << 30 and b == 16:
        traffic = measured_traffic("btree_b16_n2p30_u64_permute_traffic.txt")
    queries = {"plain_ms": round(q_ms, 4), "plain_Mq_s": round(nq / q_ms / 1e3, 2),
               "plain_timings": timing_stats(nq, q_per, 1e-6),
               "plain_roofline": qroof(q_ms, "plain"),
               "learned": learned,
               "binary_search": {"ms": round(bs_ms, 4), "Mq_s": round(nq / bs_ms / 1e3, 2),
                                 "answers_ok": bs_ok, "roofline": qroof(bs_ms, "binary")}}
    if learned is not None:
        queries["learned_roofline"] = qroof(
            learned["ms"], "learned16" if args.log2leaves == 16 else None)
    if learned_big is not None:
        queries["learned_more_leaves"] = learned_big
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
        "permute_timings": perm_stats,
        "unpermute": {"ms": round(unperm_ms, 4), "Gkeys_s": round(n / (unperm_ms * 1e-3) / 1e9, 3)},
        "permute_fixed_scratch": noalloc,
        "layouts": layouts,
        "queries": queries,
        "batch_permutes": batch,
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
            "value": round(cpu["permute"]["tp_mean"], 4), "unit": "Gkeys/s",
            "cores": cpu["cores"], "kind": cpu["kind"],
            "sample": (f"{'reference btree-cycles compiled from its sources (oracle/_ref)' if cpu['kind'] == 'reference' else 'oracle OpenMP restatement of btree-cycles (about 15x slower than the reference)'}"
                       f", n=2^{cpu['n'].bit_length() - 1} prefix of the same keys, b={b}, "
                       f"{cpu['permute']['runs']} runs with the sorted input restored before "
                       f"each; query-btree on 2^22 hits: {cpu['query']['tp_mean']:.2f} Mq/s"),
            "permute": cpu["permute"], "query": cpu["query"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
