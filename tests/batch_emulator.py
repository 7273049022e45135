"""Host replay of the uneven-batch engine's sub-task list (tl_batch_export).

The sub-tasks run one at a time in ticket order, which is one schedule the
persistent kernel may take.  Each sub-task does the slice of work the device
executor for its op type does for that sub-task number (the same index
arithmetic: runs, chunks, cycles, tiles, moves, query chunks), so a replay
that reproduces the oracle's layouts and answers checks the decomposition of
every op into sub-tasks.  Before a sub-task runs, the replay asserts what the
kernel waits for: the previous op of the same array has counted all its
sub-tasks, and every chained move the sub-task waits on carries a smaller
number.  Test infrastructure only."""
import ctypes

import numpy as np

import plan_emulator as pe

OP_QUERY = 100
REV_CHUNK = 256 * 8
COPY_CHUNK = 256 * 8
QCHUNK = 4096
THREADS = 256
CHAINED_KEYS = 256 * 32   # keys of one chained move (32 per thread)


class MoveSpec(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("P", ctypes.c_uint64), ("L", ctypes.c_uint64),
                ("k", ctypes.c_uint64), ("r", ctypes.c_uint64), ("l", ctypes.c_uint64),
                ("CH", ctypes.c_uint64), ("nchunk", ctypes.c_uint64),
                ("fuse", ctypes.c_uint32), ("rp", ctypes.c_uint64), ("lp", ctypes.c_uint64),
                ("soff", ctypes.c_uint64)]


class BOp(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("dtype", ctypes.c_uint32), ("layout", ctypes.c_uint32),
                ("A", ctypes.c_uint64), ("runs_at", ctypes.c_uint64),
                ("nruns", ctypes.c_uint64), ("nitems", ctypes.c_uint64),
                ("p", ctypes.c_uint64 * 8), ("ms", MoveSpec), ("ntasks", ctypes.c_uint64),
                ("prev", ctypes.c_int64), ("flags_at", ctypes.c_uint64),
                ("stash_at", ctypes.c_uint64), ("Q", ctypes.c_uint64), ("nq", ctypes.c_uint64),
                ("ans", ctypes.c_uint64), ("n", ctypes.c_uint64), ("b", ctypes.c_uint64),
                ("tbl_at", ctypes.c_uint64), ("idx_at", ctypes.c_uint64),
                ("ntables", ctypes.c_uint32), ("tsize", ctypes.c_uint32)]


class BTask(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("sub", ctypes.c_uint64)]


def export(lib, items, count, lds_bytes, wg_per_cu=4, num_cus=256, lds_max=163840):
    """(ops, tasks, blob, smem) of the batch the library would build"""
    from tl_amd import BatchConfig
    cfg = BatchConfig(lds_bytes, wg_per_cu, 0)
    sizes = (ctypes.c_uint64 * 5)()
    st = lib.tl_batch_export(items, count, ctypes.byref(cfg), num_cus, lds_max, None, 0, None, 0,
                             None, 0, sizes)
    if st != 0:
        raise RuntimeError(lib.tl_last_error().decode())
    nops, ntasks, nblob, smem, opsize = (int(v) for v in sizes)
    assert opsize == ctypes.sizeof(BOp), (opsize, ctypes.sizeof(BOp))
    ops = (BOp * nops)()
    tasks = (BTask * ntasks)()
    blob = np.zeros(max(nblob, 1), dtype=np.uint64)
    st = lib.tl_batch_export(items, count, ctypes.byref(cfg), num_cus, lds_max, ops,
                             ctypes.sizeof(ops), tasks, ctypes.sizeof(tasks),
                             blob.ctypes.data_as(ctypes.c_void_p), blob.size, sizes)
    if st != 0:
        raise RuntimeError(lib.tl_last_error().decode())
    return ops, tasks, blob.astype(np.int64), smem


def run_offset(blob, at, nruns, idx):
    lo, hi = 0, nruns
    while hi - lo > 1:
        mid = (lo + hi) >> 1
        if blob[at + 4 * mid + 3] <= idx:
            lo = mid
        else:
            hi = mid
    return int(blob[at + 4 * lo]) + (idx - int(blob[at + 4 * lo + 3])) * int(blob[at + 4 * lo + 1])


def rev(x, k):
    return int(format(x, f"0{k}b")[::-1], 2) if k > 0 else 0


def move_at(m, t):
    if m.kind == pe.MV_SHIFT_RIGHT:
        c = m.nchunk - 1 - t
        src = m.P + c * m.CH
        return src, src + m.k, min(m.CH, m.L - c * m.CH)
    if m.kind == pe.MV_SHIFT_LEFT:
        src = m.P + t * m.CH
        return src, src - m.k, min(m.CH, m.L - t * m.CH)
    if m.kind == pe.MV_ROW_SLIDE:
        j0 = (m.nchunk - 1 - t) * m.CH
        j1 = min(m.r, j0 + m.CH)
        return m.P + j0 * (m.l + 1), m.P + m.r + j0 * m.l, (j1 - j0) * (m.l + 1)
    if m.kind == pe.MV_ROW_SPREAD:
        j0 = t * m.CH
        j1 = min(m.r, j0 + m.CH)
        return m.P + m.r + j0 * m.l, m.P + j0 * (m.l + 1), (j1 - j0) * m.l
    fuse = getattr(m, "fuse", 0)
    if m.kind == pe.MV_SPARSE_SPREAD:
        if t < m.r * m.nchunk:
            i, c = divmod(t, m.nchunk)
            off = m.rp + c * m.CH // (m.lp + 1) * m.lp if fuse else c * m.CH
            return (m.P + m.r + m.k + i * m.l + off, m.P + i * (m.l + 1) + c * m.CH,
                    min(m.CH, m.l - c * m.CH))
        c = t - m.r * m.nchunk
        dst = m.P + m.r * (m.l + 1) + c * m.CH
        return dst + m.k, dst, min(m.CH, m.L - c * m.CH)
    ntail = (m.L + m.CH - 1) // m.CH
    if t < ntail:
        c = ntail - 1 - t
        src = m.P + m.r * (m.l + 1) + c * m.CH
        return src, src + m.k, min(m.CH, m.L - c * m.CH)
    t -= ntail
    i = m.r - 1 - t // m.nchunk
    c = m.nchunk - 1 - t % m.nchunk
    src = m.P + i * (m.l + 1) + c * m.CH
    off = m.rp + c * m.CH // (m.lp + 1) * m.lp if fuse else c * m.CH
    return src, m.P + m.r + m.k + i * m.l + off, min(m.CH, m.l - c * m.CH)


def awaited_moves(m, t, dst, ln):
    """move numbers a chained move waits for (the device's wait_deps)"""
    out = []
    if m.kind in (pe.MV_SHIFT_RIGHT, pe.MV_SHIFT_LEFT):
        if dst + ln <= m.P or dst >= m.P + m.L:
            return out
        lo = (dst - m.P) // m.CH if dst > m.P else 0
        hi = min((dst + ln - 1 - m.P) // m.CH, m.nchunk - 1)
        for c in range(lo, hi + 1):
            tt = m.nchunk - 1 - c if m.kind == pe.MV_SHIFT_RIGHT else c
            if tt < t:
                out.append(tt)
        return out
    if m.kind in (pe.MV_ROW_SPREAD, pe.MV_SPARSE_SPREAD):
        S0 = m.P + m.r + (m.k if m.kind == pe.MV_SPARSE_SPREAD else 0)
        span = ln // m.l * (m.l + 1) if m.kind == pe.MV_ROW_SPREAD else ln
        e = dst + span
        if e <= S0:
            return out
        a, z = max(dst, S0) - S0, e - 1 - S0
        if m.kind == pe.MV_ROW_SPREAD:
            lo, hi = a // (m.CH * m.l), z // (m.CH * m.l)
        else:
            leaf = m.r * m.l

            def num(x):
                if x < leaf:
                    i = x // m.l
                    w = x - i * m.l
                    if not getattr(m, "fuse", 0):
                        return i * m.nchunk + w // m.CH
                    rows = 0 if w < m.rp else (w - m.rp) // m.lp
                    return i * m.nchunk + rows // (m.CH // (m.lp + 1))
                return m.r * m.nchunk + (x - leaf) // m.CH
            lo, hi = num(a), num(z)
        return [u for u in range(lo, hi + 1) if u < t]
    if m.kind == pe.MV_ROW_SLIDE:
        dlen = ln // (m.l + 1) * m.l
        G = m.CH * (m.l + 1)
        for g in range((dst - m.P) // G, min((dst + dlen - 1 - m.P) // G, m.nchunk - 1) + 1):
            tt = m.nchunk - 1 - g
            if tt < t:
                out.append(tt)
        return out
    ntail = (m.L + m.CH - 1) // m.CH
    T0 = m.P + m.r * (m.l + 1)
    if m.L and dst + ln > T0 and dst < T0 + m.L:
        a, z = max(dst, T0) - T0, min(dst + ln, T0 + m.L) - T0
        for c in range(a // m.CH, (z - 1) // m.CH + 1):
            if ntail - 1 - c < t:
                out.append(ntail - 1 - c)
    if dst >= T0:
        return out
    L1 = m.l + 1
    rel = dst - m.P
    i_lo, i_hi = rel // L1, min((rel + ln - 1) // L1, m.r - 1)
    for i in range(i_lo, min(i_hi, m.r - 1) + 1):
        bs, be = i * L1, i * L1 + m.l
        ov_lo = rel - bs if rel > bs else 0
        ov_end = min(rel + ln, be)
        if ov_end <= bs + ov_lo:
            continue
        for c in range(ov_lo // m.CH, (ov_end - bs - 1) // m.CH + 1):
            tt = ntail + (m.r - 1 - i) * m.nchunk + (m.nchunk - 1 - c)
            if tt < t:
                out.append(tt)
    return out


class Replay:
    """arrays: {device address: numpy array (keys, viewed as stored)};
    queries: {address: numpy array}; answers: {address: numpy uint64 array}"""

    def __init__(self, ops, tasks, blob, arrays, queries, answers, search):
        self.ops, self.tasks, self.blob = ops, tasks, blob
        self.arrays, self.queries, self.answers = arrays, queries, answers
        self.search = search
        self.done = np.zeros(len(ops), dtype=np.int64)
        self.flags = {}
        self.stash = {}
        self.cost = {}

    def run(self):
        for t in range(len(self.tasks)):
            tk = self.tasks[t]
            op = self.ops[tk.op]
            if op.prev >= 0:
                assert self.done[op.prev] == self.ops[op.prev].ntasks, \
                    f"ticket {t}: op {tk.op} runs before op {op.prev} completed"
            self.step(tk.op, op, int(tk.sub))
            self.done[tk.op] += 1
        assert all(self.done[i] == self.ops[i].ntasks for i in range(len(self.ops)))

    def step(self, oid, op, sub):
        A = self.arrays.get(op.A)
        b = self.blob
        t = op.type
        p = [int(v) for v in op.p]
        key = t
        if t == OP_QUERY:
            lo, hi = sub * QCHUNK, min(op.nq, (sub + 1) * QCHUNK)
            Q = self.queries[op.Q][lo:hi]
            self.answers[op.ans][lo:hi] = self.search(op, A, Q)
            work = hi - lo
        elif t == pe.OP_LDS:
            n, kind, bb, rows, per = p[0], p[1], p[2], p[3], p[4]
            first = sub * per
            cnt = min(per, op.nitems - first)
            assert 0 < cnt <= 64 and cnt * n * op.width <= self.smem, (cnt, n)
            pos = pe.lds_perm(kind, n, bb, rows)
            for it in range(cnt):
                o = run_offset(b, op.runs_at, op.nruns, first + it)
                A[o + pos] = A[o:o + n].copy()
            work = cnt * n
        elif t == pe.OP_REVERSE:
            ln = p[0]
            half = ln // 2
            per = (half + REV_CHUNK - 1) // REV_CHUNK
            seg, chunk = sub // per, sub % per
            o = run_offset(b, op.runs_at, op.nruns, seg)
            lo, hi = chunk * REV_CHUNK, min(half, chunk * REV_CHUNK + REV_CHUNK)
            i = np.arange(lo, hi)
            x, y = A[o + i].copy(), A[o + ln - 1 - i].copy()
            A[o + i], A[o + ln - 1 - i] = y, x
            work = 2 * (hi - lo)
        elif t == pe.OP_BITREV:
            base, D, K = p[0] - (1 << 64) if p[0] >= 1 << 63 else p[0], p[1], p[2]
            R, MB, hs = 1 << K, D - 2 * K, D - K
            assert 2 * R * (R + 1) * op.width <= self.smem
            m = sub
            mr = rev(m, MB)
            work = 0
            if m <= mr:
                h, l = np.meshgrid(np.arange(R), np.arange(R), indexing="ij")
                ia = base + ((h << hs) | (m << K) | l)
                ib = base + ((h << hs) | (mr << K) | l)
                sA = np.zeros((R, R), dtype=A.dtype)
                ok = ia >= 0
                sA[ok] = A[ia[ok]]
                sB = A[ib].copy() if m != mr else None
                hr = np.vectorize(lambda v: rev(int(v), K))(l)
                lr = np.vectorize(lambda v: rev(int(v), K))(h)
                if m != mr:
                    A[ia] = sB[hr, lr]
                    A[ib] = sA[hr, lr]
                else:
                    A[ia[ok]] = sA[hr, lr][ok]
                work = 2 * R * R
        elif t == pe.OP_UNIT_MERGE:
            unit, cs_at, ncyc, cu_at = p[0], p[1], p[2], p[3]
            nchunks = (unit + THREADS - 1) // THREADS
            chunk = sub % nchunks
            cyc = (sub // nchunks) % ncyc
            parent = sub // (nchunks * ncyc)
            c0, c1 = chunk * THREADS, min(unit, (chunk + 1) * THREADS)
            s0, s1 = int(b[cs_at + cyc]), int(b[cs_at + cyc + 1])
            assert s1 > s0
            u = b[cu_at + s0: cu_at + s1]
            o = run_offset(b, op.runs_at, op.nruns, parent)
            cols = np.arange(c0, c1)
            vals = [A[o + int(x) * unit + cols].copy() for x in u]
            for k in range(len(u)):
                A[o + int(u[(k + 1) % len(u)]) * unit + cols] = vals[k]
            work = len(u) * (c1 - c0)
            key = (t, "cycle_len", len(u), "cols", c1 - c0)
        elif t == pe.OP_SPARSE_STREAM:
            r, l = p[0], p[1]
            assert r * op.width <= self.smem
            o = run_offset(b, op.runs_at, op.nruns, sub)
            reg = A[o:o + r * (l + 1)].reshape(r, l + 1).copy()
            A[o:o + r] = reg[:, l]
            A[o + r: o + r * (l + 1)] = reg[:, :l].reshape(-1)
            work = r * (l + 1)
        elif t == pe.OP_COPY:
            P, stride, count, to_stash, soff = p[0], p[1], p[2], p[3], p[4]
            lo, hi = sub * COPY_CHUNK, min(count, (sub + 1) * COPY_CHUNK)
            st = self.stash.setdefault(op.stash_at, {})
            for i in range(lo, hi):
                if to_stash:
                    st[soff + i] = A[P + i * stride]
                else:
                    A[P + i * stride] = st[soff + i]
            work = hi - lo
        elif t == pe.OP_CHAINED:
            m = op.ms
            src, dst, ln = move_at(m, sub)
            assert 0 < ln <= CHAINED_KEYS
            for tt in awaited_moves(m, sub, dst, ln):
                assert self.flags.get((oid, tt)), f"move {sub} waits on unpublished {tt}"
            v = A[src:src + ln].copy()
            self.flags[(oid, sub)] = True
            if m.kind == pe.MV_ROW_SLIDE:
                L1 = m.l + 1
                rows = v.reshape(-1, L1)
                st = self.stash.setdefault(op.stash_at, {})
                j0 = (src - m.P) // L1
                for k in range(rows.shape[0]):
                    st[j0 + k] = rows[k, m.l]
                A[dst:dst + rows.shape[0] * m.l] = rows[:, :m.l].reshape(-1)
            else:
                A[dst:dst + ln] = v
            work = ln
        else:
            raise ValueError(t)
        c = self.cost.setdefault(key, [0, 0, 0])
        c[0] += 1
        c[1] += work
        c[2] = max(c[2], work)
