"""Host replay of a serialized permutation plan (tl_plan_export).

Each launch type is replayed with its *specified* effect on the array (not the
kernel's code), using the oracle's constructions for the in-LDS sub-layouts.
Comparing the replayed array with the oracle checks the planner's
decomposition on the CPU; the GPU tests then check that every kernel has the
specified effect.  Test infrastructure only."""
import numpy as np

import oracle_lib as ol

(OP_LDS, OP_REVERSE, OP_BITREV, OP_UNIT_MERGE, OP_SPARSE_STREAM, OP_COPY, OP_CHAINED,
 OP_BITREV_LEVELS, OP_ROTATE, OP_HEADS, OP_PART_ROOTS) = range(11)
BITREV_PLAIN, BITREV_SHIFT, BITREV_UNFOLD, BITREV_HEAD, BITREV_ABSORB = range(5)
LDS_BTREE, LDS_MVEB, LDS_GATHER, LDS_LEVELREV, LDS_VEB = range(5)
MV_SHIFT_RIGHT, MV_SHIFT_LEFT, MV_SPARSE_SLIDE, MV_ROW_SLIDE, MV_ROW_SPREAD, MV_SPARSE_SPREAD = \
    range(6)


def decode(buf):
    nops, nblob = int(buf[0]), int(buf[1])
    ops = [buf[2 + 32 * k: 2 + 32 * (k + 1)].astype(np.int64) for k in range(nops)]
    blob = buf[2 + 32 * nops: 2 + 32 * nops + nblob].astype(np.int64)
    return ops, blob


def items(op, blob):
    nruns, at = int(op[1]), int(op[19])
    out = []
    for i in range(nruns):
        start, stride, count = (int(v) for v in blob[at + 4 * i: at + 4 * i + 3])
        out.extend(start + j * stride for j in range(count))
    return out


_cache = {}


def lds_perm(kind, n, b, rows):
    """slot of each rank for an in-LDS item"""
    key = (kind, n, b, rows)
    if key not in _cache:
        x = np.arange(n, dtype=np.int64)
        if kind == LDS_BTREE:
            ranks = ol.inorder_ranks(n, b).astype(np.int64)
            pos = np.empty(n, dtype=np.int64)
            pos[ranks] = x
        elif kind in (LDS_MVEB, LDS_VEB):
            ranks = (ol.mveb_struct_ranks(n) if kind == LDS_MVEB
                     else ol.classic_struct_ranks(n)).astype(np.int64)
            pos = np.empty(n, dtype=np.int64)
            pos[ranks] = x
        elif kind == LDS_GATHER:
            row, c = x // (b + 1), x % (b + 1)
            pos = np.where(row >= rows, x, np.where(c == b, row, rows + row * b + c))
        else:
            y = x + 1
            t = np.floor(np.log2(y)).astype(np.int64)
            low = y - (np.int64(1) << t)
            rev = np.array([int(format(int(v), f"0{int(tt)}b")[::-1], 2) if tt > 0 else 0
                            for v, tt in zip(low, t)], dtype=np.int64)
            pos = (np.int64(1) << t) + rev - 1
        _cache[key] = pos
    return _cache[key]


def rev_bits(y, D):
    out = np.zeros_like(y)
    for i in range(D):
        out |= ((y >> i) & 1) << (D - 1 - i)
    return out


def _level_slot(lev, y, K, t0):
    """stash slot of the row end at slot y of level lev: the level's rows
    q = (h << (lev - 2K)) | m (m the tile number) at
    ((2^lev - 2^t0) >> K) + m * 2^K + h"""
    q = y >> K
    mb = lev - 2 * K
    return (((1 << lev) - (1 << t0)) >> K) + (((q & ((1 << mb) - 1)) << K) | (q >> mb))


def _tree_row_ends(D, K, t0):
    """row ends (slot % 2^K = 2^K - 1) at slots >= 2^t0 of a tree of slots
    [0, 2^D), and their stash slots"""
    ends, slots = [], []
    for lev in range(t0, D):
        y = np.arange(1 << lev, dtype=np.int64)
        e = y[(y & ((1 << K) - 1)) == (1 << K) - 1]
        ends.append((1 << lev) + e)
        slots.append(_level_slot(lev, e, K, t0))
    return np.concatenate(ends), np.concatenate(slots)


def _fuse_blocks(blocks, fuse, rp, lp):
    """leaf blocks (rows of rp rows of (lp keys, 1 root) and lp keys) ->
    their rp + 1 lp-key blocks, each in vEB order"""
    r = blocks.shape[0]
    pos = lds_perm(LDS_MVEB if fuse == 1 else LDS_VEB, lp, 0, 0)
    rows = blocks[:, :rp * (lp + 1)].reshape(r, rp, lp + 1)[:, :, :lp]
    leaf = np.concatenate([rows, blocks[:, rp * (lp + 1):].reshape(r, 1, lp)], axis=1)
    out = np.empty_like(leaf)
    out[:, :, pos] = leaf
    return out.reshape(r, (rp + 1) * lp)


def _unfuse_blocks(final, fuse, rp, lp, stash, soff):
    """the inverse of _fuse_blocks on the final blocks (rp roots' places,
    then rp + 1 lp-key blocks in vEB order), with the row roots from the
    stash"""
    r = final.shape[0]
    pos = lds_perm(LDS_MVEB if fuse == 1 else LDS_VEB, lp, 0, 0)
    leaf = final[:, rp:].reshape(r, rp + 1, lp)[:, :, pos]
    out = np.empty((r, rp * (lp + 1) + lp), dtype=final.dtype)
    rows = out[:, :rp * (lp + 1)].reshape(r, rp, lp + 1)
    rows[:, :, :lp] = leaf[:, :rp]
    rows[:, :, lp] = stash[soff: soff + r * rp].reshape(r, rp)
    out[:, rp * (lp + 1):] = leaf[:, rp]
    return out


def replay(buf, A, stash=None):
    ops, blob = decode(buf)
    stash = np.zeros(0, dtype=A.dtype) if stash is None else stash

    def put(off, vals):
        nonlocal stash
        if stash.size < off + vals.size:
            grown = np.zeros(off + vals.size, dtype=A.dtype)
            grown[:stash.size] = stash
            stash = grown
        stash[off:off + vals.size] = vals
    for op in ops:
        t = int(op[0])
        p = [int(v) for v in op[3:11]]
        if t == OP_LDS:
            n, kind, b, rows, inv, shift = p[0], p[1], p[2], p[3], p[5], p[6]
            pos = lds_perm(kind, n, b, rows)
            for o in items(op, blob):
                if inv:   # written one place right when shifted
                    sub = A[o:o + n].copy()
                    A[o + shift:o + shift + n] = sub[pos]
                else:     # read from one place right when shifted
                    sub = A[o + shift:o + shift + n].copy()
                    A[o + pos] = sub
        elif t == OP_REVERSE:
            ln = p[0]
            for o in items(op, blob):
                A[o:o + ln] = A[o:o + ln][::-1].copy()
        elif t == OP_BITREV and p[3] != BITREV_PLAIN:
            # base 0 (a BST kept one place right); absorb reads the row ends
            # at slots >= 2^t0 from the stash
            D, K, t0 = p[1], p[2], p[4]
            y = np.arange(1 << D, dtype=np.int64)
            V = A[:1 << D].copy()
            if p[3] == BITREV_ABSORB:
                ends, slot = _tree_row_ends(D, K, t0)
                V[ends] = stash[slot]
            A[y] = V[rev_bits(y, D)]
        elif t == OP_HEADS:
            # the first 2^(t-K) slots of every level t0 .. d-1, level after level
            t0, d, K = p[0], p[1], p[2]
            put(0, np.concatenate([A[(1 << lev):(1 << lev) + (1 << (lev - K))]
                                   for lev in range(t0, d)]))
        elif t == OP_BITREV_LEVELS and p[3] == BITREV_SHIFT:
            # every level t reversed and written one slot left: A[2^t - 1 + y]
            # gets reversed slot y; the last slot of the level's range takes
            # the next level's first slot from the heads (not the last level)
            t0, d, K = p[0], p[1], p[2]
            for lev in range(t0, d):
                y = np.arange(1 << lev, dtype=np.int64)
                X = A[(1 << lev) + y][rev_bits(y, lev)].copy()
                if lev == t0:
                    A[(1 << lev) - 1] = X[0]
                A[(1 << lev) + y[:-1]] = X[1:]
                if lev + 1 < d:
                    A[(1 << (lev + 1)) - 1] = stash[((1 << (lev + 1)) - (1 << t0)) >> K]
        elif t == OP_BITREV_LEVELS and p[3] == BITREV_UNFOLD:
            # level t read from A[2^t - 1 ..], reversed, written to A[2^t ..];
            # the row ends go to the stash
            t0, d, K = p[0], p[1], p[2]
            for lev in range(t0, d):
                y = np.arange(1 << lev, dtype=np.int64)
                ends = y[(y & ((1 << K) - 1)) == (1 << K) - 1]
                slot = _level_slot(lev, ends, K, t0)
                W = A[(1 << lev) - 1 + y][rev_bits(y, lev)].copy()
                keep = (y & ((1 << K) - 1)) != (1 << K) - 1
                A[(1 << lev) + y[keep]] = W[keep]
                st = stash[:slot.max() + 1].copy() if stash.size > slot.max() else \
                    np.concatenate([stash, np.zeros(slot.max() + 1 - stash.size, A.dtype)])
                st[slot] = W[ends]
                put(0, st)
        elif t == OP_BITREV:
            base, D = np.int64(p[0]), p[1]
            y = np.arange(1 << D, dtype=np.int64)
            idx = base + y
            par = base + rev_bits(y, D)
            ok = idx >= 0
            old = A.copy()
            A[idx[ok]] = old[par[ok]]
        elif t == OP_BITREV_LEVELS:
            for lev in range(p[0], p[1]):
                y = np.arange(1 << lev, dtype=np.int64)
                idx = np.int64((1 << lev) - 1) + y
                par = np.int64((1 << lev) - 1) + rev_bits(y, lev)
                old = A.copy()
                A[idx] = old[par]
        elif t == OP_ROTATE:
            P, L, k = p[0], p[1], p[2]
            A[P:P + L] = np.roll(A[P:P + L].copy(), k)
        elif t == OP_UNIT_MERGE:
            unit, cs_at, ncyc, cu_at, inv = p[0], p[1], p[2], p[3], p[4]
            cs = blob[cs_at: cs_at + ncyc + 1]
            for o in items(op, blob):
                old = None
                for c in range(ncyc):
                    us = blob[cu_at + cs[c]: cu_at + cs[c + 1]]
                    vals = [A[o + u * unit: o + (u + 1) * unit].copy() for u in us]
                    for k in range(len(us)):
                        dst = us[(k - 1) % len(us)] if inv else us[(k + 1) % len(us)]
                        A[o + dst * unit: o + (dst + 1) * unit] = vals[k]
        elif t == OP_SPARSE_STREAM:
            r, l, inv, fuse = p[0], p[1], p[2], p[3]
            kind = LDS_MVEB if fuse == 1 else LDS_VEB
            for o in items(op, blob):
                if inv:
                    if fuse:   # undo the kids' layouts first (r + 1 leaf blocks)
                        sub = A[o:o + r].copy()
                        A[o:o + r] = sub[lds_perm(kind, r, 0, 0)]
                        blk = A[o + r: o + r + (r + 1) * l].reshape(r + 1, l)
                        blk[:, :] = blk[:, lds_perm(kind, l, 0, 0)].copy()
                    roots = A[o:o + r].copy()
                    leaves = A[o + r: o + r * (l + 1)].reshape(r, l).copy()
                    reg = A[o:o + r * (l + 1)].reshape(r, l + 1)
                    reg[:, :l] = leaves
                    reg[:, l] = roots
                    continue
                reg = A[o:o + r * (l + 1)].reshape(r, l + 1).copy()
                A[o:o + r] = reg[:, l]
                A[o + r: o + r * (l + 1)] = reg[:, :l].reshape(-1)
                if fuse:   # the roots and all r + 1 leaf blocks laid out in place
                    sub = A[o:o + r].copy()
                    A[o + lds_perm(kind, r, 0, 0)] = sub
                    blk = A[o + r: o + r + (r + 1) * l].reshape(r + 1, l)
                    new = np.empty_like(blk)
                    new[:, lds_perm(kind, l, 0, 0)] = blk
                    blk[:, :] = new
        elif t == OP_COPY:
            P, stride, count, to_stash, soff = p[0], p[1], p[2], p[3], p[4]
            if to_stash:
                put(soff, A[P: P + count * stride: stride][:count].copy())
            else:
                A[P: P + count * stride: stride][:count] = stash[soff:soff + count]
        elif t == OP_CHAINED:
            kind, P, L, k, r, l = (int(v) for v in op[11:17])
            if kind == MV_SHIFT_RIGHT:
                A[P + k: P + k + L] = A[P:P + L].copy()
            elif kind == MV_SHIFT_LEFT:
                A[P - k: P - k + L] = A[P:P + L].copy()
            elif kind == MV_ROW_SPREAD:
                # r rows of l keys back into rows of (l keys, separator from
                # the stash)
                keys = A[P + r: P + r + r * l].reshape(r, l).copy()
                reg = A[P:P + r * (l + 1)].reshape(r, l + 1)
                reg[:, :l] = keys
                reg[:, l] = stash[:r]
            elif kind == MV_SPARSE_SPREAD:
                T0 = P + r * (l + 1)
                fuse, rp, lp, soff = (int(v) for v in op[20:24])
                leaves = A[P + r + k: P + r + k + r * l].reshape(r, l).copy()
                if fuse:   # each leaf block's own gather undone, roots from the stash
                    leaves = _unfuse_blocks(leaves, fuse, rp, lp, stash, soff)
                tail = A[T0 + k: T0 + k + L].copy()
                A[P:P + r * (l + 1)].reshape(r, l + 1)[:, :l] = leaves
                A[T0:T0 + L] = tail
            elif kind == MV_ROW_SLIDE:
                # rows of (l keys, 1 separator): separators to the stash
                reg = A[P:P + r * (l + 1)].reshape(r, l + 1).copy()
                put(0, reg[:, l].copy())
                A[P + r: P + r * (l + 1)] = reg[:, :l].reshape(-1)
            else:
                # rows of (l leaf keys, 1 root) then a tail of L keys, all slid
                # right past the r roots and k extra roots (both stashed before)
                fuse, rp, lp, soff = (int(v) for v in op[20:24])
                reg = A[P:P + r * (l + 1)].reshape(r, l + 1).copy()
                tail = A[P + r * (l + 1): P + r * (l + 1) + L].copy()
                if fuse:
                    # each leaf block's own gather rides along: its row roots
                    # to the stash, its lp-key blocks to their vEB slots after
                    # its first rp places, which the slide leaves as they were
                    blocks = reg[:, :l]
                    roots = blocks[:, :rp * (lp + 1)].reshape(r, rp, lp + 1)[:, :, lp]
                    put(soff, roots.reshape(-1).copy())
                    dst = A[P + r + k: P + r + k + r * l].reshape(r, l)
                    dst[:, rp:] = _fuse_blocks(blocks, fuse, rp, lp)
                else:
                    A[P + r + k: P + r + k + r * l] = reg[:, :l].reshape(-1)
                A[P + r * (l + 1) + k: P + r * (l + 1) + k + L] = tail
        elif t == OP_PART_ROOTS:
            # the rp roots of count blocks, stash <-> their vEB slots at the
            # block fronts
            base, stride, count, rp, soff, fuse, inv = p[0], p[1], p[2], p[3], p[4], p[5], p[6]
            pos = lds_perm(LDS_MVEB if fuse == 1 else LDS_VEB, rp, 0, 0)
            for i in range(count):
                if inv:
                    put(soff + i * rp, A[base + i * stride + pos].copy())
                else:
                    A[base + i * stride + pos] = stash[soff + i * rp: soff + (i + 1) * rp]
        else:
            raise ValueError(f"unknown op {t}")
    return A
