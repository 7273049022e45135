"""The mvEB / classic vEB query keeps pos[] in 8 slots (slot(t) = slot(D[t])
+ 1, slot(0) = 0): a host model of that bookkeeping, run on the oracle's
tables, never needs more than 8 slots and gives the same positions as one
pos[] entry per depth on every root-to-leaf path it is run on."""
import numpy as np
import pytest

import oracle_lib as ol

SLOTS = 8


def slots_of(tables, base, dj):
    out = [0] * dj
    for t in range(1, dj):
        x, s = t, 0
        while x != 0:
            x = tables[base + x][2]
            s += 1
        out[t] = s
    return out


@pytest.mark.parametrize("classic", [False, True])
def test_slot_count_every_depth(classic):
    for d in range(1, 64):
        dd, nt, ts, tables, idx, tidx = ol.tableset((1 << d) - 1, classic)
        assert dd == d and nt == 1
        sl = slots_of(tables, 0, d)
        assert max(sl) <= SLOTS - 1, (d, max(sl))
        # every entry's frame root is shallower than it
        assert all(tables[t][2] < t for t in range(1, d))


def positions_full(tables, d, path):
    pos = [0] * (d + 1)
    pos[0] = 1
    bfs, out = 1, [0]
    for cd in range(1, d):
        bfs = 2 * bfs + path[cd - 1]
        L, R, D = tables[cd]
        pos[cd] = pos[D] + R + (bfs & R) * L
        out.append(pos[cd] - 1)
    return out


def positions_slots(tables, d, path, sl):
    slot = [0] * SLOTS
    slot[0] = 1
    bfs, out = 1, [0]
    for cd in range(1, d):
        bfs = 2 * bfs + path[cd - 1]
        L, R, D = tables[cd]
        p = slot[sl[cd] - 1] + R + (bfs & R) * L
        slot[sl[cd]] = p
        out.append(p - 1)
    return out


@pytest.mark.parametrize("classic", [False, True])
def test_slot_positions_match_full_pos(classic):
    rng = np.random.default_rng(3)
    for d in list(range(2, 24)) + [31, 40, 50, 63]:
        dd, nt, ts, tables, idx, tidx = ol.tableset((1 << d) - 1, classic)
        sl = slots_of(tables, 0, d)
        for _ in range(200):
            path = rng.integers(0, 2, size=d).tolist()
            assert positions_slots(tables, d, path, sl) == positions_full(tables, d, path), d


@pytest.mark.parametrize("classic", [False, True])
def test_slot_positions_nonperfect_chain(classic):
    """every structure of a non-perfect table set, walked on its own"""
    rng = np.random.default_rng(4)
    for n in [1000, 65536, 100003, 3_000_001, 300_000_000]:
        d, nt, ts, tables, idx, tidx = ol.tableset(n, classic)
        off = 0
        for j in range(nt):
            dj = int(np.floor(np.log2(n - off))) + 1
            sub = tables[tidx[j]: tidx[j] + dj]
            sl = slots_of(tables, tidx[j], dj)
            assert max(sl) <= SLOTS - 1
            for _ in range(50):
                path = rng.integers(0, 2, size=dj).tolist()
                assert positions_slots(sub, dj, path, sl) == positions_full(sub, dj, path)
            off = idx[j]
