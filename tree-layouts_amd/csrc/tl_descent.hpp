// tl_descent.hpp -- B-tree node descent by a group of G lanes (G a power of
// two >= b, at most 64), shared by the plain and the learned-start queries.
//
// Node at key index i holds keys A[i .. i+b-1]; lane j of the group loads key
// j, two ballots give (keys < q) and (keys == q) for the whole group, the
// popcount of the first is the branch and a set bit of the second at that
// position is the hit (gpu/src/query-btree.cuh:32-58; child of branch js is
// (i+1)(b+1) + js*b - 1).  Falling off the tree answers n.
//
// descent2 walks two independent queries per group in lock step, so every
// lane has two node loads in flight.
//
// descent_wide is the form for nodes of a whole number of 16-byte words
// (b * sizeof(K) a multiple of 16, 16-byte aligned array): each lane of a
// group of G >= b / KPL lanes holds KPL = 16 / sizeof(K) consecutive keys of
// the node from one 16-byte load, so a wave serves 64 / G queries per slot
// (8 for b = 16 and 8-byte keys instead of 4) with fewer instructions per
// query and level.  Nodes are numbered: node k holds keys k*b .. k*b+b-1 and
// child js of node k is node k(b+1)+1+js, the reference's key index
// (i+1)(b+1)+js*b-1 divided by b.  The hit test uses that a node's keys
// ascend: when some key of the node equals q, the first key that is not
// smaller than q (position js, the reference's stopping point) does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

namespace tl {

template <int G>
__device__ inline uint64_t group_bits(uint64_t ballot, uint32_t g) {
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1);
    return (ballot >> (g * G)) & gmask;
}

// one level of one query: branch, or record the hit; falling off ends it
__device__ inline void descent_step(uint64_t mlt, uint64_t meq, uint32_t b, uint64_t n,
                                    uint64_t &i, bool &active, uint64_t &res, bool &found) {
    const uint32_t js = (uint32_t)__popcll(mlt);
    if ((meq >> js) & 1ull) {
        res = i + js;
        found = true;
        active = false;
    } else {
        i = (i + 1) * (b + 1) + (uint64_t)js * b - 1;
        if (i >= n) active = false;
    }
}

template <typename K, int G>
__device__ inline void descent2(const K *__restrict__ A, uint64_t n, uint32_t b, uint32_t g,
                                uint32_t j, K q0, K q1, uint64_t i0, uint64_t i1, bool a0, bool a1,
                                uint64_t &r0, uint64_t &r1, bool &f0, bool &f1) {
    r0 = n;
    r1 = n;
    f0 = false;
    f1 = false;
    while (__any(a0 || a1)) {
        K k0{}, k1{};
        const bool l0 = a0 && j < b && i0 + j < n;
        const bool l1 = a1 && j < b && i1 + j < n;
        if (l0) k0 = A[i0 + j];
        if (l1) k1 = A[i1 + j];
        const uint64_t lt0 = group_bits<G>(__ballot(l0 && k0 < q0), g);
        const uint64_t eq0 = group_bits<G>(__ballot(l0 && k0 == q0), g);
        const uint64_t lt1 = group_bits<G>(__ballot(l1 && k1 < q1), g);
        const uint64_t eq1 = group_bits<G>(__ballot(l1 && k1 == q1), g);
        if (a0) descent_step(lt0, eq0, b, n, i0, a0, r0, f0);
        if (a1) descent_step(lt1, eq1, b, n, i1, a1, r1, f1);
    }
}

// KPL = 16 / sizeof(K) keys from one 16-byte aligned load
template <typename K>
__device__ inline void load16(const K *p, K (&k)[16 / sizeof(K)]) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    if constexpr (sizeof(K) == 8) {
        k[0] = (K)(((uint64_t)v.y << 32) | v.x);
        k[1] = (K)(((uint64_t)v.w << 32) | v.z);
    } else {
        k[0] = (K)v.x;
        k[1] = (K)v.y;
        k[2] = (K)v.z;
        k[3] = (K)v.w;
    }
}

// QN queries q[k] of a group from nodes node[k] (active where act[k]); the
// answer (slot, or n when absent) in res[k], found[k] on a hit.  gm is the
// group's bits of a wave ballot, j the lane's index in the group, nodes the
// number of nodes ceil(n / b) (a node number at or past it is off the tree).
// IDX holds node numbers: 32 bits when nodes * (b + 1) < 2^32.
template <typename K, int QN, typename IDX>
__device__ inline void descent_wide(const K *__restrict__ A, uint64_t n, uint32_t b, IDX nodes,
                                    uint64_t gm, uint32_t j, const K (&q)[QN], IDX (&node)[QN],
                                    bool (&act)[QN], uint64_t (&res)[QN], bool (&found)[QN]) {
    constexpr uint32_t KPL = 16 / sizeof(K);
    const uint32_t jk = j * KPL;   // node position of the lane's first key
    bool any = false;
#pragma unroll
    for (int k = 0; k < QN; k++) {
        res[k] = n;
        found[k] = false;
        any |= act[k];
    }
    while (__any(any)) {
        K key[QN][KPL];
        uint64_t pos[QN];
#pragma unroll
        for (int k = 0; k < QN; k++) {
            pos[k] = (uint64_t)node[k] * b + jk;
#pragma unroll
            for (uint32_t t = 0; t < KPL; t++) key[k][t] = K{};
            if (act[k]) {
                if (pos[k] + KPL <= n) {
                    load16<K>(A + pos[k], key[k]);
                } else {   // the array's last node: slots past n count as +inf
#pragma unroll
                    for (uint32_t t = 0; t < KPL; t++)
                        if (pos[k] + t < n) key[k][t] = A[pos[k] + t];
                }
            }
        }
        any = false;
#pragma unroll
        for (int k = 0; k < QN; k++) {
            uint32_t js = 0;
            uint64_t eq = 0;
#pragma unroll
            for (uint32_t t = 0; t < KPL; t++) {
                const bool v = act[k] && jk + t < b && pos[k] + t < n;
                js += (uint32_t)__popcll(__ballot(v && key[k][t] < q[k]) & gm);
                eq |= __ballot(v && key[k][t] == q[k]);
            }
            if (act[k]) {
                if (eq & gm) {
                    res[k] = (uint64_t)node[k] * b + js;
                    found[k] = true;
                    act[k] = false;
                } else {
                    node[k] = node[k] * (IDX)(b + 1) + (IDX)(1 + js);
                    if (node[k] >= nodes) act[k] = false;
                }
            }
            any |= act[k];
        }
    }
}

// group mask of lane `lane` for groups of G lanes
template <int G>
__device__ inline uint64_t group_mask(uint32_t lane) {
    return (G == 64) ? ~0ull : (((1ull << G) - 1) << ((lane / G) * G));
}

// the wide form applies: whole 16-byte words per node, aligned array, b <= 64
// ($TL_BTREE_WIDE=0 selects the one-key-per-lane group form instead)
inline bool wide_ok(const void *A, uint64_t b, uint32_t w) {
    const char *e = getenv("TL_BTREE_WIDE");
    if (e && atoi(e) == 0) return false;
    return b >= 1 && b <= 64 && (b * w) % 16 == 0 && ((uintptr_t)A % 16) == 0;
}

// lanes per group of the wide form: the power of two >= b * w / 16
inline uint32_t wide_group(uint64_t b, uint32_t w) {
    const uint64_t need = b * w / 16;
    uint32_t g = 1;
    while (g < need) g <<= 1;
    return g;
}

// node numbers fit 32 bits: every child number stays below 2^32
// ($TL_BTREE_IDX64=1 keeps 64-bit node numbers at every size)
inline bool wide_narrow(uint64_t n, uint64_t b) {
    const char *e = getenv("TL_BTREE_IDX64");
    if (e && atoi(e) == 1) return false;
    const uint64_t nodes = (n + b - 1) / b;
    return nodes < (1ull << 32) / (b + 1);
}

// the learned query in one launch (start nodes and descents) where the wide
// form applies with 32-bit node numbers and n < 2^32; $TL_LEARNED_FUSED=0
// keeps the two passes (start slots through the answer array)
inline bool learned_fused() {
    const char *e = getenv("TL_LEARNED_FUSED");
    return !(e && atoi(e) == 0);
}

// queries per group of the wide descents: $TL_BTREE_QPG (2 or 4) when set,
// else def.  Measured at n = 2^30, b = 16, 8-byte keys: the root descent is
// fastest with 4 (12.7 ms for 2^27 hits, 13.6 ms with 2), the learned descent
// from deep start nodes, whose loads mostly miss the caches, with 2 (10.8 ms,
// 11.8 ms with 4).
// a caller on this host thread may fix it (the batch engine's configuration)
inline int &btree_qpg_fixed() {
    static thread_local int v = 0;
    return v;
}
inline int btree_qpg(int def) {
    if (btree_qpg_fixed() == 2 || btree_qpg_fixed() == 4) return btree_qpg_fixed();
    const char *e = getenv("TL_BTREE_QPG");
    if (e && (atoi(e) == 2 || atoi(e) == 4)) return atoi(e);
    return def;
}

}   // namespace tl
