// tl_query.hip -- batched exact-match queries over the three layouts, and the
// device-side mvEB table-set build, for gfx950.
//
// Answers are the slot of the query key or n for an absent key.  The node
// visiting order is the reference's (cpu/src/query-bst.h:30-43,
// cpu/src/query-btree.h:33-54, gpu/src/query-mveb.cuh:43-68,173-213), so
// every answer equals the reference's for every query it defines and misses
// are answered instead of reading past the array or the table.
#include <algorithm>

#include "tl_internal.hpp"
#include "tl_descent.hpp"
#include "tl_queue.hpp"

namespace tl {

// ---------------------------------------------------------------------------
// BST: one lane per query.  The top levels are staged in LDS, 16 KiB of keys
// (11 levels of 8-byte keys, 12 of 4-byte keys); each lane keeps BST_QPL
// queries in flight so that independent dependent descents overlap their
// memory latency.
template <typename K>
constexpr uint32_t BST_TOP = sizeof(K) == 4 ? 4095 : 2047;
constexpr int BST_QPL = 4;   // queries in flight per lane

// queries first, first + stride, ... below end of one array: the body of the
// one-array kernel (a grid stride) and of the batch's query queue (a chunk
// of one workgroup)
template <typename K>
__device__ inline void bst_body(const K *__restrict__ A, uint64_t n, const K *top, uint32_t ntop,
                                const K *__restrict__ Q, uint64_t *__restrict__ ans,
                                uint64_t first, uint64_t end, uint64_t stride) {
    // BST_QPL independent queries per lane, `stride` apart for coalescing
    for (uint64_t q0 = first; q0 < end; q0 += stride * BST_QPL) {
        K qv[BST_QPL];
        uint64_t iv[BST_QPL], rv[BST_QPL];
#pragma unroll
        for (int k = 0; k < BST_QPL; k++) {
            const uint64_t qi = q0 + k * stride;
            qv[k] = qi < end ? Q[qi] : K{};
            iv[k] = qi < end ? 0 : n;
            rv[k] = n;
        }
        bool any = true;
        while (any) {
            K cv[BST_QPL];
#pragma unroll
            for (int k = 0; k < BST_QPL; k++)
                if (iv[k] < n) cv[k] = iv[k] < ntop ? top[iv[k]] : A[iv[k]];
            any = false;
#pragma unroll
            for (int k = 0; k < BST_QPL; k++) {
                if (iv[k] < n) {
                    if (qv[k] == cv[k]) {
                        rv[k] = iv[k];
                        iv[k] = n;
                    } else {
                        iv[k] = 2 * iv[k] + 1 + (qv[k] > cv[k]);
                    }
                }
                any |= iv[k] < n;
            }
        }
#pragma unroll
        for (int k = 0; k < BST_QPL; k++) {
            const uint64_t qi = q0 + k * stride;
            if (qi < end) ans[qi] = rv[k];
        }
    }
}

template <typename K>
__device__ inline uint32_t bst_stage(const K *__restrict__ A, uint64_t n, K *top) {
    const uint32_t ntop = (uint32_t)min((uint64_t)BST_TOP<K>, n);
    for (uint32_t i = threadIdx.x; i < ntop; i += blockDim.x) top[i] = A[i];
    __syncthreads();
    return ntop;
}

template <typename K>
__global__ __launch_bounds__(256) void k_query_bst(const K *__restrict__ A, uint64_t n,
                                                   const K *__restrict__ Q, uint64_t nq,
                                                   uint64_t *__restrict__ ans) {
    __shared__ K top[BST_TOP<K>];
    const uint32_t ntop = bst_stage<K>(A, n, top);
    bst_body<K>(A, n, top, ntop, Q, ans, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nq,
                (uint64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// B-tree, group form: G lanes (a power of two >= b, <= 64) serve one query.
// Lane j loads key j of the node (one coalesced b-key load per level); two
// wave ballots give "key < q" and "key == q"; the popcount of the group's
// "<" bits is the reference's branch index j (first key that is not smaller)
// and the "==" bit at that index is a hit.  Slots at or past n count as +inf.
template <typename K, int G>
__global__ __launch_bounds__(256) void k_query_btree_group(const K *__restrict__ A, uint64_t n,
                                                           uint32_t b, const K *__restrict__ Q,
                                                           uint64_t nq, uint64_t *__restrict__ ans) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g = lane / G, j = lane % G;
    const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / G);
    const uint64_t trips = (nq + 2 * groups - 1) / (2 * groups);   // uniform trip count
    uint64_t qa = (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    for (uint64_t it = 0; it < trips; it++, qa += 2 * groups) {
        const uint64_t qb = qa + groups;
        const bool va = qa < nq, vb = qb < nq;
        K q0{}, q1{};
        if (va) q0 = Q[qa];
        if (vb) q1 = Q[qb];
        uint64_t r0, r1;
        bool f0, f1;
        descent2<K, G>(A, n, b, g, j, q0, q1, 0, 0, va, vb, r0, r1, f0, f1);
        if (j == 0) {
            if (va) ans[qa] = r0;
            if (vb) ans[qb] = r1;
        }
    }
}

// B-tree, wide form (descent_wide): groups of G lanes holding 16 bytes of
// the node each, QN queries per group.  The body answers queries lo + first,
// + groups, ... below hi; every lane of a workgroup makes the same number of
// trips (the descent's ballots need the whole wave).
template <typename K, int G, int QN, typename IDX>
__device__ inline void btree_wide_body(const K *__restrict__ A, uint64_t n, uint32_t b,
                                       const K *__restrict__ Q, uint64_t *__restrict__ ans,
                                       uint64_t lo, uint64_t hi, uint64_t first, uint64_t groups) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t j = lane % G;
    const uint64_t gm = group_mask<G>(lane);
    const IDX nodes = (IDX)((n + b - 1) / b);
    const uint64_t trips = (hi - lo + QN * groups - 1) / (QN * groups);   // uniform trip count
    uint64_t q0 = lo + first;
    for (uint64_t it = 0; it < trips; it++, q0 += QN * groups) {
        K q[QN];
        IDX node[QN];
        bool act[QN], found[QN];
        uint64_t res[QN];
#pragma unroll
        for (int k = 0; k < QN; k++) {
            const uint64_t qi = q0 + k * groups;
            act[k] = qi < hi;
            q[k] = act[k] ? Q[qi] : K{};
            node[k] = 0;
        }
        descent_wide<K, QN, IDX>(A, n, b, nodes, gm, j, q, node, act, res, found);
        if (j == 0) {
#pragma unroll
            for (int k = 0; k < QN; k++) {
                const uint64_t qi = q0 + k * groups;
                if (qi < hi) ans[qi] = res[k];
            }
        }
    }
}

template <typename K, int G, int QN, typename IDX>
__global__ __launch_bounds__(256) void k_query_btree_wide(const K *__restrict__ A, uint64_t n,
                                                          uint32_t b, const K *__restrict__ Q,
                                                          uint64_t nq, uint64_t *__restrict__ ans) {
    const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / G);
    btree_wide_body<K, G, QN, IDX>(A, n, b, Q, ans, 0, nq,
                                   (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G,
                                   groups);
}

// B-tree, lane form (b == 1 or b > 64): the reference's linear node scan
template <typename K>
__global__ __launch_bounds__(256) void k_query_btree_lane(const K *__restrict__ A, uint64_t n,
                                                          uint64_t b, const K *__restrict__ Q,
                                                          uint64_t nq, uint64_t *__restrict__ ans) {
    for (uint64_t qi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; qi < nq;
         qi += (uint64_t)gridDim.x * blockDim.x) {
        const K q = Q[qi];
        uint64_t i = 0, res = n;
        while (i < n) {
            uint64_t j = 0;
            bool hit = false;
            for (; j < b && i + j < n; j++) {
                const K cur = A[i + j];
                if (q == cur) {
                    hit = true;
                    break;
                }
                if (q < cur) break;
            }
            if (hit) {
                res = i + j;
                break;
            }
            i = (i + 1) * (b + 1) + j * b - 1;
        }
        ans[qi] = res;
    }
}

// ---------------------------------------------------------------------------
// Binary search of the sorted array, the layouts' comparator.  For distinct
// keys it answers as the reference's loop (cpu/src/binary-search.h:30-50,
// gpu/src/binary-search.cuh:28-48): the key's index on a hit, the index of its
// predecessor on a miss, and n where no key is <= q (there the reference's
// unsigned `right` wraps and its loop does not end).
//
// Branch-free form: the window [base, base + len) always holds the last key
// <= q (or base = 0 when there is none); a step probes A[base + len/2] and
// keeps the upper or lower part.  The len sequence does not depend on q, so
// every query takes the same ceil(log2 n) steps and a wave never diverges.
// The probes of the first BIN_TOPL steps form a complete binary tree, staged
// per workgroup in LDS in heap order (16 KiB: 11 steps of 8-byte keys, 12 of
// 4-byte keys); the answer is base when A[base] <= q.
template <typename K>
constexpr uint32_t BIN_TOPL = sizeof(K) == 4 ? 12 : 11;
constexpr int BIN_QPL = 4;          // queries in flight per lane

template <typename K>
__global__ __launch_bounds__(256) void k_query_binary(const K *__restrict__ A, uint64_t n,
                                                      const K *__restrict__ Q, uint64_t nq,
                                                      uint64_t *__restrict__ ans) {
    __shared__ K top[1u << BIN_TOPL<K>];   // top[0] = A[0], node i >= 1 in heap order
    __shared__ uint64_t half[64];
    __shared__ uint32_t steps_s;
    if (threadIdx.x == 0) {
        uint64_t len = n;
        uint32_t t = 0;
        while (len > 1) {
            half[t++] = len >> 1;
            len -= len >> 1;
        }
        steps_s = t;
    }
    __syncthreads();
    const uint32_t S = steps_s, T = min(S, BIN_TOPL<K>);
    if (n) {
        // node i at level t probes base(i) + half[t], base(i) = the halves
        // taken on the path from the root (the bits of i below its top bit)
        for (uint32_t i = threadIdx.x; i < (1u << T); i += blockDim.x) {
            if (i == 0) {
                top[0] = A[0];
                continue;
            }
            const uint32_t t = 31 - __clz(i);
            uint64_t base = 0;
            for (uint32_t u = 0; u < t; u++)
                if ((i >> (t - 1 - u)) & 1) base += half[u];
            top[i] = A[base + half[t]];
        }
    }
    __syncthreads();
    const uint64_t all = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < nq;
         q0 += all * BIN_QPL) {
        K qv[BIN_QPL], vb[BIN_QPL];
        uint64_t base[BIN_QPL];
        uint32_t node[BIN_QPL];
#pragma unroll
        for (int k = 0; k < BIN_QPL; k++) {
            const uint64_t qi = q0 + k * all;
            qv[k] = qi < nq ? Q[qi] : K{};
            base[k] = 0;
            node[k] = 1;
            vb[k] = n ? top[0] : K{};
        }
        for (uint32_t t = 0; t < T; t++) {
            const uint64_t h = half[t];
#pragma unroll
            for (int k = 0; k < BIN_QPL; k++) {
                const K v = top[node[k]];
                const bool go = v <= qv[k];
                node[k] = 2 * node[k] + go;
                if (go) {
                    base[k] += h;
                    vb[k] = v;
                }
            }
        }
        for (uint32_t t = T; t < S; t++) {
            const uint64_t h = half[t];
            K v[BIN_QPL];
#pragma unroll
            for (int k = 0; k < BIN_QPL; k++) v[k] = A[base[k] + h];
#pragma unroll
            for (int k = 0; k < BIN_QPL; k++)
                if (v[k] <= qv[k]) {
                    base[k] += h;
                    vb[k] = v[k];
                }
        }
#pragma unroll
        for (int k = 0; k < BIN_QPL; k++) {
            const uint64_t qi = q0 + k * all;
            if (qi < nq) ans[qi] = (n && vb[k] <= qv[k]) ? base[k] : n;
        }
    }
}

// ---------------------------------------------------------------------------
// mvEB: one lane per query with the Brodal-Fagerberg-Jacob recurrence
// pos[t] = pos[D[t]] + R[t] + (bfs & R[t]) * L[t].  Running out of depth d_j in
// structure j means the key is absent (the reference reads table[d] there).
//
// pos[] is kept in VEB_SLOTS slots instead of one entry per depth: pos[x] is
// only read while the descent is inside the recursion frame rooted at depth
// x, frames nest, and the frame rooted at t sits one level below the frame
// rooted at D[t].  So slot(t) = slot(D[t]) + 1 (slot(0) = 0, the root) gives
// every live pos[] value its own slot: entry t reads slot(t) - 1 and writes
// slot(t).  The recursion depth of a split of d <= 64 levels is at most 7
// (checked for both splits on the CPU), so 8 slots suffice.  The slots live
// in LDS as [slot][query k][lane], small enough for QPL queries per lane to
// be in flight at full occupancy; the first VEB_TOP keys (the top frames of
// the layout) are staged in LDS as well.
constexpr uint32_t VEB_SLOTS = 8;
constexpr uint32_t VEB_TOP = 2047;

struct TblEnt {
    uint64_t L, R;
    uint32_t D, slot;
};

// LDS layout of k_query_mveb: tables, idx, tidx, the top keys, the slots
struct VebSmem {
    size_t top, pos, bytes;
};
__host__ __device__ inline VebSmem veb_smem(uint32_t table_size, uint32_t ntop, size_t kw, size_t pw,
                                            int qpl) {
    auto al = [](size_t x) { return (x + 15) & ~(size_t)15; };
    VebSmem m;
    m.top = al(table_size * sizeof(TblEnt) + 64 * 8 + 64 * 4);
    m.pos = al(m.top + ntop * kw);
    m.bytes = m.pos + (size_t)VEB_SLOTS * qpl * 256 * pw;
    return m;
}

// the LDS copies a workgroup works from: tables with every entry's pos[] slot,
// idx, table_idx, the top keys, and the pos[] slots of its lanes
template <typename K, typename PT>
struct VebView {
    TblEnt *tb;
    uint64_t *s_idx;
    uint32_t *s_tidx;
    K *top;
    PT *pos;
};

template <typename K, typename PT, int QPL>
__device__ inline VebView<K, PT> veb_stage(unsigned char *smem_raw, const K *__restrict__ A,
                                           uint64_t n, const tl_veb_table *__restrict__ tables,
                                           uint32_t table_size, const uint64_t *__restrict__ idx,
                                           const uint32_t *__restrict__ tidx, uint32_t ntables,
                                           uint32_t ntop) {
    const VebSmem lay = veb_smem(table_size, ntop, sizeof(K), sizeof(PT), QPL);
    VebView<K, PT> v;
    v.tb = reinterpret_cast<TblEnt *>(smem_raw);
    v.s_idx = reinterpret_cast<uint64_t *>(v.tb + table_size);
    v.s_tidx = reinterpret_cast<uint32_t *>(v.s_idx + 64);
    v.top = reinterpret_cast<K *>(smem_raw + lay.top);
    v.pos = reinterpret_cast<PT *>(smem_raw + lay.pos);
    TblEnt *tb = v.tb;
    for (uint32_t i = threadIdx.x; i < table_size; i += blockDim.x) {
        tb[i].L = tables[i].L;
        tb[i].R = tables[i].R;
        tb[i].D = tables[i].D;
    }
    for (uint32_t i = threadIdx.x; i < ntables; i += blockDim.x) {
        v.s_idx[i] = idx ? idx[i] : n;
        v.s_tidx[i] = tidx ? tidx[i] : 0;
    }
    for (uint32_t i = threadIdx.x; i < ntop; i += blockDim.x) v.top[i] = A[i];
    __syncthreads();
    // slot of every entry: the length of its chain t -> D[t] -> .. -> 0
    for (uint32_t i = threadIdx.x; i < table_size; i += blockDim.x) {
        uint32_t base = 0;
        for (uint32_t j = 0; j < ntables; j++)
            if (v.s_tidx[j] <= i) base = max(base, v.s_tidx[j]);
        uint32_t x = i - base, sl = 0;
        for (uint32_t it = 0; x != 0 && it < 64; it++) {
            const uint32_t nx = tb[base + x].D;
            if (nx >= x) break;   // not a valid chain (a table entry never used)
            x = nx;
            sl++;
        }
        tb[i].slot = min(max(sl, 1u), VEB_SLOTS - 1);   // entry 0 (the root) is never read
    }
    __syncthreads();
    return v;
}

// queries first, first + stride, ... below end
template <typename K, typename PT, int QPL>
__device__ inline void veb_body(const VebView<K, PT> &v, const K *__restrict__ A, uint64_t n,
                                uint32_t ntables, uint32_t ntop, const K *__restrict__ Q,
                                uint64_t *__restrict__ ans, uint64_t first, uint64_t end,
                                uint64_t stride) {
    const TblEnt *tb = v.tb;
    const uint64_t *s_idx = v.s_idx;
    const uint32_t *s_tidx = v.s_tidx;
    const K *top = v.top;
    PT *pos = v.pos;
    const uint32_t W = blockDim.x, me = threadIdx.x;
    auto P = [&](uint32_t slot, int k) -> PT & { return pos[(slot * QPL + k) * W + me]; };
    for (uint64_t q0 = first; q0 < end; q0 += stride * QPL) {
        K q[QPL];
        uint64_t res[QPL], offset[QPL], endj[QPL], bfs[QPL], index[QPL];
        uint32_t j[QPL], cd[QPL], dj[QPL], tbase[QPL];
        bool act[QPL];
        // structure j of query k: from its root, or the chain's end
        auto enter = [&](int k) {
            while (act[k]) {
                if (j[k] >= ntables) {
                    act[k] = false;
                    break;
                }
                endj[k] = s_idx[j[k]];
                if (offset[k] < endj[k]) {
                    tbase[k] = s_tidx[j[k]];
                    dj[k] = floor_log2_64(n - offset[k]) + 1;
                    bfs[k] = 1;
                    index[k] = 0;
                    cd[k] = 0;
                    P(0, k) = 1;
                    break;
                }
                offset[k] = endj[k];
                j[k]++;
            }
        };
#pragma unroll
        for (int k = 0; k < QPL; k++) {
            const uint64_t qi = q0 + k * stride;
            act[k] = qi < end;
            q[k] = act[k] ? Q[qi] : K{};
            res[k] = n;
            offset[k] = 0;
            j[k] = 0;
            enter(k);
        }
        bool any = true;
        while (__any(any)) {
            K cur[QPL];
#pragma unroll
            for (int k = 0; k < QPL; k++) {
                cur[k] = K{};
                if (act[k]) {
                    const uint64_t at = offset[k] + index[k];
                    cur[k] = at < ntop ? top[at] : A[at];
                }
            }
            any = false;
#pragma unroll
            for (int k = 0; k < QPL; k++) {
                if (act[k]) {
                    if (q[k] == cur[k]) {
                        res[k] = offset[k] + index[k];
                        act[k] = false;
                    } else {
                        bfs[k] = 2 * bfs[k] + (q[k] > cur[k]);
                        cd[k]++;
                        if (cd[k] >= dj[k]) {
                            act[k] = false;
                        } else {
                            const TblEnt &e = tb[tbase[k] + cd[k]];
                            const uint64_t p =
                                (uint64_t)P(e.slot - 1, k) + e.R + (bfs[k] & e.R) * e.L;
                            P(e.slot, k) = (PT)p;
                            index[k] = p - 1;
                            if (index[k] + offset[k] >= endj[k]) {
                                offset[k] = endj[k];
                                j[k]++;
                                enter(k);
                            }
                        }
                    }
                }
                any |= act[k];
            }
        }
#pragma unroll
        for (int k = 0; k < QPL; k++) {
            const uint64_t qi = q0 + k * stride;
            if (qi < end) ans[qi] = res[k];
        }
    }
}

template <typename K, typename PT, int QPL>
__global__ __launch_bounds__(256) void k_query_mveb(const K *__restrict__ A, uint64_t n,
                                                    const tl_veb_table *__restrict__ tables,
                                                    uint32_t table_size,
                                                    const uint64_t *__restrict__ idx,
                                                    const uint32_t *__restrict__ tidx,
                                                    uint32_t ntables, uint32_t ntop,
                                                    const K *__restrict__ Q, uint64_t nq,
                                                    uint64_t *__restrict__ ans) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const VebView<K, PT> v =
        veb_stage<K, PT, QPL>(smem_raw, A, n, tables, table_size, idx, tidx, ntables, ntop);
    veb_body<K, PT, QPL>(v, A, n, ntables, ntop, Q, ans,
                         (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, nq,
                         (uint64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// the uneven batch's query queue: one persistent launch answers every array's
// queries.  The work is cut into chunks of `chunk` queries of one array;
// workgroups take chunk numbers from an atomic ticket until none are left (no
// workgroup ever waits for another), find the chunk's array by a binary search
// of the chunks' prefix sums, stage what that array's descent keeps in LDS
// and answer the chunk with the body of the one-array kernel.
constexpr int QQ_QN = 4;    // B-tree descents per lane group
constexpr int QQ_QPL = 2;   // mvEB queries per lane

__device__ inline uint32_t qq_find(const uint64_t *__restrict__ pre, uint32_t nd, uint64_t t) {
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <typename K, int LAYOUT>
__device__ inline void qq_chunk(const QueueItem &it, uint64_t lo, uint64_t hi,
                                unsigned char *smem) {
    const K *A = (const K *)it.A;
    const K *Q = (const K *)it.Q;
    uint64_t *ans = (uint64_t *)it.ans;
    if constexpr (LAYOUT == 0) {
        K *top = reinterpret_cast<K *>(smem);
        const uint32_t ntop = bst_stage<K>(A, it.n, top);
        bst_body<K>(A, it.n, top, ntop, Q, ans, lo + threadIdx.x, hi, blockDim.x);
    } else if constexpr (LAYOUT == 1) {
        switch (it.G) {
#define TL_QW(GG)                                                                                  \
    case GG:                                                                                       \
        btree_wide_body<K, GG, QQ_QN, uint32_t>(A, it.n, it.b, Q, ans, lo, hi, threadIdx.x / GG,   \
                                                blockDim.x / GG);                                  \
        break;
            TL_QW(1) TL_QW(2) TL_QW(4) TL_QW(8) TL_QW(16)
#undef TL_QW
        default: break;
        }
    } else {
        const VebView<K, uint32_t> v = veb_stage<K, uint32_t, QQ_QPL>(
            smem, A, it.n, (const tl_veb_table *)it.tables, it.table_size,
            (const uint64_t *)it.idx, (const uint32_t *)it.tidx, it.ntables, it.ntop);
        veb_body<K, uint32_t, QQ_QPL>(v, A, it.n, it.ntables, it.ntop, Q, ans, lo + threadIdx.x,
                                      hi, blockDim.x);
    }
}

// one queue per layout, so that each kernel holds only its own descent's
// registers
template <int LAYOUT>
__global__ __launch_bounds__(256) void k_query_queue(const QueueItem *__restrict__ items,
                                                     const uint64_t *__restrict__ pre,
                                                     uint32_t nitems, uint64_t chunks,
                                                     uint64_t chunk,
                                                     unsigned long long *ticket) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ uint64_t s_t;
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(ticket, 1ull);
        __syncthreads();
        const uint64_t t = s_t;
        __syncthreads();
        if (t >= chunks) return;
        const uint32_t i = qq_find(pre, nitems, t);
        const QueueItem it = items[i];
        const uint64_t lo = (t - pre[i]) * chunk;
        const uint64_t hi = min(it.nq, lo + chunk);
        switch (it.dtype) {
        case TL_U32: qq_chunk<uint32_t, LAYOUT>(it, lo, hi, smem_raw); break;
        case TL_U64: qq_chunk<uint64_t, LAYOUT>(it, lo, hi, smem_raw); break;
        case TL_I32: qq_chunk<int32_t, LAYOUT>(it, lo, hi, smem_raw); break;
        default: qq_chunk<int64_t, LAYOUT>(it, lo, hi, smem_raw); break;
        }
        __syncthreads();   // the next chunk restages the LDS
    }
}

bool queue_takes(const tl_batch_item &it) {
    if (it.nq == 0 || it.n == 0 || it.n >= (1ull << 32)) return false;
    if (it.layout != 1) return it.layout <= 2;
    const uint32_t w = (it.dtype == TL_U32 || it.dtype == TL_I32) ? 4 : 8;
    return wide_ok(it.A, it.b, w) && wide_narrow(it.n, it.b) && wide_group(it.b, w) <= 16;
}

QueueItem queue_item(const tl_batch_item &it, uint64_t tables, uint64_t idx, uint64_t tidx,
                     uint32_t ntables, uint32_t table_size, size_t *lds) {
    const uint32_t w = (it.dtype == TL_U32 || it.dtype == TL_I32) ? 4 : 8;
    QueueItem q{};
    q.A = (uint64_t)it.A;
    q.Q = (uint64_t)it.Q;
    q.ans = (uint64_t)it.ans;
    q.n = it.n;
    q.nq = it.nq;
    q.dtype = (uint32_t)it.dtype;
    q.layout = it.layout;
    q.b = (uint32_t)it.b;
    size_t need = 0;
    if (it.layout == 0) {
        need = (size_t)(w == 4 ? BST_TOP<uint32_t> : BST_TOP<uint64_t>) * w;
    } else if (it.layout == 1) {
        q.G = wide_group(it.b, w);
    } else {
        q.tables = tables;
        q.idx = idx;
        q.tidx = tidx;
        q.ntables = ntables;
        q.table_size = table_size;
        q.ntop = (uint32_t)std::min<uint64_t>(it.n, w == 4 ? 4095u : VEB_TOP);
        need = veb_smem(table_size, q.ntop, w, 4, QQ_QPL).bytes;
        if (need > 64 * 1024) {   // a long table chain: no staged top keys
            q.ntop = 0;
            need = veb_smem(table_size, 0, w, 4, QQ_QPL).bytes;
        }
    }
    *lds = need;
    return q;
}

uint64_t queue_chunk() {
    const char *e = getenv("TL_QQ_CHUNK");
    const long long v = e ? atoll(e) : 0;
    return (v >= 1024 && v <= (1 << 20)) ? (uint64_t)v : 4096u;
}

tl_status queue_launch(Ctx *c, uint32_t layout, const QueueItem *items, const uint64_t *pre,
                       uint32_t nitems, uint64_t chunks, uint64_t chunk,
                       unsigned long long *ticket, size_t lds, hipStream_t s) {
    if (!nitems || !chunks) return TL_OK;
    const char *e = getenv("TL_QQ_GRID");
    const int per = e ? atoi(e) : 16;
    const uint64_t wg = (per >= 1 && per <= 64) ? (uint64_t)per : 16u;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(chunks, (uint64_t)c->num_cus * wg);
    if (layout == 0)
        hipLaunchKernelGGL(k_query_queue<0>, dim3(grid), dim3(256), lds, s, items, pre, nitems,
                           chunks, chunk, ticket);
    else if (layout == 1)
        hipLaunchKernelGGL(k_query_queue<1>, dim3(grid), dim3(256), lds, s, items, pre, nitems,
                           chunks, chunk, ticket);
    else
        hipLaunchKernelGGL(k_query_queue<2>, dim3(grid), dim3(256), lds, s, items, pre, nitems,
                           chunks, chunk, ticket);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// mvEB query shape: $TL_VEB_QPL (1, 2, 4) queries per lane and $TL_VEB_TOP
// (0: no staged top keys) when set; a caller on this host thread may fix the
// queries per lane (the batch engine's configuration)
int &veb_qpl_fixed() {
    static thread_local int v = 0;
    return v;
}
inline void veb_shape(int &qpl, uint32_t &topmax) {
    const char *e = getenv("TL_VEB_QPL");
    const char *f = getenv("TL_VEB_TOP");
    qpl = 2;
    topmax = VEB_TOP;
    if (e && (atoi(e) == 1 || atoi(e) == 2 || atoi(e) == 4)) qpl = atoi(e);
    const int fx = veb_qpl_fixed();
    if (fx == 1 || fx == 2 || fx == 4) qpl = fx;
    if (f && atoi(f) == 0) topmax = 0;
}

// ---------------------------------------------------------------------------
// device table-set build: one lane walks the chain of incomplete bottom trees
// (gpu/test/query-mveb-fixedN.cu:42-165) and runs buildTable
// (gpu/src/query-mveb.cu:25-47) for every structure with an explicit stack.
__device__ void build_table_dev(tl_veb_table *table, uint64_t n, uint32_t d, bool classic) {
    struct Fr {
        uint32_t base, d, root_depth;
        uint64_t n;
    };
    Fr st[130];
    int top = 0;
    st[0] = {0, d, 0, n};
    while (top >= 0) {
        const Fr f = st[top--];
        if (f.n <= 1) continue;
        uint32_t rd, ld;
        veb_split(f.d, rd, ld, classic);
        const uint64_t root_n = (1ull << rd) - 1, leaf_n = (1ull << ld) - 1;
        const uint32_t i = f.d - ld;
        table[f.base + i].L = leaf_n;
        table[f.base + i].R = root_n;
        table[f.base + i].D = f.root_depth;
        st[++top] = {f.base + i, ld, i + f.root_depth, leaf_n};
        st[++top] = {f.base, rd, f.root_depth, root_n};
    }
}

// classic = true builds the classic vEB's table set with the classic split
// (cpu/src/query-veb.cpp:25-41, cpu/test/query-veb-fixedN.cpp:37-158)
__global__ void k_tableset(uint64_t n, tl_veb_table *tables, uint64_t *idx, uint32_t *tidx,
                           uint32_t *meta, bool classic) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t d = floor_log2_64(n) + 1;
    uint64_t sizes[64];
    uint32_t depths[64];
    uint32_t nt = 1, tsize = d;
    sizes[0] = n;
    depths[0] = d;
    if (n != (1ull << d) - 1) {
        uint64_t tn = n;
        uint32_t td = d;
        while (td >= 2 && nt < 64) {
            uint32_t rd, ld;
            veb_split(td, rd, ld, classic);
            const uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
            const uint64_t nf = (tn - r) / l, inc = tn - r - nf * l;
            if (inc == 0) break;
            const uint32_t incd = floor_log2_64(inc) + 1;
            sizes[nt] = inc;
            depths[nt] = incd;
            nt++;
            tsize += incd;
            tn = inc;
            td = incd;
            if (inc == (1ull << incd) - 1) break;
        }
    }
    for (uint32_t i = 0; i < tsize; i++) tables[i] = tl_veb_table{0, 0, 0};
    uint32_t off = 0;
    for (uint32_t j = 0; j < nt; j++) {
        tidx[j] = off;
        build_table_dev(tables + off, sizes[j], depths[j], classic);
        off += depths[j];
        idx[j] = (j + 1 < nt) ? n - sizes[j + 1] : n;
    }
    meta[0] = d;
    meta[1] = nt;
    meta[2] = tsize;
}

// ---------------------------------------------------------------------------
// host side

static uint32_t tableset_size(uint64_t n, bool classic) {
    const uint32_t d = floor_log2_64(n) + 1;
    uint32_t tsize = d;
    if (n == (1ull << d) - 1) return tsize;
    uint64_t tn = n;
    uint32_t td = d, nt = 1;
    while (td >= 2 && nt < 64) {
        uint32_t rd, ld;
        veb_split(td, rd, ld, classic);
        const uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
        const uint64_t nf = (tn - r) / l, inc = tn - r - nf * l;
        if (inc == 0) break;
        const uint32_t incd = floor_log2_64(inc) + 1;
        nt++;
        tsize += incd;
        tn = inc;
        td = incd;
        if (inc == (1ull << incd) - 1) break;
    }
    return tsize;
}

static tl_status qcheck(const void *A, uint64_t n, const void *q, uint64_t nq, const void *ans) {
    if (nq && (!q || !ans)) {
        tl::set_error("null query or answer buffer");
        return TL_EINVAL;
    }
    if (n && !A) {
        tl::set_error("null array");
        return TL_EINVAL;
    }
    return TL_OK;
}

// grid of a grid-stride query kernel: one workgroup per per_block queries, at
// most per_cu workgroups per CU.  Kernels that stage keys in LDS per
// workgroup keep 16; the wide B-tree descents stage nothing and run faster
// with many more workgroups than stay resident (256 per CU: 2^27 queries at
// 2^30 keys, plain 12.7 -> 11.4 ms, learned 9.5 -> 8.6 ms against 16).
static uint32_t grid_for(Ctx *c, uint64_t work, uint32_t per_block, uint32_t per_cu = 16) {
    const uint64_t want = (work + per_block - 1) / per_block;
    const uint64_t cap = (uint64_t)c->num_cus * per_cu;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

template <typename K>
static void launch_bst(Ctx *c, const void *A, uint64_t n, const void *q, uint64_t nq, uint64_t *ans,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_query_bst<K>, dim3(grid_for(c, nq, 256 * BST_QPL)), dim3(256), 0, s,
                       (const K *)A, n, (const K *)q, nq, ans);
}

template <typename K>
static void launch_binary(Ctx *c, const void *A, uint64_t n, const void *q, uint64_t nq,
                          uint64_t *ans, hipStream_t s) {
    hipLaunchKernelGGL(k_query_binary<K>, dim3(grid_for(c, nq, 256 * BIN_QPL)), dim3(256), 0, s,
                       (const K *)A, n, (const K *)q, nq, ans);
}

template <typename K>
static void launch_btree(Ctx *c, const void *A, uint64_t n, uint64_t b, const void *q, uint64_t nq,
                         uint64_t *ans, hipStream_t s) {
    if (wide_ok(A, b, sizeof(K))) {
        const uint32_t G = wide_group(b, sizeof(K));
        const int QN = btree_qpg(4);
        const bool narrow = wide_narrow(n, b);
        const uint32_t grid = grid_for(c, nq, QN * 256 / G, 256);
#define TL_WL(GG, QQ, IX)                                                                          \
    hipLaunchKernelGGL((k_query_btree_wide<K, GG, QQ, IX>), dim3(grid), dim3(256), 0, s,           \
                       (const K *)A, n, (uint32_t)b, (const K *)q, nq, ans)
#define TL_W(GG)                                                                                   \
    case GG:                                                                                       \
        if (QN == 4) {                                                                             \
            if (narrow) TL_WL(GG, 4, uint32_t);                                                    \
            else TL_WL(GG, 4, uint64_t);                                                           \
        } else {                                                                                   \
            if (narrow) TL_WL(GG, 2, uint32_t);                                                    \
            else TL_WL(GG, 2, uint64_t);                                                           \
        }                                                                                          \
        break;
        switch (G) { TL_W(1) TL_W(2) TL_W(4) TL_W(8) TL_W(16) TL_W(32) }
#undef TL_W
#undef TL_WL
    } else if (b >= 2 && b <= 64) {
        const uint32_t G = b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : b <= 16 ? 16 : b <= 32 ? 32 : 64;
        const uint32_t grid = grid_for(c, nq, 2 * 256 / G);
        switch (G) {
#define TL_G(GG)                                                                                   \
    case GG:                                                                                       \
        hipLaunchKernelGGL((k_query_btree_group<K, GG>), dim3(grid), dim3(256), 0, s, (const K *)A, \
                           n, (uint32_t)b, (const K *)q, nq, ans);                                 \
        break;
            TL_G(2) TL_G(4) TL_G(8) TL_G(16) TL_G(32) TL_G(64)
#undef TL_G
        }
    } else {
        hipLaunchKernelGGL(k_query_btree_lane<K>, dim3(grid_for(c, nq, 256)), dim3(256), 0, s,
                           (const K *)A, n, b, (const K *)q, nq, ans);
    }
}

template <typename K>
static void launch_mveb(Ctx *c, const void *A, uint64_t n, const tl_veb_table *tables,
                        uint32_t table_size, const uint64_t *idx, const uint32_t *tidx, uint32_t nt,
                        const void *q, uint64_t nq, uint64_t *ans, hipStream_t s) {
    int qpl;
    uint32_t topmax;
    veb_shape(qpl, topmax);
    const bool narrow = n < (1ull << 32);
    const uint32_t grid = grid_for(c, nq, 256 * qpl);
    // 16 KiB of staged keys (2047 8-byte or 4095 4-byte keys), when every
    // workgroup has at least 1024 queries to spend them on
    uint32_t ntop = (uint32_t)std::min<uint64_t>(n, topmax ? (sizeof(K) == 4 ? 4095u : topmax) : 0u);
    if (nq < (uint64_t)grid * 1024) ntop = 0;
    size_t bytes = veb_smem(table_size, ntop, sizeof(K), narrow ? 4 : 8, qpl).bytes;
    if (bytes > 64 * 1024) {   // a long table chain: no staged top keys
        ntop = 0;
        bytes = veb_smem(table_size, ntop, sizeof(K), narrow ? 4 : 8, qpl).bytes;
    }
#define TL_V(QQ)                                                                                   \
    if (qpl == QQ) {                                                                               \
        if (narrow)                                                                                \
            hipLaunchKernelGGL((k_query_mveb<K, uint32_t, QQ>), dim3(grid), dim3(256), bytes, s,   \
                               (const K *)A, n, tables, table_size, idx, tidx, nt, ntop,           \
                               (const K *)q, nq, ans);                                             \
        else                                                                                       \
            hipLaunchKernelGGL((k_query_mveb<K, uint64_t, QQ>), dim3(grid), dim3(256), bytes, s,   \
                               (const K *)A, n, tables, table_size, idx, tidx, nt, ntop,           \
                               (const K *)q, nq, ans);                                             \
        return;                                                                                    \
    }
    TL_V(1) TL_V(2) TL_V(4)
#undef TL_V
}

#define TL_DISPATCH(dt, FN, ...)                                                                 \
    switch (dt) {                                                                                \
    case TL_U32: FN<uint32_t>(__VA_ARGS__); break;                                               \
    case TL_U64: FN<uint64_t>(__VA_ARGS__); break;                                               \
    case TL_I32: FN<int32_t>(__VA_ARGS__); break;                                                \
    case TL_I64: FN<int64_t>(__VA_ARGS__); break;                                                \
    default: set_error("unknown dtype"); return TL_EINVAL;                                       \
    }

}   // namespace tl

using namespace tl;

extern "C" tl_status tl_query_bst(const void *dev_A, uint64_t n, tl_dtype dtype, const void *dev_q,
                                  uint64_t nq, uint64_t *dev_ans, void *stream, float *ms) {
    TL_TRY(qcheck(dev_A, n, dev_q, nq, dev_ans));
    if (nq == 0) return TL_OK;
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    TL_DISPATCH(dtype, launch_bst, c, dev_A, n, dev_q, nq, dev_ans, s);
    TL_LAUNCH_CHECK();
    return tm.stop();
}

extern "C" tl_status tl_query_binary(const void *dev_A, uint64_t n, tl_dtype dtype,
                                     const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                                     void *stream, float *ms) {
    TL_TRY(qcheck(dev_A, n, dev_q, nq, dev_ans));
    if (nq == 0) return TL_OK;
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    TL_DISPATCH(dtype, launch_binary, c, dev_A, n, dev_q, nq, dev_ans, s);
    TL_LAUNCH_CHECK();
    return tm.stop();
}

extern "C" tl_status tl_query_btree(const void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                    const void *dev_q, uint64_t nq, uint64_t *dev_ans, void *stream,
                                    float *ms) {
    TL_TRY(qcheck(dev_A, n, dev_q, nq, dev_ans));
    if (b == 0) {
        set_error("b must be >= 1");
        return TL_EINVAL;
    }
    if (nq == 0) return TL_OK;
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    TL_DISPATCH(dtype, launch_btree, c, dev_A, n, b, dev_q, nq, dev_ans, s);
    TL_LAUNCH_CHECK();
    return tm.stop();
}

extern "C" tl_status tl_query_mveb(const void *dev_A, const tl_veb_table *dev_table, uint64_t n,
                                   uint32_t d, tl_dtype dtype, const void *dev_q, uint64_t nq,
                                   uint64_t *dev_ans, void *stream, float *ms) {
    TL_TRY(qcheck(dev_A, n, dev_q, nq, dev_ans));
    if (nq == 0) return TL_OK;
    if (n == 0 || d == 0 || d > 63 || !dev_table || n > (1ull << d) - 1) {
        set_error("tl_query_mveb: need 1 <= n <= 2^d - 1 and a table of d entries");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    TL_DISPATCH(dtype, launch_mveb, c, dev_A, n, dev_table, d, (const uint64_t *)nullptr,
                (const uint32_t *)nullptr, 1u, dev_q, nq, dev_ans, s);
    TL_LAUNCH_CHECK();
    return tm.stop();
}

static tl_status query_nonperfect(const void *dev_A, const tl_veb_table *dev_tables,
                                  const uint64_t *dev_idx, const uint32_t *dev_table_idx,
                                  uint64_t n, uint32_t num_tables, tl_dtype dtype,
                                  const void *dev_q, uint64_t nq, uint64_t *dev_ans, void *stream,
                                  float *ms, bool classic) {
    TL_TRY(qcheck(dev_A, n, dev_q, nq, dev_ans));
    if (nq == 0) return TL_OK;
    if (n == 0 || num_tables == 0 || num_tables > 64 || !dev_tables || !dev_idx ||
        !dev_table_idx) {
        set_error("non-perfect vEB query: bad table set");
        return TL_EINVAL;
    }
    // table_size is not part of the reference signature; it is fixed by n
    // (the sum of the structure depths of the chain of incomplete trees)
    const uint32_t table_size = tableset_size(n, classic);
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    TL_DISPATCH(dtype, launch_mveb, c, dev_A, n, dev_tables, table_size, dev_idx, dev_table_idx,
                num_tables, dev_q, nq, dev_ans, s);
    TL_LAUNCH_CHECK();
    return tm.stop();
}

extern "C" tl_status tl_query_mveb_nonperfect(const void *dev_A, const tl_veb_table *dev_tables,
                                              const uint64_t *dev_idx, const uint32_t *dev_table_idx,
                                              uint64_t n, uint32_t d, uint32_t num_tables,
                                              tl_dtype dtype, const void *dev_q, uint64_t nq,
                                              uint64_t *dev_ans, void *stream, float *ms) {
    (void)d;
    return query_nonperfect(dev_A, dev_tables, dev_idx, dev_table_idx, n, num_tables, dtype,
                            dev_q, nq, dev_ans, stream, ms, false);
}

extern "C" tl_status tl_query_veb(const void *dev_A, const tl_veb_table *dev_table, uint64_t n,
                                  uint32_t d, tl_dtype dtype, const void *dev_q, uint64_t nq,
                                  uint64_t *dev_ans, void *stream, float *ms) {
    // the search of a perfect tree depends only on its table (classic rule)
    return tl_query_mveb(dev_A, dev_table, n, d, dtype, dev_q, nq, dev_ans, stream, ms);
}

extern "C" tl_status tl_query_veb_nonperfect(const void *dev_A, const tl_veb_table *dev_tables,
                                             const uint64_t *dev_idx, const uint32_t *dev_table_idx,
                                             uint64_t n, uint32_t d, uint32_t num_tables,
                                             tl_dtype dtype, const void *dev_q, uint64_t nq,
                                             uint64_t *dev_ans, void *stream, float *ms) {
    (void)d;
    return query_nonperfect(dev_A, dev_tables, dev_idx, dev_table_idx, n, num_tables, dtype,
                            dev_q, nq, dev_ans, stream, ms, true);
}

static tl_status build_table_host(tl_veb_table *table, uint64_t n, uint32_t d,
                                  uint32_t root_depth, bool classic) {
    if (!table || d == 0 || d > 63) {
        set_error("vEB table build: bad arguments");
        return TL_EINVAL;
    }
    // host recursion of gpu/src/query-mveb.cu:25-47 (classic:
    // cpu/src/query-veb.cpp:25-41)
    struct Fr {
        tl_veb_table *t;
        uint64_t n;
        uint32_t d, rdp;
    };
    std::vector<Fr> st{{table, n, d, root_depth}};
    while (!st.empty()) {
        Fr f = st.back();
        st.pop_back();
        if (f.n <= 1) continue;
        uint32_t rd, ld;
        veb_split(f.d, rd, ld, classic);
        const uint64_t root_n = (1ull << rd) - 1, leaf_n = (1ull << ld) - 1;
        const uint32_t i = f.d - ld;
        f.t[i].L = leaf_n;
        f.t[i].R = root_n;
        f.t[i].D = f.rdp;
        st.push_back({f.t + i, leaf_n, ld, i + f.rdp});
        st.push_back({f.t, root_n, rd, f.rdp});
    }
    return TL_OK;
}

extern "C" tl_status tl_veb_build_table(tl_veb_table *table, uint64_t n, uint32_t d,
                                        uint32_t root_depth) {
    return build_table_host(table, n, d, root_depth, false);
}

extern "C" tl_status tl_veb_classic_build_table(tl_veb_table *table, uint64_t n, uint32_t d,
                                                uint32_t root_depth) {
    return build_table_host(table, n, d, root_depth, true);
}

static tl_status tableset_device(uint64_t n, tl_veb_table *dev_tables, uint64_t *dev_idx,
                                 uint32_t *dev_table_idx, uint32_t *dev_meta, void *stream,
                                 bool classic) {
    if (n == 0 || !dev_tables || !dev_idx || !dev_table_idx || !dev_meta) {
        set_error("vEB table set: bad arguments");
        return TL_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tableset, dim3(1), dim3(64), 0, s, n, dev_tables, dev_idx, dev_table_idx,
                       dev_meta, classic);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

extern "C" tl_status tl_veb_classic_tableset_device(uint64_t n, tl_veb_table *dev_tables,
                                                    uint64_t *dev_idx, uint32_t *dev_table_idx,
                                                    uint32_t *dev_meta, void *stream) {
    return tableset_device(n, dev_tables, dev_idx, dev_table_idx, dev_meta, stream, true);
}

extern "C" tl_status tl_veb_tableset_device(uint64_t n, tl_veb_table *dev_tables, uint64_t *dev_idx,
                                            uint32_t *dev_table_idx, uint32_t *dev_meta,
                                            void *stream) {
    return tableset_device(n, dev_tables, dev_idx, dev_table_idx, dev_meta, stream, false);
}
