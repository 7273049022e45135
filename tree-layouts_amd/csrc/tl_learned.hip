// tl_learned.hip -- learned start node for level-order B-tree queries (gfx950).
//
// Model: a recursive model index of two layers on the (key, rank) pairs of the
// sorted array.  The root is one linear map key -> leaf; each of the
// num_leaves leaves is a linear map key -> rank with an integer base (first
// key and first rank of the leaf) and an FP32 slope / intercept on the key
// offset, plus the largest rank error seen on its own keys.  Fitting is done
// on the GPU: FP64 least-squares sums by block reductions (root), one
// workgroup per leaf for the leaf sums and the error pass.  Inference is FP32.
//
// Query: leaf -> predicted rank window [pred - err, pred + err] (clipped to
// the leaf's ranks) -> slots of the window ends through the layout's
// rank -> slot closed form -> lowest common ancestor node -> exact group
// descent from that node.  A key inside the window is in that subtree, so
// hits are found there; anything else (a miss, an empty leaf) falls back to
// the plain root descent, which decides the answer.  Answers therefore equal
// tl_query_btree's bit for bit.
#include <algorithm>

#include "tl_internal.hpp"
#include "tl_descent.hpp"

namespace tl {

struct LearnedHeader {   // 64 bytes at the start of the model buffer
    uint64_t kmin;       // order-preserving unsigned image of the smallest key
    uint64_t n;
    float ra, rc;        // leaf = floor(ra * x + rc), x = (key - kmin) * inv
    float inv;           // 1 / (kmax - kmin)
    uint32_t num_leaves;
    uint32_t max_err;
    uint32_t dtype;
    uint64_t fallbacks;  // unused by the fit, kept for symmetry
    uint64_t pad[2];
};

struct LeafModel {       // 32 bytes
    uint64_t base_key;   // unsigned image of the first key of the leaf
    uint64_t base_rank;  // its rank
    float slope, icept;  // rank - base_rank ~ slope * (key - base_key) + icept
    uint32_t err;        // max |prediction - rank| over the leaf's keys
    uint32_t count;      // keys in the leaf
};

static_assert(sizeof(LearnedHeader) == 64, "header size");
static_assert(sizeof(LeafModel) == 32, "leaf size");

// order-preserving map of every key type to uint64
template <typename K> __device__ __host__ inline uint64_t ukey(K k);
template <> __device__ __host__ inline uint64_t ukey<uint32_t>(uint32_t k) { return k; }
template <> __device__ __host__ inline uint64_t ukey<uint64_t>(uint64_t k) { return k; }
template <> __device__ __host__ inline uint64_t ukey<int32_t>(int32_t k) {
    return (uint64_t)((int64_t)k) ^ (1ull << 63);
}
template <> __device__ __host__ inline uint64_t ukey<int64_t>(int64_t k) {
    return (uint64_t)k ^ (1ull << 63);
}

__device__ inline uint32_t leaf_of(uint64_t u, const LearnedHeader &h) {
    const float x = (float)(u - h.kmin) * h.inv;
    const float f = __fmaf_rn(h.ra, x, h.rc);
    int l = (int)floorf(f);
    l = l < 0 ? 0 : l;
    return (uint32_t)min((uint32_t)l, h.num_leaves - 1);
}

__device__ inline int64_t predict(uint64_t u, const LeafModel &m) {
    const float dx = (float)(int64_t)(u - m.base_key);
    const float p = __fmaf_rn(m.slope, dx, m.icept);
    return (int64_t)m.base_rank + (int64_t)rintf(p);
}

// ---------------------------------------------------------------------------
// fit

template <typename K>
__global__ __launch_bounds__(256) void k_root_sums(const K *__restrict__ keys, uint64_t n,
                                                   uint64_t kmin, double inv, double *part) {
    double s[5] = {0, 0, 0, 0, 0};   // x, y, xx, xy, count
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const double x = (double)(ukey<K>(keys[i]) - kmin) * inv;
        const double y = (double)i / (double)n;
        s[0] += x;
        s[1] += y;
        s[2] += x * x;
        s[3] += x * y;
        s[4] += 1.0;
    }
    __shared__ double red[5][256];
    for (int k = 0; k < 5; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; k++) part[(uint64_t)blockIdx.x * 5 + k] = red[k][0];
}

__global__ void k_root_solve(const double *part, uint32_t nparts, uint64_t kmin_u, double range,
                             uint64_t n, uint32_t num_leaves, uint32_t dtype, LearnedHeader *h) {
    if (threadIdx.x != 0) return;
    double s[5] = {0, 0, 0, 0, 0};
    for (uint32_t p = 0; p < nparts; p++)
        for (int k = 0; k < 5; k++) s[k] += part[p * 5 + k];
    const double c = s[4];
    const double mx = s[0] / c, my = s[1] / c;
    const double var = s[2] / c - mx * mx, cov = s[3] / c - mx * my;
    const double a = var > 0 ? cov / var : 1.0;
    const double b0 = my - a * mx;
    LearnedHeader hh{};
    hh.kmin = kmin_u;
    hh.n = n;
    hh.ra = (float)(a * num_leaves);
    hh.rc = (float)(b0 * num_leaves);
    hh.inv = range > 0 ? (float)(1.0 / range) : 0.f;
    hh.num_leaves = num_leaves;
    hh.max_err = 0;
    hh.dtype = dtype;
    *h = hh;
}

template <typename K>
__global__ __launch_bounds__(256) void k_leaf_bounds(const K *__restrict__ keys, uint64_t n,
                                                     const LearnedHeader *__restrict__ hp,
                                                     uint64_t *first) {
    const LearnedHeader h = *hp;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t li = leaf_of(ukey<K>(keys[i]), h);
        const int64_t lp = i ? (int64_t)leaf_of(ukey<K>(keys[i - 1]), h) : -1;
        for (int64_t l = lp + 1; l <= (int64_t)li; l++) first[l] = i;
        if (i == n - 1)
            for (uint32_t l = li + 1; l <= h.num_leaves; l++) first[l] = n;
    }
}

// one workgroup per leaf: least squares of (rank - base_rank) on the FP32 key
// offset, the intercept moved to centre the errors, then the largest error
// of the FP32 prediction on the leaf's keys
template <typename K>
__global__ __launch_bounds__(256) void k_leaf_fit(const K *__restrict__ keys,
                                                  const uint64_t *__restrict__ first,
                                                  LearnedHeader *h, LeafModel *leaves) {
    const uint32_t l = blockIdx.x;
    const uint64_t lo = first[l], hi = first[l + 1];
    __shared__ double red[4][256];
    __shared__ LeafModel sm;
    const uint64_t base = lo < hi ? ukey<K>(keys[lo]) : 0;
    double s[4] = {0, 0, 0, 0};
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const double dx = (double)(float)(int64_t)(ukey<K>(keys[i]) - base);
        const double dy = (double)(i - lo);
        s[0] += dx;
        s[1] += dy;
        s[2] += dx * dx;
        s[3] += dx * dy;
    }
    for (int k = 0; k < 4; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = (double)(hi - lo);
        LeafModel m{};
        m.base_key = base;
        m.base_rank = lo;
        m.count = (uint32_t)(hi - lo);
        if (c > 0) {
            const double mx = red[0][0] / c, my = red[1][0] / c;
            const double var = red[2][0] / c - mx * mx, cov = red[3][0] / c - mx * my;
            const double a = var > 1e-30 ? cov / var : 0.0;
            m.slope = (float)a;
            m.icept = (float)(my - a * mx);
        }
        sm = m;
    }
    __syncthreads();
    // centre the window: the least-squares line's largest errors above and
    // below the ranks (P, N) are rarely equal, and moving the intercept by
    // (P - N) / 2 makes the largest |error| about (P + N) / 2
    {
        int64_t up = 0, dn = 0;
        for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
            const int64_t d = predict(ukey<K>(keys[i]), sm) - (int64_t)i;
            up = max(up, d);
            dn = max(dn, -d);
        }
        red[0][threadIdx.x] = (double)up;
        red[1][threadIdx.x] = (double)dn;
        __syncthreads();
        for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + w]);
                red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double shift = rint((red[0][0] - red[1][0]) / 2.0);
            if (shift != 0.0 && fabs(shift) < 16777216.0) sm.icept = (float)((double)sm.icept - shift);
        }
        __syncthreads();
    }
    // the window's half width: the largest error of the model as stored
    const LeafModel m = sm;
    uint32_t e = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int64_t d = predict(ukey<K>(keys[i]), m) - (int64_t)i;
        const uint64_t ad = (uint64_t)(d < 0 ? -d : d);
        e = max(e, (uint32_t)min(ad, (uint64_t)0xFFFFFFFFu));
    }
    __shared__ uint32_t ered[256];
    ered[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) ered[threadIdx.x] = max(ered[threadIdx.x], ered[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        LeafModel out = m;
        out.err = ered[0];
        leaves[l] = out;
        atomicMax(&h->max_err, ered[0]);
    }
}

// ---------------------------------------------------------------------------
// query

// Start node of a rank window [lo, hi] in the level-order B-tree, in closed
// form.  A rank is placed among the internal tree's keys: internal rank r sits
// at (p, q) = (r, r+1) and every key of leaf node j at (j, j) (between
// internal ranks j-1 and j).  The internal node at height h (bottom internal
// level h = 1) with block m holds exactly the positions with
// floor(p / k1^h) = floor(q / k1^h) = m, so the lowest node holding the whole
// window is the first height where floor(p_lo / k1^h) = floor(q_hi / k1^h).
// Both keys in one leaf node start at that leaf node.
struct BtreeConst {
    uint64_t n, b, k1;
    uint64_t perfect, nint, span, rem, nsep;
    uint64_t mk1;              // ceil(2^64 / k1)
    uint32_t D, pad;           // levels of the internal (whole, if perfect) tree
    uint64_t pw[41];           // k1^h
    uint64_t mag[41];          // ceil(2^64 / k1^h)
    uint64_t nodestart[41];    // heap index of the first node at height h
};

// x / k1^h.  With FAST (x < 2^32, k1^h < 2^32) a multiply-high by
// ceil(2^64 / k1^h) is exact: x * e < 2^64 for the rounding excess e < k1^h.
template <bool FAST>
__device__ inline uint64_t div_pw(uint64_t x, uint32_t h, const uint64_t *pw, const uint64_t *mag) {
    return FAST ? __umul64hi(x, mag[h]) : x / pw[h];
}

template <bool FAST>
__device__ inline void rank_place(uint64_t x, const BtreeConst &c, const uint64_t *pw,
                                  const uint64_t *mag, uint64_t &p, uint64_t &q, bool &leaf) {
    leaf = false;
    if (!c.perfect) {
        if (x < c.span) {
            const uint64_t row = div_pw<FAST>(x, 1, pw, mag), cc = x - row * c.k1;
            p = row;
            q = cc < c.b ? row : row + 1;
            leaf = cc < c.b;
            return;
        }
        if (x < c.span + c.rem) {
            p = q = c.nsep;
            leaf = true;
            return;
        }
        x = c.nsep + (x - c.span - c.rem);
    }
    p = x;
    q = x + 1;
}

template <bool FAST>
__device__ inline uint64_t start_slot(uint64_t lo, uint64_t hi, const BtreeConst &c,
                                      const uint64_t *pw, const uint64_t *mag,
                                      const uint64_t *nodestart) {
    uint64_t plo, qlo, phi, qhi;
    bool llo, lhi;
    rank_place<FAST>(lo, c, pw, mag, plo, qlo, llo);
    rank_place<FAST>(hi, c, pw, mag, phi, qhi, lhi);
    if (llo && lhi && plo == phi) return c.nint + plo * c.b;   // one leaf node
    uint32_t h = 1;
    while (h < c.D && div_pw<FAST>(plo, h, pw, mag) != div_pw<FAST>(qhi, h, pw, mag)) h++;
    return (nodestart[h] + div_pw<FAST>(plo, h, pw, mag)) * c.b;
}

// the start slot of one query key divided by b (a node number; every start
// slot is a node start), 0 (the root) when the model has no usable window
template <typename K, bool FAST>
__device__ inline uint64_t start_node(K key, const LearnedHeader &h,
                                      const LeafModel *__restrict__ leaves, const BtreeConst &c,
                                      const uint64_t *pw, const uint64_t *mag,
                                      const uint64_t *nodestart, uint64_t nint_nodes,
                                      bool &fallback) {
    const uint64_t u = ukey<K>(key);
    const LeafModel m = leaves[leaf_of(u, h)];
    fallback = true;
    if (!m.count) return 0;
    const int64_t pred = predict(u, m);
    int64_t lo = pred - (int64_t)m.err, hi = pred + (int64_t)m.err;
    lo = max(lo, (int64_t)m.base_rank);
    hi = min(hi, (int64_t)(m.base_rank + m.count - 1));
    if (lo > hi) return 0;
    fallback = false;
    uint64_t plo, qlo, phi, qhi;
    bool llo, lhi;
    rank_place<FAST>((uint64_t)lo, c, pw, mag, plo, qlo, llo);
    rank_place<FAST>((uint64_t)hi, c, pw, mag, phi, qhi, lhi);
    if (llo && lhi && plo == phi) return nint_nodes + plo;   // one leaf node
    uint32_t ht = 1;
    while (ht < c.D && div_pw<FAST>(plo, ht, pw, mag) != div_pw<FAST>(qhi, ht, pw, mag)) ht++;
    return nodestart[ht] + div_pw<FAST>(plo, ht, pw, mag);
}

constexpr int START_QPL = 4;   // start-pass queries per lane

// Pass 1: the start slot of every query, one lane per query, written into the
// answer array (no dependent chain, so the model loads and the index
// arithmetic run at throughput).  Queries without a usable leaf start at the
// root (slot 0) and are counted as fallbacks.
template <typename K, bool FAST>
__global__ __launch_bounds__(256) void k_learned_start(BtreeConst bc,
                                                       const LearnedHeader *__restrict__ hp,
                                                       const LeafModel *__restrict__ leaves,
                                                       const K *__restrict__ Q, uint64_t nq,
                                                       uint64_t *__restrict__ ans,
                                                       unsigned long long *stats) {
    __shared__ uint64_t s_pw[41], s_mag[41], s_ns[41];
    for (uint32_t t = threadIdx.x; t < 41; t += blockDim.x) {
        s_pw[t] = bc.pw[t];
        s_mag[t] = bc.mag[t];
        s_ns[t] = bc.nodestart[t];
    }
    __syncthreads();
    const LearnedHeader h = *hp;
    uint32_t fb = 0;
    // START_QPL queries per lane: their model loads are in flight together
    const uint64_t all = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < nq;
         q0 += all * START_QPL) {
        uint64_t u[START_QPL];
        LeafModel m[START_QPL];
#pragma unroll
        for (int k = 0; k < START_QPL; k++) {
            const uint64_t qi = q0 + k * all;
            u[k] = qi < nq ? ukey<K>(Q[qi]) : 0;
        }
#pragma unroll
        for (int k = 0; k < START_QPL; k++)
            if (q0 + k * all < nq) m[k] = leaves[leaf_of(u[k], h)];
#pragma unroll
        for (int k = 0; k < START_QPL; k++) {
            const uint64_t qi = q0 + k * all;
            if (qi >= nq) continue;
            uint64_t start = 0;
            if (m[k].count) {
                const int64_t pred = predict(u[k], m[k]);
                int64_t lo = pred - (int64_t)m[k].err, hi = pred + (int64_t)m[k].err;
                lo = max(lo, (int64_t)m[k].base_rank);
                hi = min(hi, (int64_t)(m[k].base_rank + m[k].count - 1));
                if (lo <= hi)
                    start = start_slot<FAST>((uint64_t)lo, (uint64_t)hi, bc, s_pw, s_mag, s_ns);
                else
                    fb++;
            } else {
                fb++;
            }
            ans[qi] = start;
        }
    }
    if (stats) {
        for (int o = 32; o > 0; o >>= 1) fb += __shfl_xor(fb, o);
        if ((threadIdx.x & 63) == 0 && fb) atomicAdd(&stats[0], (unsigned long long)fb);
    }
}

// Pass 2: group descent from each query's start slot (read from the answer
// array), two queries per group.  A descent that started below the root and
// did not find the key is redone from the root: the plain descent decides
// every other case, so the answers equal tl_query_btree's.
template <typename K, int G>
__global__ __launch_bounds__(256) void k_query_btree_from(const K *__restrict__ A, uint64_t n,
                                                          uint32_t b, const K *__restrict__ Q,
                                                          uint64_t nq, uint64_t *__restrict__ ans,
                                                          unsigned long long *stats) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g = lane / G, j = lane % G;
    const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / G);
    const uint64_t trips = (nq + 2 * groups - 1) / (2 * groups);   // uniform trip count
    uint64_t qa = (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    uint32_t fb = 0;
    for (uint64_t it = 0; it < trips; it++, qa += 2 * groups) {
        const uint64_t qb = qa + groups;
        const bool va = qa < nq, vb = qb < nq;
        K q0{}, q1{};
        uint64_t s0 = 0, s1 = 0;
        if (va) {
            q0 = Q[qa];
            s0 = ans[qa];
        }
        if (vb) {
            q1 = Q[qb];
            s1 = ans[qb];
        }
        uint64_t r0, r1;
        bool f0, f1;
        descent2<K, G>(A, n, b, g, j, q0, q1, s0, s1, va, vb, r0, r1, f0, f1);
        const bool again0 = va && s0 != 0 && !f0, again1 = vb && s1 != 0 && !f1;
        if (j == 0) fb += again0 + again1;
        uint64_t p0, p1;
        descent2<K, G>(A, n, b, g, j, q0, q1, 0, 0, again0, again1, p0, p1, f0, f1);
        if (again0) r0 = p0;
        if (again1) r1 = p1;
        if (j == 0) {
            if (va) ans[qa] = r0;
            if (vb) ans[qb] = r1;
        }
    }
    if (stats) {
        for (int o = 32; o > 0; o >>= 1) fb += __shfl_xor(fb, o);
        if (lane == 0 && fb) atomicAdd(&stats[0], (unsigned long long)fb);
    }
}

// Pass 2, wide form (descent_wide): the start slots are node starts (k * b),
// QN queries per group; the same redo from the root.
template <typename K, int G, int QN, typename IDX>
__global__ __launch_bounds__(256) void k_query_btree_from_wide(const K *__restrict__ A, uint64_t n,
                                                               uint32_t b, const K *__restrict__ Q,
                                                               uint64_t nq,
                                                               uint64_t *__restrict__ ans,
                                                               unsigned long long *stats) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t j = lane % G;
    const uint64_t gm = group_mask<G>(lane);
    const IDX nodes = (IDX)((n + b - 1) / b);
    const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / G);
    const uint64_t trips = (nq + QN * groups - 1) / (QN * groups);   // uniform trip count
    uint64_t q0 = (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    uint32_t fb = 0;
    for (uint64_t it = 0; it < trips; it++, q0 += QN * groups) {
        K q[QN];
        IDX node[QN];
        bool act[QN], found[QN], redo[QN];
        uint64_t st[QN], res[QN], res2[QN];
#pragma unroll
        for (int k = 0; k < QN; k++) {
            const uint64_t qi = q0 + k * groups;
            act[k] = qi < nq;
            q[k] = act[k] ? Q[qi] : K{};
            st[k] = act[k] ? ans[qi] : 0;
            node[k] = (IDX)(st[k] / b);
        }
        descent_wide<K, QN, IDX>(A, n, b, nodes, gm, j, q, node, act, res, found);
#pragma unroll
        for (int k = 0; k < QN; k++) {
            redo[k] = q0 + k * groups < nq && st[k] != 0 && !found[k];
            act[k] = redo[k];
            node[k] = 0;
            if (j == 0) fb += redo[k];
        }
        descent_wide<K, QN, IDX>(A, n, b, nodes, gm, j, q, node, act, res2, found);
        if (j == 0) {
#pragma unroll
            for (int k = 0; k < QN; k++) {
                const uint64_t qi = q0 + k * groups;
                if (qi < nq) ans[qi] = redo[k] ? res2[k] : res[k];
            }
        }
    }
    if (stats) {
        for (int o = 32; o > 0; o >>= 1) fb += __shfl_xor(fb, o);
        if (lane == 0 && fb) atomicAdd(&stats[0], (unsigned long long)fb);
    }
}

// One pass (wide form, n < 2^32, 32-bit node numbers): a group of G lanes
// takes G consecutive queries per trip; lane j computes the start node of
// query j (the group's model loads in flight together, no start slots through
// HBM), then the group descends the G queries QN at a time, keys and start
// nodes broadcast by lane shuffles, and lane j stores query j's answer.  The
// fallback count and the redo from the root are those of the two passes.
template <typename K, int G, int QN>
__global__ __launch_bounds__(256) void k_query_btree_learned_fused(
    BtreeConst bc, const LearnedHeader *__restrict__ hp, const LeafModel *__restrict__ leaves,
    const K *__restrict__ A, const K *__restrict__ Q, uint64_t nq, uint64_t *__restrict__ ans,
    unsigned long long *stats) {
    static_assert(G >= QN && G % QN == 0, "a group descends its queries QN at a time");
    __shared__ uint64_t s_pw[41], s_mag[41], s_ns[41];
    for (uint32_t t = threadIdx.x; t < 41; t += blockDim.x) {
        s_pw[t] = bc.pw[t];
        s_mag[t] = bc.mag[t];
        s_ns[t] = bc.nodestart[t];
    }
    __syncthreads();
    const LearnedHeader h = *hp;
    const uint64_t n = bc.n;
    const uint32_t b = (uint32_t)bc.b;
    const uint64_t nint_nodes = bc.nint / bc.b;
    const uint32_t nodes = (uint32_t)((n + b - 1) / b);
    const uint32_t lane = threadIdx.x & 63, j = lane % G, gbase = lane - j;
    const uint64_t gm = group_mask<G>(lane);
    const uint64_t groups = (uint64_t)gridDim.x * (blockDim.x / G);
    const uint64_t gid = (uint64_t)blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
    const uint64_t trips = (nq + groups * G - 1) / (groups * G);   // uniform trip count
    uint32_t fb = 0;
    for (uint64_t it = 0; it < trips; it++) {
        const uint64_t qi = (it * groups + gid) * G + j;
        const bool valid = qi < nq;
        const K myq = valid ? Q[qi] : K{};
        uint32_t mystart = 0;
        if (valid) {
            bool fall;
            mystart = (uint32_t)start_node<K, true>(myq, h, leaves, bc, s_pw, s_mag, s_ns,
                                                    nint_nodes, fall);
            fb += fall;
        }
        uint64_t myres = n;
#pragma unroll 1
        for (uint32_t r = 0; r < G / QN; r++) {
            K q[QN];
            uint32_t node[QN], st[QN];
            bool act[QN], found[QN], redo[QN];
            uint64_t res[QN], res2[QN];
#pragma unroll
            for (int k = 0; k < QN; k++) {
                const int src = (int)(gbase + r * QN + k);
                q[k] = __shfl(myq, src);
                st[k] = (uint32_t)__shfl((int)mystart, src);
                act[k] = __shfl((int)valid, src) != 0;
                node[k] = st[k];
            }
            descent_wide<K, QN, uint32_t>(A, n, b, nodes, gm, j, q, node, act, res, found);
#pragma unroll
            for (int k = 0; k < QN; k++) {
                const bool v = __shfl((int)valid, (int)(gbase + r * QN + k)) != 0;
                redo[k] = v && st[k] != 0 && !found[k];
                act[k] = redo[k];
                node[k] = 0;
            }
            descent_wide<K, QN, uint32_t>(A, n, b, nodes, gm, j, q, node, act, res2, found);
#pragma unroll
            for (int k = 0; k < QN; k++)
                if (j == r * QN + k) {
                    myres = redo[k] ? res2[k] : res[k];
                    fb += redo[k];
                }
        }
        if (valid) ans[qi] = myres;
    }
    if (stats) {
        for (int o = 32; o > 0; o >>= 1) fb += __shfl_xor(fb, o);
        if (lane == 0 && fb) atomicAdd(&stats[0], (unsigned long long)fb);
    }
}

// ---------------------------------------------------------------------------

template <typename K>
static tl_status fit(Ctx *c, const void *dev_sorted, uint64_t n, uint32_t num_leaves,
                     uint32_t dtype, void *dev_model, hipStream_t s) {
    const K *keys = (const K *)dev_sorted;
    K kmin_k, kmax_k;
    TL_HIP(hipMemcpyAsync(&kmin_k, keys, sizeof(K), hipMemcpyDeviceToHost, s));
    TL_HIP(hipMemcpyAsync(&kmax_k, keys + (n - 1), sizeof(K), hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    const uint64_t kmin_u = ukey<K>(kmin_k), kmax_u = ukey<K>(kmax_k);
    if (kmax_u < kmin_u) {
        set_error("tl_learned_fit: keys are not sorted");
        return TL_EINVAL;
    }
    const double range = (double)(kmax_u - kmin_u);
    const double inv = range > 0 ? 1.0 / range : 0.0;
    const uint32_t nparts = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c->num_cus * 8);
    double *part;
    TL_TRY(scratch(c, 3, (size_t)nparts * 5 * sizeof(double) + (num_leaves + 1) * 8ull,
                   (void **)&part));
    uint64_t *first = (uint64_t *)(part + (size_t)nparts * 5);
    LearnedHeader *h = (LearnedHeader *)dev_model;
    LeafModel *leaves = (LeafModel *)((char *)dev_model + sizeof(LearnedHeader));
    hipLaunchKernelGGL(k_root_sums<K>, dim3(nparts), dim3(256), 0, s, keys, n, kmin_u, inv, part);
    TL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_root_solve, dim3(1), dim3(64), 0, s, part, nparts, kmin_u, range, n,
                       num_leaves, dtype, h);
    TL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_leaf_bounds<K>, dim3(nparts), dim3(256), 0, s, keys, n,
                       (const LearnedHeader *)h, first);
    TL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_leaf_fit<K>, dim3(num_leaves), dim3(256), 0, s, keys,
                       (const uint64_t *)first, h, leaves);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

template <typename K>
static tl_status query(Ctx *c, const void *A, uint64_t n, uint64_t b, const void *model,
                       const void *q, uint64_t nq, uint64_t *ans, unsigned long long *stats,
                       hipStream_t s) {
    BtreeConst bc{};
    bc.n = n;
    bc.b = b;
    bc.k1 = b + 1;
    const uint32_t h = ilog_base(n, b + 1);
    const uint64_t P = ipow64(b + 1, h);
    bc.perfect = (n == P * (b + 1) - 1);
    if (bc.perfect) {
        bc.nint = n;
    } else {
        const uint64_t nint = P - 1, nleaf = n - nint;
        const uint64_t rem = nleaf % b, l = nleaf - rem, i = l / b;
        bc.nint = nint;
        bc.span = l + i;
        bc.rem = rem;
        bc.nsep = i;
    }
    // ceil(2^64 / d) = floor((2^64 - 1) / d) + 1 for every d > 1
    auto ceil_recip = [](uint64_t d) -> uint64_t { return ~0ull / d + 1; };
    bc.mk1 = ceil_recip(bc.k1);
    bc.D = bc.perfect ? h + 1 : h;   // levels of the internal (or whole) tree
    {
        uint64_t p = 1;
        for (uint32_t t = 0; t <= 40; t++) {
            bc.pw[t] = p;
            bc.mag[t] = t ? ceil_recip(p) : 0;
            if (t >= 1 && t <= bc.D) bc.nodestart[t] = (ipow64(bc.k1, bc.D - t) - 1) / b;
            if (p > (~0ull) / bc.k1) p = ~0ull;   // saturate; never reached below height D
            else p *= bc.k1;
        }
    }
    const bool fast = n < (1ull << 32);
    const LearnedHeader *hp = (const LearnedHeader *)model;
    const LeafModel *leaves = (const LeafModel *)((const char *)model + sizeof(LearnedHeader));
    if (learned_fused() && fast && wide_ok(A, b, sizeof(K)) && wide_narrow(n, b) &&
        wide_group(b, sizeof(K)) >= 2) {
        const uint32_t G = wide_group(b, sizeof(K));
        const uint64_t want = (nq + 256 - 1) / 256;
        // many more workgroups than stay resident (256 per CU) run faster
        // for this kernel, which stages no keys (8.6 against 9.5 ms with 16)
        const uint32_t grid = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(want, (uint64_t)c->num_cus * 256));
        switch (G) {
#define TL_F(GG)                                                                                   \
    case GG:                                                                                       \
        hipLaunchKernelGGL((k_query_btree_learned_fused<K, GG, 2>), dim3(grid), dim3(256), 0, s,  \
                           bc, hp, leaves, (const K *)A, (const K *)q, nq, ans, stats);            \
        break;
            TL_F(2) TL_F(4) TL_F(8) TL_F(16) TL_F(32)
#undef TL_F
        }
        TL_LAUNCH_CHECK();
        return TL_OK;
    }
    const uint64_t sgrid = std::max<uint64_t>(
        1, std::min<uint64_t>((nq + 256 * START_QPL - 1) / (256 * START_QPL),
                              (uint64_t)c->num_cus * 16));
    if (fast)
        hipLaunchKernelGGL((k_learned_start<K, true>), dim3((uint32_t)sgrid), dim3(256), 0, s, bc, hp,
                           leaves, (const K *)q, nq, ans, stats);
    else
        hipLaunchKernelGGL((k_learned_start<K, false>), dim3((uint32_t)sgrid), dim3(256), 0, s, bc,
                           hp, leaves, (const K *)q, nq, ans, stats);
    TL_LAUNCH_CHECK();
    if (wide_ok(A, b, sizeof(K))) {
        const uint32_t G = wide_group(b, sizeof(K));
        const int QN = btree_qpg(2);
        const bool narrow = wide_narrow(n, b);
        const uint64_t want = (nq + QN * 256 / G - 1) / (QN * 256 / G);
        const uint32_t grid = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(want, (uint64_t)c->num_cus * 16));
#define TL_WL(GG, QQ, IX)                                                                          \
    hipLaunchKernelGGL((k_query_btree_from_wide<K, GG, QQ, IX>), dim3(grid), dim3(256), 0, s,      \
                       (const K *)A, n, (uint32_t)b, (const K *)q, nq, ans, stats)
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
        TL_LAUNCH_CHECK();
        return TL_OK;
    }
    const uint32_t G = b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : b <= 16 ? 16 : b <= 32 ? 32 : 64;
    const uint64_t want = (nq + 2 * 256 / G - 1) / (2 * 256 / G);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)c->num_cus * 16));
    switch (G) {
#define TL_G(GG)                                                                                    \
    case GG:                                                                                        \
        hipLaunchKernelGGL((k_query_btree_from<K, GG>), dim3(grid), dim3(256), 0, s, (const K *)A,  \
                           n, (uint32_t)b, (const K *)q, nq, ans, stats);                           \
        break;
        TL_G(2) TL_G(4) TL_G(8) TL_G(16) TL_G(32) TL_G(64)
#undef TL_G
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// per key: the model's predicted rank and its leaf's error bound (-1 for both
// when the key maps to an empty leaf), with the query path's own arithmetic
template <typename K>
__global__ __launch_bounds__(256) void k_learned_predict(const LearnedHeader *__restrict__ hp,
                                                         const LeafModel *__restrict__ leaves,
                                                         const K *__restrict__ keys, uint64_t n,
                                                         int64_t *__restrict__ pred,
                                                         int64_t *__restrict__ err) {
    const LearnedHeader h = *hp;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = ukey<K>(keys[i]);
        const LeafModel m = leaves[leaf_of(u, h)];
        pred[i] = m.count ? predict(u, m) : -1;
        err[i] = m.count ? (int64_t)m.err : -1;
    }
}

}   // namespace tl

using namespace tl;

extern "C" tl_status tl_learned_predict(const void *dev_model, const void *dev_keys, uint64_t n,
                                        tl_dtype dtype, int64_t *dev_pred, int64_t *dev_err,
                                        void *stream) {
    if (n == 0) return TL_OK;
    if (!dev_model || !dev_keys || !dev_pred || !dev_err) {
        set_error("tl_learned_predict: null buffer");
        return TL_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const LearnedHeader *h = (const LearnedHeader *)dev_model;
    const LeafModel *lv = (const LeafModel *)(h + 1);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
    switch (dtype) {
    case TL_U32:
        hipLaunchKernelGGL(k_learned_predict<uint32_t>, dim3(grid), dim3(256), 0, s, h, lv,
                           (const uint32_t *)dev_keys, n, dev_pred, dev_err);
        break;
    case TL_U64:
        hipLaunchKernelGGL(k_learned_predict<uint64_t>, dim3(grid), dim3(256), 0, s, h, lv,
                           (const uint64_t *)dev_keys, n, dev_pred, dev_err);
        break;
    case TL_I32:
        hipLaunchKernelGGL(k_learned_predict<int32_t>, dim3(grid), dim3(256), 0, s, h, lv,
                           (const int32_t *)dev_keys, n, dev_pred, dev_err);
        break;
    case TL_I64:
        hipLaunchKernelGGL(k_learned_predict<int64_t>, dim3(grid), dim3(256), 0, s, h, lv,
                           (const int64_t *)dev_keys, n, dev_pred, dev_err);
        break;
    default: set_error("unknown dtype"); return TL_EINVAL;
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

extern "C" uint64_t tl_learned_model_bytes(uint32_t num_leaves) {
    return sizeof(LearnedHeader) + (uint64_t)num_leaves * sizeof(LeafModel);
}

extern "C" tl_status tl_learned_fit(const void *dev_sorted, uint64_t n, tl_dtype dtype,
                                    uint32_t num_leaves, void *dev_model, void *stream,
                                    float *ms) {
    if (!dev_sorted || !dev_model || n < 2 || num_leaves == 0 || num_leaves > (1u << 24)) {
        set_error("tl_learned_fit: need n >= 2 sorted keys, 1..2^24 leaves and a model buffer");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    tl_status st;
    switch (dtype) {
    case TL_U32: st = fit<uint32_t>(c, dev_sorted, n, num_leaves, dtype, dev_model, s); break;
    case TL_U64: st = fit<uint64_t>(c, dev_sorted, n, num_leaves, dtype, dev_model, s); break;
    case TL_I32: st = fit<int32_t>(c, dev_sorted, n, num_leaves, dtype, dev_model, s); break;
    case TL_I64: st = fit<int64_t>(c, dev_sorted, n, num_leaves, dtype, dev_model, s); break;
    default: set_error("unknown dtype"); return TL_EINVAL;
    }
    if (st != TL_OK) return st;
    return tm.stop();
}

extern "C" tl_status tl_query_btree_learned(const void *dev_A, uint64_t n, uint64_t b,
                                            tl_dtype dtype, const void *dev_model,
                                            const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                                            uint64_t *dev_stats, void *stream, float *ms) {
    if (nq == 0) return TL_OK;
    if (!dev_A || !dev_model || !dev_q || !dev_ans || n < 2 || b < 2 || b > 64) {
        set_error("tl_query_btree_learned: need n >= 2, 2 <= b <= 64 and device buffers");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    unsigned long long *st = (unsigned long long *)dev_stats;
    tl_status r;
    switch (dtype) {
    case TL_U32: r = query<uint32_t>(c, dev_A, n, b, dev_model, dev_q, nq, dev_ans, st, s); break;
    case TL_U64: r = query<uint64_t>(c, dev_A, n, b, dev_model, dev_q, nq, dev_ans, st, s); break;
    case TL_I32: r = query<int32_t>(c, dev_A, n, b, dev_model, dev_q, nq, dev_ans, st, s); break;
    case TL_I64: r = query<int64_t>(c, dev_A, n, b, dev_model, dev_q, nq, dev_ans, st, s); break;
    default: set_error("unknown dtype"); return TL_EINVAL;
    }
    if (r != TL_OK) return r;
    return tm.stop();
}
