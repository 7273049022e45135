// tl_involution.hip -- the involution family of in-place layout permutations
// (the reference's second set of programs, gpu/src/bst-involutions.cuh,
// gpu/src/btree-involutions.cuh, gpu/src/mveb-involutions.cuh) and its
// primitives: the k-way shuffle and unshuffle of gpu/src/involutions.cuh and
// the rotation by reversals of gpu/src/common.cuh:62-110.
//
// Every pass is an involution: each key swaps with a partner given by a closed
// form, pairs are disjoint, and the lane holding the smaller index of a pair
// does the swap, so a pass needs no scratch, no waits and no order between
// lanes.  The layouts equal those of tl_permute.hip's gather / cycle-leader
// path (the reference's two families agree); these passes move more bytes
// (two or four passes per level, the modular partners scattered), so the
// involution family is the slower one and is kept for callers that ask for it.
//
// Passes over whole arrays run as grid-stride launches, one lane per index.
// The mvEB's base cases (partitions of 255, 15 and 3 keys, and any later
// round whose partitions fit the LDS budget) run in LDS: one workgroup loads a
// partition, runs every remaining pass on it there and writes it back once.
#include "tl_internal.hpp"
#include "tl_modinv.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

namespace tl {

// ---------------------------------------------------------------------------
// integer helpers

// floor(x / d) for a divisor fixed per pass.  With M = floor((2^64 - 1) / d) + 1
// and 2 <= d < 2^32, floor(x / d) = (M * x) >> 64 for every x < 2^32 (Lemire,
// Kaser and Kurz, "Faster remainder by direct computation", 2019, theorem 1):
// one 64-bit high multiply in place of a 64-bit division.  Larger x or d take
// the division.
struct FastDiv {
    uint64_t d = 1;
    uint64_t M = 0;
};

static FastDiv make_div(uint64_t d) {
    FastDiv f;
    f.d = d ? d : 1;
    f.M = (f.d >= 2 && f.d < (1ull << 32)) ? (~0ull / f.d + 1) : 0;
    return f;
}

__device__ inline uint64_t fdiv(uint64_t x, const FastDiv &f) {
    if (f.M && x < (1ull << 32)) return __umul64hi(f.M, x);
    return f.d == 1 ? x : x / f.d;
}

// x with its nd lowest base-k digits reversed (the higher digits kept):
// rev_base_d / rev_base_b of gpu/src/involutions.cu:34-69, digit by digit from
// the low end instead of from a power computed with pow()
__device__ inline uint64_t rev_digits(uint64_t x, const FastDiv &k, uint32_t nd) {
    uint64_t r = 0, p = 1;
    for (uint32_t t = 0; t < nd; t++) {
        const uint64_t q = fdiv(x, k);
        r = r * k.d + (x - q * k.d);
        x = q;
        p *= k.d;
    }
    return x * p + r;
}

// ---------------------------------------------------------------------------
// one pass

enum : uint32_t {
    IV_DIGITS = 0,   // view index i <-> rev_digits(i + 1, k, nd) - 1 (unshuffle phases)
    IV_MODINV = 1,   // i in [1, len - 1) <-> mod_partner(i, len - 1, k)  (shuffle_dk phases)
    IV_REV2 = 2,     // [0, k) and [k, len) each reversed (k = 0: the whole view)
};

struct IPass {
    uint32_t kind = IV_DIGITS;
    uint32_t nd = 0;
    int64_t base = 0;      // first key of partition 0's view (array offset)
    uint64_t parts = 1;    // partitions, `stride` keys apart
    uint64_t stride = 0;
    uint64_t len = 0;      // keys in each partition's view
    uint64_t k = 0;        // digit base / multiplier / first part of IV_REV2
    FastDiv kd;            // division by k (IV_DIGITS)
    FastDiv lend;          // division by len (a lane's partition)
    uint64_t tab = 0;      // device table of the partners of 0 .. len-1 (0: computed per key)
};

// partners come from a table when a pass's shape repeats over partitions
constexpr uint64_t IV_TABLE_MAX = 1u << 17;

__device__ inline bool partner(const IPass &p, uint32_t kind, uint64_t i, uint64_t &j) {
    if (kind == IV_DIGITS) {
        j = rev_digits(i + 1, p.kd, p.nd) - 1;
    } else if (kind == IV_MODINV) {
        if (i == 0 || i + 1 >= p.len) return false;
        j = mod_partner(i, p.len - 1, p.k);
    } else {
        j = i < p.k ? p.k - 1 - i : p.k + p.len - 1 - i;
    }
    return j > i;
}

// the partner of i from the pass's table (an entry j <= i: no swap), or
// computed
__device__ inline bool partner_of(const IPass &p, uint32_t kind, uint64_t i, uint64_t &j) {
    if (p.tab) {
        j = ((const uint32_t *)p.tab)[i];
        return j > i;
    }
    return partner(p, kind, i, j);
}

// the table of one pass shape: entry i = i's partner when it is larger, else i
__global__ __launch_bounds__(256) void k_partner_table(IPass p, uint32_t *__restrict__ tab) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < p.len;
         i += (uint64_t)gridDim.x * 256) {
        uint64_t j;
        tab[i] = (uint32_t)(partner(p, p.kind, i, j) ? j : i);
    }
}

template <typename T, uint32_t KIND>
__global__ __launch_bounds__(256) void k_involution(T *__restrict__ A, IPass p) {
    const uint64_t total = p.parts * p.len;
    for (uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * 256) {
        const uint64_t q = p.parts == 1 ? 0 : fdiv(g, p.lend);
        const uint64_t i = g - q * p.len;
        uint64_t j;
        if (!partner_of(p, KIND, i, j)) continue;
        const int64_t o = p.base + (int64_t)(q * p.stride);
        const T a = A[o + (int64_t)i], b = A[o + (int64_t)j];
        A[o + (int64_t)i] = b;
        A[o + (int64_t)j] = a;
    }
}

// every partition of m keys (base + q * stride) through LDS: load, the passes
// of `prog` (offsets relative to the partition), store
template <typename T>
__global__ __launch_bounds__(256) void k_involution_lds(T *__restrict__ A, int64_t base,
                                                        uint64_t parts, uint64_t stride,
                                                        uint32_t m, const IPass *__restrict__ prog,
                                                        uint32_t nprog) {
    extern __shared__ __align__(16) unsigned char smem_inv[];
    T *s = (T *)smem_inv;
    for (uint64_t q = blockIdx.x; q < parts; q += gridDim.x) {
        T *a = A + base + (int64_t)(q * stride);
        for (uint32_t i = threadIdx.x; i < m; i += 256) s[i] = a[i];
        __syncthreads();
        for (uint32_t t = 0; t < nprog; t++) {
            const IPass p = prog[t];
            const uint64_t total = p.parts * p.len;
            for (uint64_t g = threadIdx.x; g < total; g += 256) {
                const uint64_t pq = p.parts == 1 ? 0 : fdiv(g, p.lend);
                const uint64_t i = g - pq * p.len;
                uint64_t j;
                if (!partner_of(p, p.kind, i, j)) continue;
                const int64_t o = p.base + (int64_t)(pq * p.stride);
                const T x = s[o + (int64_t)i];
                s[o + (int64_t)i] = s[o + (int64_t)j];
                s[o + (int64_t)j] = x;
            }
            __syncthreads();
        }
        for (uint32_t i = threadIdx.x; i < m; i += 256) a[i] = s[i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host: the pass list of a permutation

struct IStep {
    bool lds = false;
    IPass pass;                 // a global pass
    int64_t base = 0;           // an LDS group: partitions of m keys ...
    uint64_t parts = 0, stride = 0;
    uint32_t m = 0;
    std::vector<IPass> prog;    // ... and their passes, relative to a partition
};

struct IPlan {
    std::vector<IStep> steps;
    uint64_t lds_keys = 0;      // largest partition an LDS group may hold
};

// the passes of one recursion go to `out` (a plan's steps or a group's prog)
struct IEmit {
    IPlan *plan = nullptr;
    std::vector<IPass> *prog = nullptr;

    void pass(uint32_t kind, int64_t base, uint64_t parts, uint64_t stride, uint64_t len,
              uint64_t k, uint32_t nd) {
        if (parts == 0 || len < 2) return;
        if (kind == IV_MODINV && len < 3) return;   // no index in [1, len - 1) moves
        IPass p;
        p.kind = kind;
        p.nd = nd;
        p.base = base;
        p.parts = parts;
        p.stride = stride;
        p.len = len;
        p.k = k;
        p.kd = make_div(kind == IV_DIGITS ? k : 1);
        p.lend = make_div(len);
        if (prog) {
            prog->push_back(p);
        } else {
            IStep st;
            st.pass = p;
            plan->steps.push_back(st);
        }
    }
    // k-way unshuffle of n = k^d - 1 keys (gpu/src/involutions.cuh:38,61)
    void unshuffle(int64_t base, uint64_t k, uint64_t n, uint32_t d) {
        part_unshuffle(base, n, k, n, d);
    }
    // the same on each partition of m keys of [base, base + n)
    // (gpu/src/involutions.cuh:129,157)
    void part_unshuffle(int64_t base, uint64_t n, uint64_t k, uint64_t m, uint32_t d) {
        pass(IV_DIGITS, base, n / m, m, m, k, d);
        pass(IV_DIGITS, base, n / m, m, m, k, d - 1);
    }
    // k-way shuffle of n = d * k keys: involution(1, .) then involution(k, .)
    // (gpu/src/involutions.cuh:84,106); unshuffle_dk the two in reverse order
    void shuffle_dk(int64_t base, uint64_t k, uint64_t n) {
        pass(IV_MODINV, base, 1, 0, n, 1, 0);
        pass(IV_MODINV, base, 1, 0, n, k, 0);
    }
    void unshuffle_dk(int64_t base, uint64_t k, uint64_t n) {
        pass(IV_MODINV, base, 1, 0, n, k, 0);
        pass(IV_MODINV, base, 1, 0, n, 1, 0);
    }
    // k-way shuffle of the last m - s keys of each partition of m
    // (shuffle_dk_skip_grid_phaseOne/Two, gpu/src/involutions.cuh:213,241)
    void part_shuffle_skip(int64_t base, uint64_t n, uint64_t k, uint64_t m, uint64_t s) {
        pass(IV_MODINV, base + (int64_t)s, n / m, m, m - s, 1, 0);
        pass(IV_MODINV, base + (int64_t)s, n / m, m, m - s, k, 0);
    }
    // rotation right by k by two reversal passes (gpu/src/common.cuh:62-110)
    void shift_right(int64_t base, uint64_t n, uint64_t k) {
        pass(IV_REV2, base, 1, 0, n, 0, 0);
        pass(IV_REV2, base, 1, 0, n, k, 0);
    }
};

static bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

static uint64_t isqrt64(uint64_t m) {
    uint64_t r = (uint64_t)std::sqrt((double)m);
    while (r * r > m) r--;
    while ((r + 1) * (r + 1) <= m) r++;
    return r;
}

// ---- BST (gpu/src/bst-involutions.cuh:71-140, 167-195): the non-full last
// level to the end by an unshuffle, rotations and a shuffle; the perfect part
// then takes the two bit-reversal involutions (phaseOne, phaseTwo), which
// tl_permute_bst runs as Carter-Gatlin tiles.
static void plan_bst_leaves(IEmit &e, uint64_t n, uint64_t &perfect) {
    const uint32_t h = floor_log2_64(n);
    perfect = n;
    if (n == (2ull << h) - 1) return;
    const uint64_t ni = (1ull << h) - 1, nl = n - ni;
    e.unshuffle_dk(0, 2, 2 * nl);
    e.shift_right(0, 2 * nl, nl);
    // the reference's shuffle_dk with k = 1 here applies involution(1, .)
    // twice: the identity, left out
    e.shift_right((int64_t)nl, ni, ni - nl);
    perfect = n - nl;
}

// ---- B-tree (gpu/src/btree-involutions.cuh:30-170): per level a (b+1)-way
// unshuffle of the whole prefix, then a b-way shuffle of its leaf part; a
// non-full last level first goes to the end.  Empty shuffles are left out
// (the reference's loop over [1, n - 1) with n = 0 does not end).
static void plan_btree_perfect(IEmit &e, uint64_t n, uint64_t b, uint32_t d) {
    while (d > 1) {
        e.unshuffle(0, b + 1, n, d);
        e.shuffle_dk((int64_t)(n / (b + 1)), b, ipow64(b + 1, d - 1) * b);
        n /= b + 1;
        d--;
    }
}

static void plan_btree(IEmit &e, uint64_t n, uint64_t b) {
    const uint32_t h = ilog_base(n, b + 1);
    if (n != ipow64(b + 1, h + 1) - 1) {
        const uint64_t ni = ipow64(b + 1, h) - 1, nl = n - ni;
        const uint64_t l = nl - nl % b, i = l / b;
        e.unshuffle_dk(0, b + 1, l + i);
        e.shift_right(0, l + i, i);
        e.shuffle_dk((int64_t)i, b, l);
        e.shift_right((int64_t)i, nl + ni - i, ni - i);
        plan_btree_perfect(e, n - nl, b, h);
    } else {
        plan_btree_perfect(e, n, b, h + 1);
    }
}

// ---- mvEB (gpu/src/mveb-involutions.cuh:30-523)

// rounds of the balanced layout on every partition of m = 2^d - 1 keys (d a
// power of two) of [base, base + n): an unshuffle of the roots to the front and
// a shuffle of the leaves, until the partitions are 255, 15 or 3 keys (the
// base cases, gpu/src/mveb-involutions.cuh:82-217, here the same passes)
static void veb_balanced_passes(IEmit &e, int64_t base, uint64_t n, uint64_t m, uint32_t d) {
    while (d > 8) {
        const uint64_t k = isqrt64(m);
        e.part_unshuffle(base, n, k + 1, m, 2);
        e.part_shuffle_skip(base, n, k, m, k);
        m = k;
        d /= 2;
    }
    if (m == 255) {
        e.part_unshuffle(base, n, 16, 255, 2);
        e.part_shuffle_skip(base, n, 15, 255, 15);
        m = 15;
    }
    if (m == 15) {
        e.part_unshuffle(base, n, 4, 15, 2);
        e.part_shuffle_skip(base, n, 3, 15, 3);
        m = 3;
    }
    if (m == 3) e.pass(IV_REV2, base, n / 3, 3, 2, 0, 0);   // swap the first two of each 3
}

// the same, with the partitions that fit the LDS budget done in LDS groups
static void veb_balanced(IEmit &e, int64_t base, uint64_t n, uint64_t m, uint32_t d) {
    if (n == 0 || m < 3) return;
    if (e.prog) {   // already inside a group
        veb_balanced_passes(e, base, n, m, d);
        return;
    }
    while (d > 8 && m > e.plan->lds_keys) {
        const uint64_t k = isqrt64(m);
        e.part_unshuffle(base, n, k + 1, m, 2);
        e.part_shuffle_skip(base, n, k, m, k);
        m = k;
        d /= 2;
    }
    if (m > e.plan->lds_keys) {   // an LDS budget below 255 keys
        veb_balanced_passes(e, base, n, m, d);
        return;
    }
    IStep st;
    st.lds = true;
    st.base = base;
    st.parts = n / m;
    st.stride = m;
    st.m = (uint32_t)m;
    IEmit g;
    g.prog = &st.prog;
    veb_balanced_passes(g, 0, m, m, d);
    if (!st.prog.empty()) e.plan->steps.push_back(std::move(st));
}

static void veb_perfect(IEmit &e, int64_t base, uint64_t n, uint32_t d) {
    if (n <= 1) return;
    uint32_t rd, ld;
    mveb_split(d, rd, ld);
    const uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
    const uint32_t h = ilog_base(n, l + 1);
    if (n != ipow64(l + 1, h + 1) - 1) {
        // the roots to the front: unshuffle_dk of the n + 1 keys from base - 1
        // (the key before the array is index 0 of the view and never moves)
        e.pass(IV_MODINV, base - 1, 1, 0, n + 1, l + 1, 0);
        e.pass(IV_MODINV, base - 1, 1, 0, n + 1, 1, 0);
    } else {
        e.unshuffle(base, l + 1, n, h + 1);
    }
    e.shuffle_dk(base + (int64_t)r, l, n - r);
    veb_balanced(e, base + (int64_t)r, n - r, l, ld);
    if (!is_pow2(rd))
        veb_perfect(e, base, r, rd);
    else
        veb_balanced(e, base, r, r, rd);
}

static void veb_nonperfect(IEmit &e, int64_t base, uint64_t n, uint32_t d) {
    if (n <= 1) return;
    uint32_t rd, ld;
    mveb_split(d, rd, ld);
    const uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
    const uint64_t nf = (n - r) / l, inc = n - r - nf * l, t = nf * (l + 1);
    e.unshuffle_dk(base, l + 1, t);
    e.shift_right(base, t, nf);
    e.shuffle_dk(base + (int64_t)nf, l, t - nf);
    if (nf < r) e.shift_right(base + (int64_t)nf, n - nf, r - nf);
    if (rd == ld) {
        veb_balanced(e, base, n - inc, r, rd);
    } else {
        veb_balanced(e, base + (int64_t)r, nf * l, l, ld);
        if (!is_pow2(rd))
            veb_perfect(e, base, r, rd);
        else
            veb_balanced(e, base, r, r, rd);
    }
    if (inc > 0) {
        const uint32_t id = floor_log2_64(inc) + 1;
        const int64_t ib = base + (int64_t)(n - inc);
        if (inc != (1ull << id) - 1)
            veb_nonperfect(e, ib, inc, id);
        else if (!is_pow2(id))
            veb_perfect(e, ib, inc, id);
        else
            veb_balanced(e, ib, inc, inc, id);
    }
}

static void plan_mveb(IEmit &e, uint64_t n) {
    const uint32_t d = floor_log2_64(n) + 1;
    if (n != (1ull << d) - 1)
        veb_nonperfect(e, 0, n, d);
    else if (is_pow2(d))
        veb_balanced(e, 0, n, n, d);
    else
        veb_perfect(e, 0, n, d);
}

// ---------------------------------------------------------------------------
// host: launches

template <typename T>
static tl_status launch_pass(Ctx *c, T *A, const IPass &p, hipStream_t s) {
    const uint64_t total = p.parts * p.len;
    const uint32_t g =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, c->num_cus * 32ull));
    switch (p.kind) {
    case IV_DIGITS:
        hipLaunchKernelGGL((k_involution<T, IV_DIGITS>), dim3(g), dim3(256), 0, s, A, p);
        break;
    case IV_MODINV:
        hipLaunchKernelGGL((k_involution<T, IV_MODINV>), dim3(g), dim3(256), 0, s, A, p);
        break;
    default:
        hipLaunchKernelGGL((k_involution<T, IV_REV2>), dim3(g), dim3(256), 0, s, A, p);
        break;
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// every LDS group's passes in one table, uploaded before the launches
static tl_status upload_groups(Ctx *c, const IPlan &pl, hipStream_t s, IPass **dev_table,
                               std::vector<size_t> &at) {
    std::vector<IPass> table;
    for (const IStep &st : pl.steps) {
        at.push_back(table.size());
        if (st.lds) table.insert(table.end(), st.prog.begin(), st.prog.end());
    }
    *dev_table = nullptr;
    if (!table.empty())
        TL_TRY(copy_to_device(c, 2, table.data(), table.size() * sizeof(IPass), s,
                              (void **)dev_table));
    return TL_OK;
}

template <typename T>
static tl_status run_iplan(Ctx *c, const IPlan &pl, T *A, const IPass *dev_table,
                           const std::vector<size_t> &at, hipStream_t s) {
    for (size_t i = 0; i < pl.steps.size(); i++) {
        const IStep &st = pl.steps[i];
        if (!st.lds) {
            TL_TRY(launch_pass<T>(c, A, st.pass, s));
            continue;
        }
        const uint32_t g = (uint32_t)std::min<uint64_t>(st.parts, c->num_cus * 16ull);
        hipLaunchKernelGGL(k_involution_lds<T>, dim3(g), dim3(256), st.m * sizeof(T), s, A,
                           st.base, st.parts, st.stride, st.m, dev_table + at[i],
                           (uint32_t)st.prog.size());
        TL_LAUNCH_CHECK();
    }
    return TL_OK;
}

static tl_status width(tl_dtype dt, uint32_t &W) {
    if (dt == TL_U32 || dt == TL_I32) {
        W = 4;
    } else if (dt == TL_U64 || dt == TL_I64) {
        W = 8;
    } else {
        set_error("unknown dtype");
        return TL_EINVAL;
    }
    return TL_OK;
}

static tl_status run_timed(IPlan pl, void *dev_A, uint32_t W, void *stream, float *ms,
                           uint64_t bst_perfect = 0, tl_dtype dt = TL_U64) {
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    // partner tables: one per pass shape (kind, digits, length, base or
    // multiplier) that repeats over partitions or runs in LDS; reversals
    // stay computed
    std::map<std::tuple<uint32_t, uint32_t, uint64_t, uint64_t>, uint64_t> shapes;
    std::vector<IPass> reps;
    uint64_t entries = 0;
    auto want = [&](IPass &p) {
        if (p.kind == IV_REV2 || p.len > IV_TABLE_MAX) return;
        const auto key = std::make_tuple(p.kind, p.nd, p.len, p.k);
        auto it = shapes.find(key);
        if (it == shapes.end()) {
            it = shapes.emplace(key, entries).first;
            reps.push_back(p);
            reps.back().tab = entries;
            entries += p.len;
        }
        p.tab = it->second + 1;   // entry offset + 1 until the tables are placed
    };
    for (IStep &st : pl.steps) {
        if (st.lds)
            for (IPass &p : st.prog) want(p);
        else if (st.pass.parts > 1)
            want(st.pass);
    }
    uint32_t *tabs = nullptr;
    if (entries) {
        TL_TRY(scratch(c, 3, entries * 4, (void **)&tabs));
        auto place = [&](IPass &p) {
            if (p.tab) p.tab = (uint64_t)(tabs + (p.tab - 1));
        };
        for (IStep &st : pl.steps) {
            place(st.pass);
            for (IPass &p : st.prog) place(p);
        }
    }
    IPass *dev_table;
    std::vector<size_t> at;
    TL_TRY(upload_groups(c, pl, s, &dev_table, at));
    Timer tm;
    TL_TRY(tm.start(s, ms));
    for (const IPass &r : reps) {
        IPass q = r;
        q.tab = 0;
        const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((r.len + 255) / 256, 4096));
        hipLaunchKernelGGL(k_partner_table, dim3(g), dim3(256), 0, s, q, tabs + r.tab);
        TL_LAUNCH_CHECK();
    }
    TL_TRY(W == 4 ? run_iplan<uint32_t>(c, pl, (uint32_t *)dev_A, dev_table, at, s)
                  : run_iplan<uint64_t>(c, pl, (uint64_t *)dev_A, dev_table, at, s));
    if (bst_perfect > 1) TL_TRY(tl_permute_bst(dev_A, bst_perfect, dt, stream, nullptr));
    TL_TRY(tm.stop());
    return TL_OK;
}

static tl_status check_array(const void *dev_A, uint64_t n, tl_dtype dt, uint32_t &W) {
    TL_TRY(width(dt, W));
    if (n > 1 && !dev_A) {
        set_error("null array");
        return TL_EINVAL;
    }
    if (n >= (1ull << 40)) {
        set_error("n too large");
        return TL_EINVAL;
    }
    return TL_OK;
}

static IPlan new_plan(uint32_t W) {
    IPlan pl;
    Ctx *c = nullptr;
    const uint64_t budget = (get_ctx(&c) == TL_OK && c) ? c->lds_max : 65536;
    pl.lds_keys = std::min<uint64_t>(budget, 64u << 10) / W;
    return pl;
}

}   // namespace tl

using namespace tl;

extern "C" tl_status tl_shuffle_dk(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype,
                                   void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    if (k == 0 || n % k) {
        set_error("tl_shuffle_dk: n must be a multiple of k >= 1");
        return TL_EINVAL;
    }
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    e.shuffle_dk(0, k, n);
    return run_timed(pl, dev_A, W, stream, ms);
}

extern "C" tl_status tl_unshuffle_dk(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype,
                                     void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    if (k == 0 || n % k) {
        set_error("tl_unshuffle_dk: n must be a multiple of k >= 1");
        return TL_EINVAL;
    }
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    e.unshuffle_dk(0, k, n);
    return run_timed(pl, dev_A, W, stream, ms);
}

extern "C" tl_status tl_unshuffle(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype,
                                  void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    const uint32_t d = k >= 2 ? ilog_base(n + 1, k) : 0;
    if (k < 2 || n == 0 || ipow64(k, d) - 1 != n) {
        set_error("tl_unshuffle: n must be k^d - 1 with k >= 2, d >= 1");
        return TL_EINVAL;
    }
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    e.unshuffle(0, k, n, d);
    return run_timed(pl, dev_A, W, stream, ms);
}

extern "C" tl_status tl_shift_right(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype,
                                    void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    if (k > n) {
        set_error("tl_shift_right: k must be <= n");
        return TL_EINVAL;
    }
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    e.shift_right(0, n, k);
    return run_timed(pl, dev_A, W, stream, ms);
}

extern "C" tl_status tl_permute_bst_involutions(void *dev_A, uint64_t n, tl_dtype dtype,
                                                void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    uint64_t perfect = n;
    if (n > 1) plan_bst_leaves(e, n, perfect);
    return run_timed(pl, dev_A, W, stream, ms, n > 1 ? perfect : 0, dtype);
}

extern "C" tl_status tl_permute_btree_involutions(void *dev_A, uint64_t n, uint64_t b,
                                                  tl_dtype dtype, void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    if (b == 0 || b > 4096) {
        set_error("b must be in [1, 4096]");
        return TL_EINVAL;
    }
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    if (n > 1) plan_btree(e, n, b);
    return run_timed(pl, dev_A, W, stream, ms);
}

extern "C" tl_status tl_permute_mveb_involutions(void *dev_A, uint64_t n, tl_dtype dtype,
                                                 void *stream, float *ms) {
    uint32_t W;
    TL_TRY(check_array(dev_A, n, dtype, W));
    IPlan pl = new_plan(W);
    IEmit e;
    e.plan = &pl;
    if (n > 1) plan_mveb(e, n);
    return run_timed(pl, dev_A, W, stream, ms);
}
