// tl_plan.hpp -- host planner of the in-place layout permutations, shared by
// the single-array entry points and the uneven-batch engine.
//
// A plan is a list of launches (Op) over "runs" of equally sized, equally
// spaced sub-arrays; its item lists and cycle tables live in one blob.
#pragma once

#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tl_internal.hpp"
#include "tl_kernels.hpp"

namespace tl {

// ---------------------------------------------------------------------------
// item runs: item i of a run lives at start + i*stride

struct Run {
    uint64_t start, stride, count;
};

__device__ inline uint64_t run_offset(const uint64_t *runs, uint32_t nruns, uint64_t idx) {
    uint32_t lo = 0, hi = nruns;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (runs[4 * mid + 3] <= idx)
            lo = mid;
        else
            hi = mid;
    }
    return runs[4 * lo] + (idx - runs[4 * lo + 3]) * runs[4 * lo + 1];
}

// ---------------------------------------------------------------------------
// plan

enum OpType {
    OP_LDS,
    OP_REVERSE,
    OP_BITREV,
    OP_UNIT_MERGE,
    OP_SPARSE_STREAM,
    OP_COPY,
    OP_CHAINED,
    OP_BITREV_LEVELS,  // the per-level reversals of levels p[0] .. p[1]-1, tiles 2^p[2]
    OP_ROTATE,         // [p0, p0+p1) rotated right by p2 in one pass, p3 lanes (stash)
    OP_HEADS,          // the heads of levels p[0] .. p[1]-1 (tiles 2^p[2]) to the stash
    OP_PART_ROOTS      // the p[3] roots of p[2] blocks at p[0] + i*p[1] from stash[p[4] ..]
                       // to their vEB slots (p[5]: 1 mvEB, 2 classic; p[6]: inverse)
};

// A BST of 2^D keys has one leaf, the smallest key, which goes last; every
// other key's final place is one left of where the tree kept one place right
// (slot y at A[y]) has it, and that is where the sorted keys already are.
// The plan keeps the tree there until its last phase, which writes every
// level one slot left.  Forms of the bit-reversal launches (p[3]; tiles of
// 2^K keys a side, t0 = 2K = p[4] for OP_BITREV):
//   OP_BITREV, base 0: BITREV_HEAD (phase one of the kept tree: plain; its
//     inverse is BITREV_ABSORB) or BITREV_ABSORB (the row ends at slots
//     >= 2^t0 read from the stash, where BITREV_UNFOLD put them);
//   OP_BITREV_LEVELS: BITREV_SHIFT (every level reversed in place and
//     written one slot left; the first slots of the rows after each tile's
//     own come from the heads that OP_HEADS put in the stash) or
//     BITREV_UNFOLD (the inverse: level t read from A[2^t - 1 ..] and
//     written to A[2^t ..], the row ends to the stash).
// OP_HEADS has no inverse (the inverse plan drops it).  OP_LDS of kind
// LDS_LEVELREV with p[6] = 1 reads its keys one place right (inverse: writes
// them one place right).
enum : uint64_t {
    BITREV_PLAIN = 0,
    BITREV_SHIFT = 1,
    BITREV_UNFOLD = 2,
    BITREV_HEAD = 3,
    BITREV_ABSORB = 4
};

// stash slots of the row ends of levels t0 .. D-1 (rows of 2^K keys)
inline uint64_t rowend_slots(uint32_t D, uint32_t K, uint32_t t0) {
    return ((1ull << D) - (1ull << t0)) >> K;
}

struct Op {
    OpType type;
    uint64_t runs_at = 0, nruns = 0, nitems = 0;   // item runs in the blob
    uint64_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};      // op parameters
    MoveSpec ms{};
};

struct Plan {
    std::vector<Op> ops;
    std::vector<uint64_t> blob;
    uint64_t max_stash = 0;   // elements
    uint64_t max_moves = 0;

    void put_runs(Op &op, const std::vector<Run> &runs) {
        op.runs_at = blob.size();
        op.nruns = runs.size();
        uint64_t pre = 0;
        for (const Run &r : runs) {
            blob.push_back(r.start);
            blob.push_back(r.stride);
            blob.push_back(r.count);
            blob.push_back(pre);
            pre += r.count;
        }
        op.nitems = pre;
    }
    uint64_t put(const std::vector<uint64_t> &v) {
        uint64_t at = blob.size();
        blob.insert(blob.end(), v.begin(), v.end());
        return at;
    }
};

// bytes of chained-move flags (and the ticket behind them) a plan needs
inline size_t plan_flag_bytes(const Plan &pl) {
    return (((pl.max_moves + 15) / 16) * 16 + 16) * 4;
}

// bytes of flags when every chained op of the plan has a region of its own
inline size_t plan_flag_bytes_all(const Plan &pl) {
    size_t words = 0;
    for (const Op &op : pl.ops)
        if (op.type == OP_CHAINED) words += ((op.p[0] + 15) / 16) * 16 + 16;
    return words * 4;
}

// launch a plan whose blob is already on the device, with caller-owned stash
// and flags (plan_flag_bytes_all bytes, zeroed by the caller), so that the
// launches can be captured
tl_status plan_launch(Ctx *c, const Plan &pl, void *A, uint32_t W, hipStream_t s,
                      const uint64_t *dev_blob, void *stash, uint32_t *flags);
// p[0, count) = v by a kernel
tl_status fill_u32_launch(uint32_t *p, uint64_t count, uint32_t v, hipStream_t s);
// whole arrays of at most max_n keys in LDS, one workgroup per (address, n)
tl_status lds_items_launch(const uint64_t *dev_desc, uint32_t count, uint32_t W, uint32_t kind,
                           uint32_t b, uint32_t max_n, hipStream_t s);

// ---- multi-array launches: one launch runs the same kind of op for many
//      independent arrays (the uneven batch runs a plan's k-th op of every
//      array together).  Descriptors hold absolute device addresses; `pre`
//      is the prefix sum of each descriptor's workgroup (or move) count.
struct MCopy {   // dst[i*dstride] = src[i*sstride], i < count, keys of w bytes
    uint64_t dst, src, dstride, sstride, count, w;
};
struct MLds {    // one OP_LDS of one array, keys of W bytes
    uint64_t A, runs, nitems;
    uint32_t nruns, per, n, kind, b, rows, magic, inv, shift, W;
};
struct MStream {   // one OP_SPARSE_STREAM of one array, keys of W bytes
    uint64_t A, runs;
    uint32_t nruns, r, l, inv, fuse, W;
};
struct MChain {    // one OP_CHAINED of one array, keys of W bytes
    uint64_t A, flags, stash;
    MoveSpec ms;
    uint64_t W;
};
struct MBitrev {   // OP_BITREV (levels = 0) or OP_BITREV_LEVELS (levels = 1)
    uint64_t A;
    int64_t base;
    uint32_t D, t0, d, levels;
    uint64_t stash;   // BITREV_SHIFT levels: the array's level heads
    uint64_t W;       // key width in bytes
};
struct MWalk {     // the lane walk of one OP_ROTATE (its stash copied before); keys of W bytes
    uint64_t A, stash, L, c, w, W;
};
struct MPRoots {   // one OP_PART_ROOTS of one array (stash: its first root), keys of W bytes
    uint64_t A, stash, base, stride, count;
    uint32_t rp, classic, inv, W;
};

enum MultiKind : uint32_t {
    MK_COPY,
    MK_LDS,
    MK_STREAM,
    MK_FUSED,
    MK_CHAIN,
    MK_BITREV,
    MK_WALK,
    MK_PROOTS
};

struct MultiLaunch {
    uint32_t kind, W, param;        // W: 0 (descriptors carry their key width); param:
                                    // chained mode, or bitrev K
    uint32_t nd;                    // descriptors
    uint64_t desc, pre;             // device addresses of descriptors and prefix sums
    uint64_t tasks;                 // workgroups (chained: moves)
    uint64_t counter;               // chained: ticket (zeroed with the flags)
    size_t lds;                     // dynamic LDS bytes
    uint32_t threads;
};
tl_status multi_launch(Ctx *c, const MultiLaunch &m, hipStream_t s);

// LDS kept free for kernels' static shared variables next to the dynamic area
constexpr size_t LDS_STATIC_RESERVE = 2048;

struct Tuning {
    size_t lds_bytes = 65536;   // LDS budget of one in-LDS layout workgroup
    uint64_t fmax = 256;        // largest merge radix
    uint64_t stash_max = 1ull << 24;   // elements a rotation may stash
    uint64_t row_stash_max = 1ull << 24;   // separators a row slide may stash
    int64_t row_stash_env = -1;            // $TL_ROW_STASH_MAX when set (0: no row slides)
    uint64_t pack_bytes = 32768;           // keys of several small in-LDS items per workgroup
    uint32_t bitrev_k = 0;                 // bit-reversal tile 2^K x 2^K (0: 5, both key widths)
    bool rotate_walk = true;               // one-pass rotations (else three reversals)
    bool fuse_leaves = true;               // vEB leaf layouts fused into streamed gathers
    uint64_t walk_lanes = 1ull << 24;      // lanes (stashed keys) of a one-pass rotation
    int64_t walk_env = -1;                 // $TL_WALK_LANES when set (0: no one-pass rotations)
    bool bst_fold = true;                  // BST of 2^D keys: kept one place right
    bool fuse_slide = true;                // vEB: leaf blocks' gathers ride in the sparse slide
};

inline void tuning_env(Tuning &t) {
    const char *r = getenv("TL_ROW_STASH_MAX");
    if (r) t.row_stash_env = atoll(r);
    const char *k = getenv("TL_BITREV_K");
    if (k && (atoi(k) == 5 || atoi(k) == 6)) t.bitrev_k = (uint32_t)atoi(k);
    const char *w = getenv("TL_WALK_LANES");
    if (w) t.walk_env = atoll(w);
    const char *f = getenv("TL_BST_FOLD");
    if (f && atoi(f) == 0) t.bst_fold = false;
    const char *fs = getenv("TL_FUSE_SLIDE");
    if (fs && atoi(fs) == 0) t.fuse_slide = false;
}

inline uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

inline Tuning tuning_for(Ctx *c) {
    Tuning t;
    tuning_env(t);
    const char *e = getenv("TL_LDS_BYTES");
    size_t want = e ? (size_t)atoll(e) : 163840;
    // the in-LDS kernel keeps 512 bytes of static LDS beside the dynamic area
    t.lds_bytes = std::min(want, c->lds_max - LDS_STATIC_RESERVE);
    const char *f = getenv("TL_FMAX");
    if (f) t.fmax = std::max<uint64_t>(2, atoll(f));
    return t;
}

// cycles of the unit permutation of one dense-gather merge: F child
// partitions, each [1 separator unit][b leaf units] -> [F separator units]
// [F*b leaf units].  Fixed points are dropped.
struct Cycles {
    std::vector<uint64_t> start, units;
};

inline const Cycles &merge_cycles(uint64_t F, uint64_t b) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, uint64_t>, Cycles> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(F, b);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const uint64_t N = F * (b + 1);
    std::vector<uint8_t> seen(N, 0);
    Cycles cy;
    for (uint64_t v0 = 0; v0 < N; v0++) {
        if (seen[v0]) continue;
        uint64_t v = v0, len = 0;
        const size_t at = cy.units.size();
        do {
            seen[v] = 1;
            cy.units.push_back(v);
            len++;
            const uint64_t j = v / (b + 1), c = v % (b + 1);
            v = (c == 0) ? j : F + j * b + c - 1;
        } while (v != v0);
        if (len == 1)
            cy.units.resize(at);
        else
            cy.start.push_back(at);
    }
    cy.start.push_back(cy.units.size());
    return cache.emplace(key, std::move(cy)).first->second;
}

// ---------------------------------------------------------------------------
// planners

struct Planner {
    Plan &pl;
    Tuning tu;
    uint32_t W;   // key width in bytes

    uint64_t cap() const { return tu.lds_bytes / W; }

    void lds(const std::vector<Run> &runs, uint64_t n, uint32_t kind, uint64_t b = 1,
             uint64_t rows = 0) {
        if (runs.empty() || n <= 1) return;
        Op op;
        op.type = OP_LDS;
        pl.put_runs(op, runs);
        if (op.nitems == 0) return;
        // small items share a workgroup up to tu.pack_bytes of keys, so that
        // several workgroups fit a compute unit
        const uint64_t pack = std::min<uint64_t>(tu.lds_bytes, tu.pack_bytes) / W;
        uint64_t per = std::max<uint64_t>(1, pack / n);
        if (per > 64) per = 64;
        op.p[0] = n;
        op.p[1] = kind;
        op.p[2] = b;
        op.p[3] = rows;
        op.p[4] = per;
        pl.ops.push_back(op);
    }

    // rotate right by k every region [start, start+len) of `runs`
    void rotate_right(const std::vector<Run> &runs, uint64_t len, uint64_t k) {
        if (runs.empty() || len < 2) return;
        k %= len;
        if (k == 0) return;
        uint64_t items = 0;
        for (auto &r : runs) items += r.count;
        const uint64_t small = std::min(k, len - k);
        if (items == 1 && small <= tu.stash_max && len > 4 * small) {
            const uint64_t P = runs[0].start;
            const uint64_t CH = CHAINED_KEYS;
            if (k <= len - k) {
                // [X][Y] -> [Y][X] with Y small: stash Y, slide X right by k
                copy(P + len - k, 1, k, /*to_stash=*/true);
                chained_shift(P, len - k, k, true, CH);
                copy(P, 1, k, /*to_stash=*/false);
            } else {
                // X = [P, P+len-k) small: stash X, slide Y left by len-k
                const uint64_t x = len - k;
                copy(P, 1, x, true);
                chained_shift(P + x, k, x, false, CH);
                copy(P + k, 1, x, false);
            }
            return;
        }
        // one pass: w lanes start at the first w keys (stashed), each walks
        // its part of a cycle of p -> p + k (mod len); every cycle needs a
        // start, so w >= gcd(len, k).  About 16 steps per lane keep the stash
        // traffic at 1/8 of a pass.
        if (items == 1 && tu.rotate_walk) {
            const uint64_t g = gcd64(len, k);
            const uint64_t w = std::min<uint64_t>(tu.walk_lanes, std::max<uint64_t>(len / 16, g));
            if (w >= g && w >= 1) {
                Op op;
                op.type = OP_ROTATE;
                op.p[0] = runs[0].start;
                op.p[1] = len;
                op.p[2] = k;
                op.p[3] = w;
                pl.max_stash = std::max(pl.max_stash, w);
                pl.ops.push_back(op);
                return;
            }
        }
        // three reversals: whole, then the two parts
        Op a;
        a.type = OP_REVERSE;
        pl.put_runs(a, runs);
        a.p[0] = len;
        pl.ops.push_back(a);
        std::vector<Run> r1 = runs, r2 = runs;
        for (auto &r : r2) r.start += k;
        Op b1;
        b1.type = OP_REVERSE;
        pl.put_runs(b1, r1);
        b1.p[0] = k;
        pl.ops.push_back(b1);
        Op b2;
        b2.type = OP_REVERSE;
        pl.put_runs(b2, r2);
        b2.p[0] = len - k;
        pl.ops.push_back(b2);
    }

    // stash (to_stash) or unstash `count` keys at P + i*stride, i < count,
    // to / from stash[soff + i]
    void copy(uint64_t P, uint64_t stride, uint64_t count, bool to_stash, uint64_t soff = 0) {
        Op op;
        op.type = OP_COPY;
        op.p[0] = P;
        op.p[1] = stride;
        op.p[2] = count;
        op.p[3] = to_stash;
        op.p[4] = soff;
        pl.max_stash = std::max(pl.max_stash, soff + count);
        pl.ops.push_back(op);
    }

    void chained_shift(uint64_t P, uint64_t L, uint64_t k, bool right, uint64_t CH) {
        Op op;
        op.type = OP_CHAINED;
        op.ms.kind = right ? MV_SHIFT_RIGHT : MV_SHIFT_LEFT;
        op.ms.P = P;
        op.ms.L = L;
        op.ms.k = k;
        op.ms.CH = CH;
        op.ms.nchunk = (L + CH - 1) / CH;
        op.p[0] = op.ms.nchunk;
        pl.max_moves = std::max(pl.max_moves, op.p[0]);
        pl.ops.push_back(op);
    }

    // ---- dense gather: every item of `runs` holds M rows of b leaf keys and
    //      one separator -> [M separators][M*b leaf keys]
    // one chained pass over r rows of (l keys, 1 separator) at P: separators
    // to the stash, keys slid right into place, separators back to the front
    void row_slide(uint64_t P, uint64_t r, uint64_t l) {
        Op op;
        op.type = OP_CHAINED;
        op.ms.kind = MV_ROW_SLIDE;
        op.ms.P = P;
        op.ms.r = r;
        op.ms.l = l;
        op.ms.CH = std::max<uint64_t>(1, (CHAINED_KEYS) / (l + 1));
        op.ms.nchunk = (r + op.ms.CH - 1) / op.ms.CH;
        op.p[0] = op.ms.nchunk;
        pl.max_moves = std::max(pl.max_moves, op.p[0]);
        pl.max_stash = std::max(pl.max_stash, r);
        pl.ops.push_back(op);
        copy(P, 1, r, /*to_stash=*/false);
    }

    void gather_dense(const std::vector<Run> &runs, uint64_t M, uint64_t b) {
        if (runs.empty() || M == 0) return;
        const uint64_t k1 = b + 1;
        // a single row always goes to LDS, whatever the budget
        if (M * k1 <= cap() || M == 1) {
            lds(runs, M * k1, LDS_GATHER, b, M);
            return;
        }
        // one large region whose separators fit the stash: one chained pass
        if (runs.size() == 1 && runs[0].count == 1 && k1 <= 2048 && M <= tu.row_stash_max) {
            row_slide(runs[0].start, M, b);
            return;
        }
        const uint64_t cap_r = std::max<uint64_t>(1, cap() / k1);
        const uint64_t tiles = (M + cap_r - 1) / cap_r;
        const double lv = ceil(log((double)tiles) / log((double)tu.fmax) - 1e-9);
        const uint64_t levels = std::max<uint64_t>(1, (uint64_t)lv);
        uint64_t F = (uint64_t)ceil(pow((double)tiles, 1.0 / (double)levels) - 1e-9);
        F = std::max<uint64_t>(2, std::min(F, M));
        const uint64_t q = (M + F - 1) / F;
        const uint64_t Fp = M / q;
        const uint64_t qf = M - Fp * q;
        // children of every parent item: a front partition of qf rows and Fp
        // partitions of q rows.  A run of parents yields either one run per
        // parent (children equally spaced inside it) or one run per child slot
        // (children equally spaced across parents), whichever is shorter.
        std::vector<Run> full, front, parents;
        for (const Run &r : runs) {
            if (r.count <= Fp) {
                for (uint64_t i = 0; i < r.count; i++)
                    full.push_back({r.start + i * r.stride + qf * k1, q * k1, Fp});
            } else {
                for (uint64_t j = 0; j < Fp; j++)
                    full.push_back({r.start + (qf + j * q) * k1, r.stride, r.count});
            }
            front.push_back({r.start, r.stride, r.count});
            parents.push_back({r.start + qf * k1, r.stride, r.count});
        }
        gather_dense(full, q, b);
        if (qf) gather_dense(front, qf, b);
        // merge the Fp children of every parent
        const Cycles &cy = merge_cycles(Fp, b);
        if (cy.start.size() > 1) {
            Op op;
            op.type = OP_UNIT_MERGE;
            pl.put_runs(op, parents);
            op.p[0] = q;                       // unit
            op.p[1] = pl.put(cy.start);        // cycle starts
            op.p[2] = cy.start.size() - 1;     // cycles
            op.p[3] = pl.put(cy.units);        // cycle units
            pl.ops.push_back(op);
        }
        // [S_f][L_f][S_*][L_*] -> [S_f][S_*][L_f][L_*]: rotate [L_f][S_*] left
        if (qf) {
            std::vector<Run> mid;
            for (const Run &r : runs) mid.push_back({r.start + qf, r.stride, r.count});
            const uint64_t len = qf * b + Fp * q;
            rotate_right(mid, len, Fp * q);
        }
    }

    // ---- sparse gather for the mvEB: every item holds r rows of (l leaf keys,
    //      1 root) followed by untouched keys
    // returns true when the kids' layouts were fused into the gather (fuse:
    // 1 mvEB, 2 classic vEB; only for a streamed gather of short rows whose
    // kids both fit in LDS)
    bool gather_sparse(const std::vector<Run> &runs, uint64_t r, uint64_t l, uint32_t fuse = 0) {
        if (runs.empty() || r == 0) return false;
        uint64_t items = 0;
        for (auto &x : runs) items += x.count;
        const uint64_t part = r * (l + 1);
        const bool roots_fit = r * W <= 65536;
        // one workgroup streams each partition: worth it with enough
        // partitions to fill the device, or partitions short enough
        if (roots_fit && (items >= 64 || part <= 131072)) {
            Op op;
            op.type = OP_SPARSE_STREAM;
            pl.put_runs(op, runs);
            op.p[0] = r;
            op.p[1] = l;
            const bool fused = fuse && tu.fuse_leaves && l + 1 <= 2048 && r <= cap() && l <= cap();
            op.p[3] = fused ? fuse : 0;
            pl.ops.push_back(op);
            return fused;
        }
        for (const Run &x : runs)
            for (uint64_t i = 0; i < x.count; i++) sparse_slide(x.start + i * x.stride, r, l, 0, 0);
        return false;
    }

    // one chained pass over r rows of (l leaf keys, 1 root) at P, followed by
    // `tail` keys and then `extra` roots: the roots go to the stash (row roots
    // first, then the extra roots), leaf blocks and tail slide right by `extra`
    // more than a plain gather, and the stash comes back as the first r+extra
    // keys.  With tail = extra = 0 this is the sparse gather.
    // fuse (1 mvEB, 2 classic vEB): every leaf block is a perfect tree of rp
    // rows of (lp keys, 1 root) and a last block of lp keys, whose own gather
    // and whose lp-key blocks' layouts ride in the same moves; its roots go
    // to the stash behind the slide's own and are laid out by OP_PART_ROOTS.
    void sparse_slide(uint64_t P, uint64_t r, uint64_t l, uint64_t tail, uint64_t extra,
                      uint32_t fuse = 0, uint64_t rp = 0, uint64_t lp = 0) {
        copy(P + l, l + 1, r, true);                               // row roots
        if (extra) copy(P + r * (l + 1) + tail, 1, extra, true, r);   // trailing roots
        Op op;
        op.type = OP_CHAINED;
        op.ms.kind = MV_SPARSE_SLIDE;
        op.ms.P = P;
        op.ms.r = r;
        op.ms.l = l;
        op.ms.L = tail;
        op.ms.k = extra;
        op.ms.CH = std::min<uint64_t>(l, CHAINED_KEYS);
        if (fuse) {   // whole rows per move
            op.ms.fuse = fuse;
            op.ms.rp = rp;
            op.ms.lp = lp;
            op.ms.soff = r + extra;
            op.ms.CH = std::min<uint64_t>(CHAINED_KEYS / (lp + 1), rp) * (lp + 1);
            pl.max_stash = std::max(pl.max_stash, op.ms.soff + r * rp);
        }
        op.ms.nchunk = (l + op.ms.CH - 1) / op.ms.CH;
        op.p[0] = (tail + op.ms.CH - 1) / op.ms.CH + r * op.ms.nchunk;
        pl.max_moves = std::max(pl.max_moves, op.p[0]);
        pl.ops.push_back(op);
        copy(P, 1, r + extra, false);   // roots to the front
        if (fuse) {
            Op pr;
            pr.type = OP_PART_ROOTS;
            pr.p[0] = P + r + extra;
            pr.p[1] = l;
            pr.p[2] = r;
            pr.p[3] = rp;
            pr.p[4] = r + extra;
            pr.p[5] = fuse;
            pl.ops.push_back(pr);
        }
    }
};

// ---- B-tree (gpu/src/btree-cycles.cuh:62-104) -----------------------------
inline void plan_btree(Planner &P, uint64_t n, uint64_t b) {
    if (n <= b) return;   // one node, already sorted
    if (n <= P.cap()) {
        P.lds({{0, 0, 1}}, n, LDS_BTREE, b);
        return;
    }
    const uint64_t k1 = b + 1;
    const uint32_t h = ilog_base(n, k1);
    const uint64_t Pw = ipow64(k1, h);
    if (n != Pw * k1 - 1) {
        const uint64_t nint = Pw - 1, nleaf = n - nint;
        const uint64_t rem = nleaf % b, l = nleaf - rem, i = l / b;
        P.gather_dense({{0, 0, 1}}, i, b);
        P.rotate_right({{i, 0, 1}}, nleaf + nint - i, nint - i);
        n = nint;
    }
    while (n > b) {
        if (n <= P.cap()) {
            P.lds({{0, 0, 1}}, n, LDS_BTREE, b);
            break;
        }
        P.gather_dense({{0, 0, 1}}, n / k1, b);
        n /= k1;
    }
}

// ---- BST (gpu/src/bst-hybrid.cuh:95-148) ----------------------------------
inline void plan_bst(Planner &P, uint64_t n) {
    if (n <= 1) return;
    if (n <= P.cap()) {
        P.lds({{0, 0, 1}}, n, LDS_BTREE, 1);
        return;
    }
    const uint32_t h = floor_log2_64(n);
    const uint32_t K = P.tu.bitrev_k ? P.tu.bitrev_k : 5u;
    // A non-perfect tree: 2^h - 1 internal keys and L >= 1 leaves.  The plan
    // keeps the internal tree one place right, behind the smallest leaf:
    // [leaf 0 | internal keys | leaves 1 .. L-1].  Key 1 is internal and the
    // pairs (leaf, internal) from key 2 on are gathered (internals first),
    // then the remaining internal keys rotate in front of the L - 1 leaves.
    // For n = 2^h (L = 1) both are empty: the sorted keys are already there,
    // and the leaf's move to the end, a whole pass as a gather, folds into
    // the last phase.
    const uint64_t nint_f = (1ull << h) - 1;
    if (P.tu.bst_fold && n != (1ull << (h + 1)) - 1 && nint_f > P.cap() && h >= 2 * K + 2) {
        const uint64_t L = n - nint_f;
        if (L > 1) {
            P.gather_dense({{2, 0, 1}}, L - 1, 1);
            P.rotate_right({{L + 1, 0, 1}}, n - L - 1, nint_f - L);
        }
        const uint32_t t0 = 2 * K;
        const uint64_t leaf = rowend_slots(h, K, t0);   // the leaf's stash slot
        Op one;
        one.type = OP_BITREV;
        one.p[0] = 0;
        one.p[1] = h;
        one.p[2] = K;
        one.p[3] = BITREV_HEAD;
        one.p[4] = t0;
        P.pl.max_stash = std::max<uint64_t>(P.pl.max_stash, leaf + 1);
        P.pl.ops.push_back(one);
        Op heads;
        heads.type = OP_HEADS;
        heads.p[0] = t0;
        heads.p[1] = h;
        heads.p[2] = K;
        P.pl.ops.push_back(heads);
        P.copy(0, 1, 1, true, leaf);
        P.lds({{0, 0, 1}}, (1ull << t0) - 1, LDS_LEVELREV);
        P.pl.ops.back().p[6] = 1;   // read one place right
        Op two;
        two.type = OP_BITREV_LEVELS;
        two.p[0] = t0;
        two.p[1] = h;
        two.p[2] = K;
        two.p[3] = BITREV_SHIFT;
        P.pl.ops.push_back(two);
        P.copy(nint_f, 1, 1, false, leaf);   // leaf 0 in front of the others
        return;
    }
    uint32_t d;
    if (n != (1ull << (h + 1)) - 1) {
        const uint64_t nint = (1ull << h) - 1, nleaf = n - nint;
        P.gather_dense({{0, 0, 1}}, nleaf, 1);
        P.rotate_right({{nleaf, 0, 1}}, nint, nint - nleaf);
        n = nint;
        d = h;
    } else {
        d = h + 1;
    }
    if (n <= P.cap() || d < 2 * K + 2) {
        P.lds({{0, 0, 1}}, n, LDS_BTREE, 1);
        return;
    }
    Op one;
    one.type = OP_BITREV;
    one.p[0] = (uint64_t)(int64_t)-1;
    one.p[1] = d;
    one.p[2] = K;
    P.pl.ops.push_back(one);
    if (d > 2 * K) {   // phase two: every level from 2K on, in one launch
        Op two;
        two.type = OP_BITREV_LEVELS;
        two.p[0] = 2 * K;
        two.p[1] = d;
        two.p[2] = K;
        P.pl.ops.push_back(two);
    }
    P.lds({{0, 0, 1}}, std::min<uint64_t>(n, (1ull << (2 * K)) - 1), LDS_LEVELREV);
}

// ---- modified vEB (gpu/src/mveb-cycles.cuh:114-440) -----------------------
struct VebBatch {
    uint64_t n;
    std::vector<Run> runs;
};

inline void plan_mveb_rec(Planner &P, std::vector<std::vector<Op>> &depth_ops,
                          std::vector<VebBatch> &leaves_lds, VebBatch B, uint32_t depth,
                          bool classic) {
    if (B.n <= 1 || B.runs.empty()) return;
    if (B.n <= P.cap()) {
        leaves_lds.push_back(B);
        return;
    }
    if (depth_ops.size() <= depth) depth_ops.resize(depth + 1);
    // ops of this depth are appended to a private plan and moved afterwards
    Plan sub;
    Planner Q{sub, P.tu, P.W};
    sub.blob.swap(P.pl.blob);
    const uint64_t n = B.n;
    const uint32_t d = floor_log2_64(n) + 1;
    uint32_t rd, ld;
    veb_split(d, rd, ld, classic);
    const uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
    std::vector<VebBatch> kids;
    // a chained gather of whole leaf blocks (perfect trees of depth ld) can
    // carry their own gathers and lp-key layouts (fused sparse slide)
    uint32_t rd2 = 0, ld2 = 0;
    veb_split(ld, rd2, ld2, classic);
    const uint64_t rp = (1ull << rd2) - 1, lp = (1ull << ld2) - 1;
    auto fusable = [&](uint64_t rows) {
        return Q.tu.fuse_slide && ld >= 2 && lp <= FUSED_LP_MAX && lp + 1 <= CHAINED_KEYS &&
               rp <= 65535 && r + rows * rp <= Q.tu.row_stash_max;
    };
    if (n == (1ull << d) - 1) {
        uint64_t items = 0;
        for (auto &x : B.runs) items += x.count;
        const bool streamed = r * Q.W <= 65536 && (items >= 64 || r * (l + 1) <= 131072);
        if (!streamed && r <= Q.tu.row_stash_max && fusable(r)) {
            // the rows' leaf blocks are done; the last leaf block stays
            // where it is and is laid out on its own
            VebBatch last{l, {}};
            for (auto &x : B.runs)
                for (uint64_t i = 0; i < x.count; i++) {
                    Q.sparse_slide(x.start + i * x.stride, r, l, 0, 0, classic ? 2u : 1u, rp, lp);
                    last.runs.push_back({x.start + i * x.stride + r * (l + 1), 0, 1});
                }
            kids.push_back({r, B.runs});
            kids.push_back(last);
        } else {
            if (Q.gather_sparse(B.runs, r, l, classic ? 2u : 1u)) {
                // both kids were laid out inside the gather
                sub.blob.swap(P.pl.blob);
                P.pl.max_stash = std::max(P.pl.max_stash, sub.max_stash);
                P.pl.max_moves = std::max(P.pl.max_moves, sub.max_moves);
                for (auto &op : sub.ops) depth_ops[depth].push_back(op);
                return;
            }
            bool contiguous = (rd == ld);
            for (auto &x : B.runs)
                if (x.count > 1 && x.stride != n) contiguous = false;
            if (contiguous) {
                VebBatch c{l, {}};
                for (auto &x : B.runs) c.runs.push_back({x.start, l, x.count * (n / l)});
                kids.push_back(c);
            } else {
                VebBatch roots{r, B.runs};
                VebBatch lv{l, {}};
                for (auto &x : B.runs)
                    for (uint64_t i = 0; i < x.count; i++)
                        lv.runs.push_back({x.start + i * x.stride + r, l, r + 1});
                kids.push_back(roots);
                kids.push_back(lv);
            }
        }
    } else {
        for (auto &x : B.runs)
            for (uint64_t i = 0; i < x.count; i++) {
                const uint64_t off = x.start + i * x.stride;
                const uint64_t nf = (n - r) / l, inc = n - r - nf * l;
                // the roots past the full blocks (whose bottom trees are the
                // incomplete one and empty ones) follow the incomplete block;
                // when the gather is a chained pass they ride along in it
                const bool chained = nf && !(nf * Q.W <= 65536 && nf * (l + 1) <= 131072);
                // the full leaf blocks' own gathers ride in the slide
                const bool fuse = chained && fusable(nf);
                if (chained && r <= Q.tu.row_stash_max) {
                    if (fuse)
                        Q.sparse_slide(off, nf, l, inc, r - nf, classic ? 2u : 1u, rp, lp);
                    else
                        Q.sparse_slide(off, nf, l, inc, r - nf);
                } else {
                    Q.gather_sparse({{off, 0, 1}}, nf, l);
                    if (nf < r) Q.rotate_right({{off + nf, 0, 1}}, n - nf, r - nf);
                }
                kids.push_back({r, {{off, 0, 1}}});
                if (nf && !(fuse && chained && r <= Q.tu.row_stash_max))
                    kids.push_back({l, {{off + r, l, nf}}});
                if (inc) kids.push_back({inc, {{off + n - inc, 0, 1}}});
            }
    }
    sub.blob.swap(P.pl.blob);
    P.pl.max_stash = std::max(P.pl.max_stash, sub.max_stash);
    P.pl.max_moves = std::max(P.pl.max_moves, sub.max_moves);
    for (auto &op : sub.ops) depth_ops[depth].push_back(op);
    for (auto &k : kids) plan_mveb_rec(P, depth_ops, leaves_lds, k, depth + 1, classic);
}

// classic = true: the CPU vEB layout (cpu/src/veb-cycles.h:34-160), the same
// recursion with the classic split
inline void plan_mveb(Planner &P, uint64_t n, bool classic = false) {
    if (n <= 1) return;
    std::vector<std::vector<Op>> depth_ops;
    std::vector<VebBatch> leaves;
    plan_mveb_rec(P, depth_ops, leaves, VebBatch{n, {{0, 0, 1}}}, 0, classic);
    for (auto &ops : depth_ops)
        for (auto &op : ops) P.pl.ops.push_back(op);
    // group LDS leaves by size
    std::map<uint64_t, std::vector<Run>> by_n;
    for (auto &lb : leaves)
        for (auto &r : lb.runs) by_n[lb.n].push_back(r);
    for (auto &kv : by_n) P.lds(kv.second, kv.first, classic ? LDS_VEB : LDS_MVEB);
}

// ---------------------------------------------------------------------------
// host-side bounds check of a whole plan: every launch must stay inside
// [0, N).  A failure is an internal planning error, reported before any
// kernel runs.
inline uint64_t runs_end(const Plan &pl, const Op &op, uint64_t item_len) {
    uint64_t end = 0;
    for (uint64_t i = 0; i < op.nruns; i++) {
        const uint64_t *r = &pl.blob[op.runs_at + 4 * i];
        if (r[2] == 0) continue;
        end = std::max(end, r[0] + (r[2] - 1) * r[1] + item_len);
    }
    return end;
}

// a fused sparse slide / spread: leaf blocks of rp rows of (lp keys, 1 root)
// and lp keys, lp + 1 a power of 2 dividing the chunk, roots in the stash
inline bool fused_spec_ok(const Plan &pl, const MoveSpec &m) {
    const uint64_t L1 = m.lp + 1;
    return m.fuse <= 2 && m.lp >= 1 && m.lp <= FUSED_LP_MAX && (L1 & m.lp) == 0 &&
           m.CH % L1 == 0 && m.l == m.rp * L1 + m.lp && m.soff + m.r * m.rp <= pl.max_stash;
}

inline bool validate_plan(const Plan &pl, uint64_t N, std::string &why) {
    for (size_t k = 0; k < pl.ops.size(); k++) {
        const Op &op = pl.ops[k];
        uint64_t end = 0;
        bool neg = false;
        switch (op.type) {
        case OP_LDS:
            end = runs_end(pl, op, op.p[0]) + (op.p[6] ? 1 : 0);
            if (op.p[6] && (op.p[1] != LDS_LEVELREV || op.p[6] != 1)) neg = true;
            break;
        case OP_REVERSE: end = runs_end(pl, op, op.p[0]); break;
        case OP_BITREV: {
            const int64_t base = (int64_t)op.p[0];
            const uint32_t D = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            if (base < -1 || D < 2 * K || D > 62) neg = true;
            end = (uint64_t)(base + (int64_t)(1ull << D));
            if (op.p[3] == BITREV_HEAD || op.p[3] == BITREV_ABSORB) {
                if (base != 0 || op.p[4] != 2 * K || rowend_slots(D, K, 2 * K) > pl.max_stash)
                    neg = true;
            } else if (op.p[3] != BITREV_PLAIN) {
                neg = true;
            }
            break;
        }
        case OP_BITREV_LEVELS: {
            const uint32_t t0 = (uint32_t)op.p[0], d = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            if (t0 != 2 * K || d <= t0 || d > 62 || op.p[3] > BITREV_UNFOLD) neg = true;
            end = (1ull << d) - 1;
            if (op.p[3] != BITREV_PLAIN) {   // one place right: up to A[2^d - 1]
                end = 1ull << d;
                if (rowend_slots(d, K, t0) > pl.max_stash) neg = true;
            }
            break;
        }
        case OP_PART_ROOTS:
            end = op.p[2] ? op.p[0] + (op.p[2] - 1) * op.p[1] + op.p[3] : 0;
            if (op.p[3] > op.p[1] || op.p[3] > 65535 || op.p[5] < 1 || op.p[5] > 2 ||
                op.p[4] + op.p[2] * op.p[3] > pl.max_stash)
                neg = true;
            break;
        case OP_HEADS: {
            const uint32_t t0 = (uint32_t)op.p[0], d = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            if (t0 != 2 * K || d <= t0 || d > 62 || rowend_slots(d, K, t0) > pl.max_stash)
                neg = true;
            end = 1ull << d;
            break;
        }
        case OP_ROTATE:
            end = op.p[0] + op.p[1];
            if (op.p[2] == 0 || op.p[2] >= op.p[1] || op.p[3] == 0 || op.p[3] > op.p[1] ||
                op.p[3] > pl.max_stash || op.p[3] < gcd64(op.p[1], op.p[2]))
                neg = true;
            break;
        case OP_UNIT_MERGE: {
            uint64_t mx = 0;
            const uint64_t ncyc = op.p[2];
            const uint64_t units = pl.blob[op.p[1] + ncyc];
            for (uint64_t i = 0; i < units; i++) mx = std::max(mx, pl.blob[op.p[3] + i]);
            end = runs_end(pl, op, (mx + 1) * op.p[0]);
            break;
        }
        case OP_SPARSE_STREAM:   // a fused gather also covers the last leaf block
            end = runs_end(pl, op, op.p[0] * (op.p[1] + 1) + (op.p[3] ? op.p[1] : 0));
            if (op.p[3] > 2 || (op.p[3] && (op.p[1] + 1 > 2048 || op.p[0] * 8 > 65536))) neg = true;
            break;
        case OP_COPY:
            end = op.p[0] + (op.p[2] ? (op.p[2] - 1) * op.p[1] + 1 : 0);
            if (op.p[4] + op.p[2] > pl.max_stash) neg = true;
            break;
        case OP_CHAINED: {
            const MoveSpec &m = op.ms;
            if (m.kind == MV_SPARSE_SLIDE) {
                end = m.P + m.r * (m.l + 1) + m.L + m.k;
                if (m.CH == 0 || m.CH > CHAINED_KEYS || m.CH > m.l ||
                    op.p[0] != (m.L + m.CH - 1) / m.CH + m.r * m.nchunk)
                    neg = true;
                if (m.fuse && !fused_spec_ok(pl, m)) neg = true;
            } else if (m.kind == MV_SPARSE_SPREAD) {
                end = m.P + m.r * (m.l + 1) + m.L + m.k;
                if (m.CH == 0 || m.CH > CHAINED_KEYS || m.CH > m.l || m.nchunk * m.CH < m.l ||
                    op.p[0] != (m.L + m.CH - 1) / m.CH + m.r * m.nchunk)
                    neg = true;
                if (m.fuse && !fused_spec_ok(pl, m)) neg = true;
                break;
            } else if (m.kind == MV_ROW_SLIDE || m.kind == MV_ROW_SPREAD) {
                end = m.P + m.r * (m.l + 1);
                if (m.CH == 0 || m.CH * (m.l + 1) > CHAINED_KEYS || m.nchunk * m.CH < m.r ||
                    op.p[0] != m.nchunk || m.l == 0)
                    neg = true;
                if (m.r > pl.max_stash) neg = true;
                break;
            } else if (m.kind == MV_SHIFT_RIGHT) {
                end = m.P + m.L + m.k;
            } else {
                end = m.P + m.L;
                if (m.k > m.P) neg = true;
            }
            if (m.CH == 0 || m.CH > CHAINED_KEYS) neg = true;
            break;
        }
        }
        if (neg || end > N) {
            why = "internal plan error: op " + std::to_string(k) + " (type " +
                  std::to_string((int)op.type) + ") reaches " + std::to_string(end) +
                  " of " + std::to_string(N);
            return false;
        }
    }
    return true;
}

inline bool width_of(tl_dtype dt, uint32_t &W) {
    switch (dt) {
    case TL_U32:
    case TL_I32: W = 4; return true;
    case TL_U64:
    case TL_I64: W = 8; return true;
    }
    return false;
}

enum Layout { L_BST, L_BTREE, L_MVEB, L_VEB };

// flag in a layout number: the inverse permutation (layout -> sorted order)
constexpr uint32_t PLAN_INVERSE = 0x100;

// The inverse of a forward plan (layout -> sorted order): the launches in
// reverse order, each replaced by its inverse.  Reversals and bit-reversal
// tiles are involutions (for a BST of 2^D keys the shifted level reversal
// becomes the unfold, which puts the row ends in the stash, and phase one
// then absorbs them from there; the heads copy is dropped); LDS layouts, unit
// merges and streamed root gathers run with their inverse flag; copies swap
// direction; a shift becomes the
// opposite shift of the moved region; a one-pass rotation by k becomes the
// rotation by len - k (it fills its own stash); a row slide becomes a row spread (which
// takes the separators back from the stash the inverted copy before it
// fills); a sparse slide becomes a sparse spread (its roots are put back by
// the inverted copies around it).
// Runs, cycle tables, stash and flag sizes are the forward plan's.
inline void invert_plan(const Plan &f, Plan &g) {
    g.ops.clear();
    g.blob = f.blob;
    g.max_stash = f.max_stash;
    g.max_moves = f.max_moves;
    for (size_t i = f.ops.size(); i-- > 0;) {
        Op op = f.ops[i];
        switch (op.type) {
        case OP_LDS: op.p[5] ^= 1; break;
        case OP_UNIT_MERGE: op.p[4] ^= 1; break;
        case OP_SPARSE_STREAM: op.p[2] ^= 1; break;
        case OP_COPY: op.p[3] ^= 1; break;
        case OP_REVERSE: break;
        case OP_BITREV:   // an involution; the kept tree's phase one absorbs row ends
            if (op.p[3] != BITREV_PLAIN) op.p[3] = BITREV_HEAD + BITREV_ABSORB - op.p[3];
            break;
        case OP_BITREV_LEVELS:   // the shift becomes the unfold
            if (op.p[3] != BITREV_PLAIN) op.p[3] = BITREV_SHIFT + BITREV_UNFOLD - op.p[3];
            break;
        case OP_HEADS: continue;   // the unfold keeps row ends instead
        case OP_PART_ROOTS: op.p[6] ^= 1; break;
        case OP_ROTATE: op.p[2] = op.p[1] - op.p[2]; break;
        case OP_CHAINED: {
            MoveSpec &m = op.ms;
            if (m.kind == MV_SHIFT_RIGHT) {
                m.kind = MV_SHIFT_LEFT;
                m.P += m.k;
            } else if (m.kind == MV_SHIFT_LEFT) {
                m.kind = MV_SHIFT_RIGHT;
                m.P -= m.k;
            } else if (m.kind == MV_SPARSE_SLIDE) {
                m.kind = MV_SPARSE_SPREAD;
            } else if (m.kind == MV_ROW_SLIDE) {
                m.kind = MV_ROW_SPREAD;
            }
            break;
        }
        }
        g.ops.push_back(op);
    }
}

// stash_cap (elements, 0 = none): a hard bound on every stash the plan uses,
// for callers whose scratch is fixed in advance
inline void build_plan(Plan &pl, const Tuning &tu0, Layout lay, uint64_t n, uint64_t b, uint32_t W,
                       uint64_t stash_cap = 0) {
    // scratch stays at most n/16 keys (or 2^24 keys for small arrays)
    Tuning tu = tu0;
    tu.row_stash_max = tu.row_stash_env >= 0 ? (uint64_t)tu.row_stash_env
                                              : std::max<uint64_t>(tu.stash_max, n / 16);
    // one-pass rotations use the same budget: more lanes, shorter walks
    if (tu.walk_env >= 0) {
        tu.rotate_walk = tu.rotate_walk && tu.walk_env > 0;
        tu.walk_lanes = (uint64_t)tu.walk_env;
    } else {
        tu.walk_lanes = std::max<uint64_t>(tu.stash_max, n / 16);
    }
    if (stash_cap) {
        tu.stash_max = std::min(tu.stash_max, stash_cap);
        tu.row_stash_max = std::min(tu.row_stash_max, stash_cap);
        tu.walk_lanes = std::min(tu.walk_lanes, stash_cap);
    }
    Planner P{pl, tu, W};
    if (lay == L_BST)
        plan_bst(P, n);
    else if (lay == L_BTREE)
        plan_btree(P, n, b);
    else
        plan_mveb(P, n, lay == L_VEB);
}

}   // namespace tl
