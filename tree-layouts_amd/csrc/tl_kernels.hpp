// tl_kernels.hpp -- gfx950 kernels of the in-place layout permutations.
//
// Every kernel here moves keys of width sizeof(T) (4 or 8 bytes); layouts do
// not depend on key values, so one instantiation per width serves every key
// type.  All global accesses are coalesced runs: a workgroup reads and writes
// contiguous spans of at least 256 elements (the bit-reversal tiles read and
// write rows of 2^K elements).
#pragma once

#include "tl_internal.hpp"

namespace tl {

// ---------------------------------------------------------------------------
// kind of in-LDS layout: per-level bit reversal of a BST's top levels
enum : uint32_t { LDS_LEVELREV = 3 };

__device__ inline uint32_t rev_low_bits32(uint32_t x, uint32_t bits) {
    return bits ? (__brev(x) >> (32 - bits)) : 0u;
}

// ---------------------------------------------------------------------------
// tiled bit-reversal involution: the key at base + y swaps with the key at
//     base + rev_D(y), y in [0, 2^D).  y = (h | m | l) with K high bits, D-2K
//     middle bits and K low bits; rev(y) = (rev l | rev m | rev h), so tile m
//     (2^K rows of 2^K consecutive keys) exchanges with tile rev(m) through
//     LDS, transposed.  A key whose index base + y is negative does not exist
//     (the 1-based BST index 0).  Carter-Gatlin style blocking.
//     Tile numbers are visited so that consecutive workgroups share the pages
//     of both tiles of their pair: with 2^p tiles per 2 MiB page, the tile
//     number's middle 2p - MB bits (which change neither page) vary fastest.
__device__ inline uint64_t bitrev_tile_order(uint64_t t, uint32_t MB, uint32_t p) {
    const uint32_t c = (2 * p > MB) ? min(2 * p - MB, MB) : 0;
    const uint32_t sb = (MB - c) / 2;
    const uint64_t C = t & ((1ull << c) - 1);
    const uint64_t B = (t >> c) & ((1ull << sb) - 1);
    const uint64_t Ahi = t >> (c + sb);
    return (Ahi << (sb + c)) | (C << sb) | B;
}

// stash slot of the row end at slot Y >= 2^t0 (Y % 2^K = 2^K - 1) of a BST
// stored one place right (slot Y at A[Y]), t0 = 2K: level t = floor(log2 Y)
// holds its rows q = (h << (t - 2K)) | m (m its tile number) at
// ((2^t - 2^t0) >> K) + m * 2^K + h, so that each tile's row ends are
// contiguous
template <int K>
__device__ inline uint64_t rowend_slot(uint64_t Y, uint32_t t0) {
    const uint32_t t = 63u - (uint32_t)__clzll(Y);
    const uint32_t MBt = t - 2 * K;
    const uint64_t q = (Y - (1ull << t)) >> K;
    const uint64_t h = q >> MBt, m = q & ((1ull << MBt) - 1);
    return (((1ull << t) - (1ull << t0)) >> K) + ((m << K) | h);
}

// one tile pair of a D-bit reversal at base: tile number t of 2^(D-2K).
// The forms (FOLD), with slot y of the tile and l its place in its row:
//   0: slot y is A[base + y] (slot -1 does not exist) and stays there;
//   2 (unfold; a level of a BST whose final places are one left of where it
//      is kept): read A[base + y], write A[base + 1 + y], or, when
//      l = 2^K - 1, S[m * 2^K + h] (the tile's row ends contiguous);
//   4 (absorb, base 0): as 0, with the row ends at slots >= 2^t0 read from
//      S[rowend_slot(y)], where unfolding the levels left them.
// A tile pair reads all its slots before it writes any, and every key a tile
// writes over is one of its own slots' or a row end put in S, so tiles need
// no order among themselves.
// The LDS of a tile pair: two tiles of 2^K x (2^K + 1) keys of up to 8 bytes
// and two row-end vectors, one buffer per tile size, so a kernel that handles
// both key widths holds it once.
template <int K>
__device__ inline uint64_t *tile_lds() {
    __shared__ uint64_t buf[2 * (1 << K) * ((1 << K) + 1) + 2 * (1 << K)];
    return buf;
}

template <typename T, int K, int FOLD = 0>
__device__ inline void bitrev_tile(T *__restrict__ A, int64_t base, uint32_t D, uint64_t t,
                                   T *__restrict__ S = nullptr, uint32_t t0 = 0) {
    constexpr int R = 1 << K;
    constexpr int E = R * R / 256;   // elements per thread and tile
    T (*sA)[R + 1] = reinterpret_cast<T (*)[R + 1]>(tile_lds<K>());
    T (*sB)[R + 1] = sA + R;
    const uint32_t MB = D - 2 * K;
    constexpr uint32_t row_log2 = (uint32_t)K + (sizeof(T) == 8 ? 3u : 2u);
    const uint64_t m = bitrev_tile_order(t, MB, 21u - row_log2);
    const uint64_t mr = MB ? (__brevll(m) >> (64 - MB)) : 0;
    if (m > mr) return;
    const uint32_t hs = D - K;
    auto load = [&](uint64_t mm, uint32_t h, uint32_t l) -> T {
        const uint64_t y = ((uint64_t)h << hs) | (mm << K) | l;
        if constexpr (FOLD == 4) {
            return (l == R - 1 && (y >> t0) != 0) ? S[rowend_slot<K>(y, t0)] : A[y];
        } else {
            return A[base + (int64_t)y];
        }
    };
    auto store = [&](uint64_t mm, uint32_t h, uint32_t l, T v) {
        const uint64_t y = ((uint64_t)h << hs) | (mm << K) | l;
        if constexpr (FOLD == 2) {
            if (l == R - 1)
                S[(mm << K) | h] = v;
            else
                A[base + 1 + (int64_t)y] = v;
        } else {
            A[base + (int64_t)y] = v;
        }
    };
    // slot 0 of tile 0 does not exist when base = -1
    const bool slot0 = base >= 0;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint32_t e = threadIdx.x + k * 256;
        const uint32_t h = e >> K, l = e & (R - 1);
        if (m != 0 || e != 0 || slot0) sA[h][l] = load(m, h, l);
        if (m != mr) sB[h][l] = load(mr, h, l);
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < R * R; e += blockDim.x) {
        const uint32_t h = e >> K, l = e & (R - 1);
        const uint32_t hr = rev_low_bits32(l, K), lr = rev_low_bits32(h, K);
        if (m != mr) {
            store(m, h, l, sB[hr][lr]);
            store(mr, h, l, sA[hr][lr]);
        } else if (m != 0 || e != 0 || slot0) {
            store(m, h, l, sA[hr][lr]);
        }
    }
}

// one tile pair of the reversal of one level of a BST kept one place right of
// its final places (slot y of the level at A[base + 1 + y], D its depth),
// written in place one slot left: the kept row (h, m) gets the values of its
// slots 1 .. 2^K - 1 and, last, of the first slot of the next row.  That value
// comes from the heads H: the level's first 2^(D-K) slots (row 0 of every
// tile) as they were before the launch, then the next level's first slot.
// All reads and writes are whole aligned rows of the tile pair's own.
// first: tile 0 also writes A[base], the final place of the level's first
// slot; last: the level's last kept slot is not written (the tree's leaf goes
// there).
template <typename T, int K>
__device__ inline void bitrev_tile_shift(T *__restrict__ A, int64_t base, uint32_t D, uint64_t t,
                                         const T *__restrict__ H, bool first, bool last) {
    constexpr int R = 1 << K;
    constexpr int E = R * R / 256;
    T (*sA)[R + 1] = reinterpret_cast<T (*)[R + 1]>(tile_lds<K>());
    T (*sB)[R + 1] = sA + R;
    T *nA = reinterpret_cast<T *>(sB + R);
    T *nB = nA + R;
    const uint32_t MB = D - 2 * K;
    constexpr uint32_t row_log2 = (uint32_t)K + (sizeof(T) == 8 ? 3u : 2u);
    const uint64_t m = bitrev_tile_order(t, MB, 21u - row_log2);
    const uint64_t mr = MB ? (__brevll(m) >> (64 - MB)) : 0;
    if (m > mr) return;
    const uint32_t hs = D - K;
    T *Y = A + base + 1;   // slot y of the level at Y[y]
#pragma unroll
    for (int k = 0; k < E; k++) {
        const uint32_t e = threadIdx.x + k * 256;
        const uint32_t h = e >> K, l = e & (R - 1);
        sA[h][l] = Y[((uint64_t)h << hs) | (m << K) | l];
        if (m != mr) sB[h][l] = Y[((uint64_t)h << hs) | (mr << K) | l];
    }
    // first slot of the row after (h, mm): row (h, mm + 1), or (h + 1, 0), or
    // the next level's first slot; after the reversal slot (h', m', 0) holds
    // what slot (0, rev m', rev h') held, which is H[rev m' * 2^K + rev h']
    if (threadIdx.x < 2 * R) {
        const uint32_t h = threadIdx.x & (R - 1);
        const uint64_t mm = threadIdx.x < R ? m : mr;
        T v{};
        if (mm + 1 < (1ull << MB)) {
            const uint64_t mn = MB ? (__brevll(mm + 1) >> (64 - MB)) : 0;
            v = H[(mn << K) | rev_low_bits32(h, K)];
        } else if (h + 1 < (uint32_t)R) {
            v = H[rev_low_bits32(h + 1, K)];
        } else if (!last) {
            v = H[1ull << (D - K)];
        }
        if (threadIdx.x < R)
            nA[h] = v;
        else
            nB[h] = v;
    }
    __syncthreads();
    const uint64_t lastrow = (1ull << MB) - 1;
    for (uint32_t e = threadIdx.x; e < R * R; e += blockDim.x) {
        const uint32_t h = e >> K, j = e & (R - 1);
        const bool end = j == R - 1;
        const uint32_t hr = end ? 0u : rev_low_bits32(j + 1, K), lr = rev_low_bits32(h, K);
        const bool skip = last && end && h == R - 1;   // the level's last kept slot
        if (m != mr) {
            if (!(skip && m == lastrow))
                Y[((uint64_t)h << hs) | (m << K) | j] = end ? nA[h] : sB[hr][lr];
            if (!(skip && mr == lastrow))
                Y[((uint64_t)h << hs) | (mr << K) | j] = end ? nB[h] : sA[hr][lr];
        } else if (!(skip && m == lastrow)) {
            Y[((uint64_t)h << hs) | (m << K) | j] = end ? nA[h] : sA[hr][lr];
        }
    }
    if (first && m == 0 && threadIdx.x == 0) A[base] = sA[0][0];
}

// tile number of workgroup b when `xcd` is set: workgroups go round robin to
// the 8 XCDs (b % 8), each with its own L2 and address translation cache, so
// XCD x takes the x-th eighth of the tile numbers in order and neighbouring
// tiles (which share pages and the cache lines at their row ends) run on one
// XCD.  The launch makes its workgroup count a multiple of 8.
__device__ inline uint64_t tile_of(uint32_t xcd) {
    const uint64_t b = (uint64_t)blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
    if (!xcd) return b;
    const uint64_t per = ((uint64_t)gridDim.x * gridDim.y) >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename T, int K, int FOLD = 0>
__global__ __launch_bounds__(256) void k_bitrev_tiles(T *__restrict__ A, int64_t base,
                                                      uint32_t D, T *__restrict__ S,
                                                      uint32_t t0, uint32_t xcd) {
    const uint32_t MB = D - 2 * K;
    const uint64_t t = tile_of(xcd);
    if (MB < 64 && (t >> MB) != 0) return;
    bitrev_tile<T, K, FOLD>(A, base, D, t, S, t0);
}

// the same tiles with a grid of fewer workgroups, each taking tile numbers
// t = blockIdx.x, + gridDim.x, ...: the half of the numbers whose pair
// belongs to the partner tile cost a loop trip instead of a workgroup
template <typename T, int K, int FOLD = 0>
__global__ __launch_bounds__(256) void k_bitrev_tiles_loop(T *__restrict__ A, int64_t base,
                                                           uint32_t D, uint64_t tiles,
                                                           T *__restrict__ S, uint32_t t0) {
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        bitrev_tile<T, K, FOLD>(A, base, D, t, S, t0);
        __syncthreads();   // the tile buffers are reused by the next trip
    }
}

// the per-level reversals of levels t0 .. d-1 (t0 = 2K) in one launch: level t
// holds 2^t keys at base 2^t - 1 and 2^(t-2K) tiles; tile number g of the
// launch is tile g - (2^j - 1) of level t0 + j, j = floor(log2(g + 1)).
// FOLD 1 (shift) / 2 (unfold): the tree kept one place right; level t's part
// of S (heads or row ends) starts at (2^t - 2^t0) >> K
template <typename T, int K, int FOLD>
__device__ inline void level_tile(T *__restrict__ A, uint32_t t0, uint32_t d, uint64_t g,
                                  T *__restrict__ S) {
    const uint32_t j = 63u - (uint32_t)__clzll(g + 1);
    const uint32_t t = t0 + j;
    const int64_t base = (int64_t)((1ull << t) - 1);
    const uint64_t tile = g - ((1ull << j) - 1);
    if constexpr (FOLD == 1) {
        bitrev_tile_shift<T, K>(A, base, t, tile, S + (((1ull << t) - (1ull << t0)) >> K),
                                t == t0, t + 1 == d);
    } else if constexpr (FOLD == 2) {
        bitrev_tile<T, K, 2>(A, base, t, tile, S + (((1ull << t) - (1ull << t0)) >> K));
    } else {
        bitrev_tile<T, K>(A, base, t, tile);
    }
}

template <typename T, int K, int FOLD = 0>
__global__ __launch_bounds__(256) void k_bitrev_levels(T *__restrict__ A, uint32_t t0, uint32_t d,
                                                       uint32_t xcd, T *__restrict__ S) {
    const uint64_t g = tile_of(xcd);
    if (g >= (1ull << (d - t0)) - 1) return;
    level_tile<T, K, FOLD>(A, t0, d, g, S);
}

template <typename T, int K, int FOLD = 0>
__global__ __launch_bounds__(256) void k_bitrev_levels_loop(T *__restrict__ A, uint32_t t0,
                                                            uint32_t d, T *__restrict__ S) {
    const uint64_t total = (1ull << (d - t0)) - 1;
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x) {
        level_tile<T, K, FOLD>(A, t0, d, g, S);
        __syncthreads();   // the tile buffers are reused by the next trip
    }
}

// the heads of levels t0 .. d-1 of a BST kept one place right (level t's
// first 2^(t-K) slots, at A[2^t ..]) to S, level after level: what the
// shifted level reversal reads for the first slots of the rows after its own
template <typename T, int K>
__global__ __launch_bounds__(256) void k_level_heads(const T *__restrict__ A, T *__restrict__ S,
                                                     uint32_t t0, uint32_t d) {
    const uint64_t total = ((1ull << d) - (1ull << t0)) >> K;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t = 63u - (uint32_t)__clzll((i << K) + (1ull << t0));
        S[i] = A[(1ull << t) + i - (((1ull << t) - (1ull << t0)) >> K)];
    }
}

// ---------------------------------------------------------------------------
// decoupled chained moves.  A move copies a chunk of at most 256*PER keys
//     from src to dst inside one array.  Moves are numbered so that every
//     chunk whose source overlaps a move's destination has a smaller number.
//     Workgroups take move numbers from an atomic ticket, load the chunk,
//     publish "loaded", wait until the lower-numbered moves whose sources
//     they overwrite have published, then store.  Every awaited move belongs
//     to a workgroup that already started and publishes before it waits, so
//     the scheme cannot deadlock; the wait is still bounded and reports a
//     timeout through err instead of hanging.
//   MV_ROW_SLIDE: r rows of (l keys, 1 separator) at P become
//     [r separators (left in the stash)][r*l keys]; a move carries CH whole
//     rows, writes their keys contiguously and their separators to the stash.
//   MV_SPARSE_SLIDE with fuse = 1 (mvEB) or 2 (classic vEB): every leaf
//     block of l keys is itself a perfect tree of rp rows of (lp keys,
//     1 root) and a last block of lp keys (lp + 1 a power of 2), and its own
//     gather rides along: a leaf-block move carries whole rows (CH a multiple
//     of lp + 1), writes each row's lp keys in their block's vEB order to
//     the block's final place (after its rp roots) and the row's root to
//     stash[soff + i*rp + row]; OP_PART_ROOTS lays the roots out afterwards.
//   MV_ROW_SPREAD, MV_SPARSE_SPREAD: the inverses (layout -> sorted order).
//     The keys move left, back into rows of l keys with one gap after each
//     row; moves are numbered from the left, so every source a move
//     overwrites belongs to a move with a smaller number.  A row spread also
//     fills each gap with the row's separator from the stash (stash[j] for
//     row j); the roots of a sparse spread are put back by strided copies.
enum : uint32_t {
    MV_SHIFT_RIGHT = 0,
    MV_SHIFT_LEFT = 1,
    MV_SPARSE_SLIDE = 2,
    MV_ROW_SLIDE = 3,
    MV_ROW_SPREAD = 4,
    MV_SPARSE_SPREAD = 5
};

// keys per thread of a chained move (256 threads per move)
constexpr int CHAINED_PER = 32;
constexpr uint64_t CHAINED_KEYS = 256 * CHAINED_PER;

struct MoveSpec {
    uint32_t kind;
    uint64_t P;       // region start
    uint64_t L;       // shift: region length
    uint64_t k;       // shift distance
    uint64_t r, l;    // slides: rows and leaf-block length
    uint64_t CH;      // chunk length (rows per move for the row slide)
    uint64_t nchunk;  // chunks per region / per leaf block, or row-slide moves
    uint32_t fuse;    // sparse slide / spread: leaf blocks' own gather (1 mvEB, 2 vEB)
    uint64_t rp, lp;  // ... its rows and leaf-block length
    uint64_t soff;    // ... stash slot of block 0's first row root
};

// a fused leaf-block move of a sparse slide (t counted from 0 in move order)
__device__ inline bool fused_slide_move(const MoveSpec &m, uint64_t t) {
    return m.fuse && t >= (m.L + m.CH - 1) / m.CH;
}

// ... of a sparse spread (leaf blocks first)
__device__ inline bool fused_spread_move(const MoveSpec &m, uint64_t t) {
    return m.fuse && t < m.r * m.nchunk;
}

// leaf keys of a fused move of len source keys: its full rows each hold one
// root, the leaf block's last lp keys (a short last row) none
__device__ inline uint64_t fused_leaves(const MoveSpec &m, uint64_t len) {
    return len - len / (m.lp + 1);
}

__device__ inline void move_at(const MoveSpec &m, uint64_t t, uint64_t &src, uint64_t &dst,
                               uint64_t &len) {
    if (m.kind == MV_SHIFT_RIGHT) {
        const uint64_t c = m.nchunk - 1 - t;
        src = m.P + c * m.CH;
        len = min(m.CH, m.L - c * m.CH);
        dst = src + m.k;
    } else if (m.kind == MV_SHIFT_LEFT) {
        src = m.P + t * m.CH;
        len = min(m.CH, m.L - t * m.CH);
        dst = src - m.k;
    } else if (m.kind == MV_ROW_SLIDE) {
        const uint64_t j0 = (m.nchunk - 1 - t) * m.CH;
        const uint64_t j1 = min(m.r, j0 + m.CH);
        src = m.P + j0 * (m.l + 1);
        len = (j1 - j0) * (m.l + 1);
        dst = m.P + m.r + j0 * m.l;
    } else {
        // tail chunks first (rightmost first), then the leaf blocks
        const uint64_t ntail = (m.L + m.CH - 1) / m.CH;
        if (t < ntail) {
            const uint64_t c = ntail - 1 - t;
            src = m.P + m.r * (m.l + 1) + c * m.CH;
            len = min(m.CH, m.L - c * m.CH);
            dst = src + m.k;
            return;
        }
        t -= ntail;
        const uint64_t i = m.r - 1 - t / m.nchunk;
        const uint64_t c = m.nchunk - 1 - t % m.nchunk;
        src = m.P + i * (m.l + 1) + c * m.CH;
        len = min(m.CH, m.l - c * m.CH);
        dst = m.P + m.r + m.k + i * m.l +
              (m.fuse ? m.rp + c * m.CH / (m.lp + 1) * m.lp : c * m.CH);
    }
}

// moves of the inverse kinds (MV_ROW_SPREAD, MV_SPARSE_SPREAD), numbered
// from the left
__device__ inline void move_at_spread(const MoveSpec &m, uint64_t t, uint64_t &src, uint64_t &dst,
                                      uint64_t &len) {
    if (m.kind == MV_ROW_SPREAD) {
        // rows j0 .. j1-1: keys [P + r + j0*l, +len) -> rows at P + j*(l+1)
        const uint64_t j0 = t * m.CH;
        const uint64_t j1 = min(m.r, j0 + m.CH);
        src = m.P + m.r + j0 * m.l;
        len = (j1 - j0) * m.l;
        dst = m.P + j0 * (m.l + 1);
        return;
    }
    // leaf blocks first (block 0, chunk 0 first), then the tail chunks
    const uint64_t nleaf = m.r * m.nchunk;
    if (t < nleaf) {
        // len counts destination keys (a fused move's roots included)
        const uint64_t i = t / m.nchunk, c = t - i * m.nchunk;
        src = m.P + m.r + m.k + i * m.l +
              (m.fuse ? m.rp + c * m.CH / (m.lp + 1) * m.lp : c * m.CH);
        len = min(m.CH, m.l - c * m.CH);
        dst = m.P + i * (m.l + 1) + c * m.CH;
        return;
    }
    const uint64_t c = t - nleaf;
    dst = m.P + m.r * (m.l + 1) + c * m.CH;
    len = min(m.CH, m.L - c * m.CH);
    src = dst + m.k;
}

// Bounded spin wait.  Polls with relaxed agent-scope loads (no cache
// maintenance per poll), backs off with s_sleep, and issues one agent-scope
// acquire fence once `ready()` holds.  Gives up after `limit` ticks of the
// constant-rate wall clock, or soon after another workgroup gave up
// (err != 0); a give-up ORs `code` into err and is reported as a timeout.
template <bool Acquire = true, typename Ready>
__device__ inline bool bounded_wait(Ready ready, uint32_t *err, uint32_t code, uint64_t limit) {
    if (!ready()) {
        const uint64_t t0 = (uint64_t)wall_clock64();
        for (uint32_t spins = 1;; spins++) {
            if (spins < 16)
                __builtin_amdgcn_s_sleep(1);
            else if (spins < 256)
                __builtin_amdgcn_s_sleep(8);
            else
                __builtin_amdgcn_s_sleep(32);
            if (ready()) break;
            if ((spins & 7u) == 0u) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
                    return false;
                if ((uint64_t)wall_clock64() - t0 > limit) {
                    atomicOr(err, code);
                    return false;
                }
            }
        }
    }
    if (Acquire) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// chained-move flag: "the chunk was read".  The waiter only overwrites that
// chunk and never reads data the flag publishes, so no acquire is needed.
__device__ inline bool wait_flag(uint32_t *f, uint32_t *err, uint64_t limit) {
    return bounded_wait<false>(
        [f]() { return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u; },
        err, 1u, limit);
}

// every load this wave issued has returned its data (vmcnt = 0); cheaper than
// a release fence, which also writes back the L2
__device__ inline void loads_done() { __builtin_amdgcn_s_waitcnt(0); }

// waits for every move t' < t whose source overlaps [dst, dst+len)
__device__ inline bool wait_deps(const MoveSpec &m, uint64_t t, uint64_t dst, uint64_t len,
                                 uint32_t *flags, uint32_t *err, uint64_t limit) {
    if (m.kind == MV_SHIFT_RIGHT || m.kind == MV_SHIFT_LEFT) {
        if (dst + len <= m.P || dst >= m.P + m.L) return true;
        const uint64_t lo = dst > m.P ? (dst - m.P) / m.CH : 0;
        uint64_t hi = (dst + len - 1 - m.P) / m.CH;
        if (hi >= m.nchunk) hi = m.nchunk - 1;
        for (uint64_t c = lo; c <= hi; c++) {
            const uint64_t tt = (m.kind == MV_SHIFT_RIGHT) ? m.nchunk - 1 - c : c;
            if (tt < t && !wait_flag(&flags[tt], err, limit)) return false;
        }
        return true;
    }
    if (m.kind == MV_ROW_SLIDE) {
        // destination keys [dst, dst + rows*l); sources are groups of CH rows
        const uint64_t dlen = len / (m.l + 1) * m.l;
        if (dlen == 0) return true;
        const uint64_t G = m.CH * (m.l + 1);
        const uint64_t glo = (dst - m.P) / G;
        uint64_t ghi = (dst + dlen - 1 - m.P) / G;
        if (ghi >= m.nchunk) ghi = m.nchunk - 1;
        for (uint64_t g = glo; g <= ghi; g++) {
            const uint64_t tt = m.nchunk - 1 - g;
            if (tt < t && !wait_flag(&flags[tt], err, limit)) return false;
        }
        return true;
    }
    // sparse slide: sources are the tail chunks (numbers 0..ntail-1, the
    // rightmost first) and the leaf blocks of rows r-1..0 after them
    const uint64_t ntail = (m.L + m.CH - 1) / m.CH;
    const uint64_t T0 = m.P + m.r * (m.l + 1);
    if (m.L && dst + len > T0 && dst < T0 + m.L) {
        const uint64_t a = (dst > T0 ? dst : T0) - T0;
        const uint64_t z = (dst + len < T0 + m.L ? dst + len : T0 + m.L) - T0;
        for (uint64_t c = a / m.CH; c <= (z - 1) / m.CH; c++) {
            const uint64_t tt = ntail - 1 - c;
            if (tt < t && !wait_flag(&flags[tt], err, limit)) return false;
        }
    }
    if (dst >= T0) return true;
    const uint64_t L1 = m.l + 1;
    const uint64_t rel = dst - m.P;
    uint64_t i_lo = rel / L1, i_hi = (rel + len - 1) / L1;
    if (i_hi >= m.r) i_hi = m.r - 1;
    for (uint64_t i = i_lo; i <= i_hi && i < m.r; i++) {
        const uint64_t bs = i * L1, be = bs + m.l;   // block i source [bs, be)
        const uint64_t ov_lo = rel > bs ? rel - bs : 0;
        const uint64_t ov_end = (rel + len < be ? rel + len : be);
        if (ov_end <= bs + ov_lo) continue;
        const uint64_t c_lo = ov_lo / m.CH, c_hi = (ov_end - bs - 1) / m.CH;
        for (uint64_t c = c_lo; c <= c_hi; c++) {
            const uint64_t tt = ntail + (m.r - 1 - i) * m.nchunk + (m.nchunk - 1 - c);
            if (tt < t && !wait_flag(&flags[tt], err, limit)) return false;
        }
    }
    return true;
}

// waits of the inverse kinds: the sources of all moves tile [S0, S1) left
// to right in move order, so the moves overlapping [dst, dst + span) form one
// range of numbers
__device__ inline bool wait_deps_spread(const MoveSpec &m, uint64_t t, uint64_t dst, uint64_t len,
                                        uint32_t *flags, uint32_t *err, uint64_t limit) {
    const uint64_t S0 = m.P + m.r + (m.kind == MV_SPARSE_SPREAD ? m.k : 0);
    const uint64_t span = m.kind == MV_ROW_SPREAD ? len / m.l * (m.l + 1) : len;
    const uint64_t e = dst + span;   // written keys lie in [dst, e)
    if (e <= S0) return true;
    const uint64_t a = (dst > S0 ? dst : S0) - S0, z = e - 1 - S0;
    uint64_t ulo, uhi;
    if (m.kind == MV_ROW_SPREAD) {
        ulo = a / (m.CH * m.l);
        uhi = z / (m.CH * m.l);
    } else {
        const uint64_t leaf = m.r * m.l;
        // the move whose source holds x; fused: a block's roots (its first rp
        // keys, read by OP_PART_ROOTS before) count as its first move's
        auto num = [&](uint64_t x) {
            if (x < leaf) {
                const uint64_t i = x / m.l, w = x - i * m.l;
                if (!m.fuse) return i * m.nchunk + w / m.CH;
                const uint64_t rows = w < m.rp ? 0 : (w - m.rp) / m.lp;
                return i * m.nchunk + rows / (m.CH / (m.lp + 1));
            }
            return m.r * m.nchunk + (x - leaf) / m.CH;
        };
        ulo = num(a);
        uhi = num(z);
    }
    for (uint64_t u = ulo; u <= uhi && u < t; u++)
        if (!wait_flag(&flags[u], err, limit)) return false;
    return true;
}

// loads of a row-spread move, indexed by the destination, which is rows of
// (l keys, 1 separator): key c of row j comes from src + j*l + c and the
// separator from stash[j0 + j], so that both the loads and the stores of
// every instruction cover contiguous addresses.  The inverse kinds have a
// kernel instantiation each (MODE), which keeps every kernel on one straight
// path and its registers low.
template <typename T, int PER>
__device__ inline void load_row_spread(const T *A, const T *stash, const MoveSpec &ms,
                                       uint64_t src, uint64_t len, T (&v)[PER]) {
    const uint32_t l = (uint32_t)ms.l, L1 = l + 1;
    const uint32_t tot = (uint32_t)(len / l) * L1;
    const T *st = stash + (src - ms.P - ms.r) / l;
    // (row, column) of destination index threadIdx.x + j*256, stepped
    uint32_t row = threadIdx.x / L1, c = threadIdx.x - row * L1;
    const uint32_t rq = 256u / L1, cq = 256u - rq * L1;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t idx = threadIdx.x + (uint32_t)j * 256u;
        if (idx < tot) v[j] = c == l ? st[row] : A[src + (uint64_t)row * l + c];
        row += rq;
        c += cq;
        if (c >= L1) {
            c -= L1;
            row++;
        }
    }
}

// store of one loaded move: contiguous (a row spread's v is already in
// destination order), or (row slide) keys contiguous and every (l+1)-th
// element to the stash
// kernel modes: the forward kinds (shifts, slides), a row spread, a sparse
// spread
enum : int { CH_FORWARD = 0, CH_ROW_SPREAD = 1, CH_SPARSE_SPREAD = 2 };

// loads of a fused sparse-spread move, indexed by the destination (rows of
// lp keys and a root): key c of row a comes from its block's vEB slot
// src + a*lp + slot[c], the root from the stash
template <typename T, int PER>
__device__ inline void load_fused_spread(const T *A, const T *stash, const MoveSpec &ms,
                                         uint64_t src, uint64_t dst, uint64_t len,
                                         const uint32_t *slot, T (&v)[PER]) {
    const uint32_t lg = 63u - (uint32_t)__clzll(ms.lp + 1), lp = (uint32_t)ms.lp;
    const uint64_t rel = dst - ms.P, i = rel / (ms.l + 1);
    const T *st = stash + ms.soff + i * ms.rp + ((rel - i * (ms.l + 1)) >> lg);
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t idx = threadIdx.x + (uint32_t)j * 256u;
        if (idx < len) {
            const uint32_t a = idx >> lg, c = idx & (lp);
            v[j] = c == lp ? st[a] : A[src + (uint64_t)a * lp + slot[c]];
        }
    }
}

template <typename T, int PER, int MODE = CH_FORWARD>
__device__ inline void store_move(T *A, T *stash, const MoveSpec &ms, uint64_t src, uint64_t dst,
                                  uint64_t len, const T (&v)[PER], bool fused = false,
                                  const uint32_t *slot = nullptr) {
    if (MODE == CH_FORWARD && fused) {
        // rows of (lp keys, 1 root): keys to their vEB slots, roots to the stash
        const uint32_t lg = 63u - (uint32_t)__clzll(ms.lp + 1), lp = (uint32_t)ms.lp;
        const uint64_t rel = src - ms.P, i = rel / (ms.l + 1);
        T *st = stash + ms.soff + i * ms.rp + ((rel - i * (ms.l + 1)) >> lg);
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t idx = threadIdx.x + (uint32_t)j * blockDim.x;
            if (idx < len) {
                const uint32_t a = idx >> lg, c = idx & lp;
                if (c == lp)
                    st[a] = v[j];
                else
                    A[dst + (uint64_t)a * lp + slot[c]] = v[j];
            }
        }
        return;
    }
    if constexpr (MODE == CH_ROW_SPREAD) {
        const uint64_t tot = len / ms.l * (ms.l + 1);
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint64_t idx = threadIdx.x + (uint64_t)j * blockDim.x;
            if (idx < tot) A[dst + idx] = v[j];
        }
        return;
    }
    if (ms.kind == MV_ROW_SLIDE) {
        const uint32_t L1 = (uint32_t)(ms.l + 1), l = (uint32_t)ms.l;
        T *st = stash + (src - ms.P) / L1;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t idx = threadIdx.x + (uint32_t)j * blockDim.x;
            if (idx < len) {
                const uint32_t row = idx / L1, c = idx - row * L1;
                if (c == l)
                    st[row] = v[j];
                else
                    A[dst + (uint64_t)row * l + c] = v[j];
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint64_t idx = threadIdx.x + (uint64_t)j * blockDim.x;
        if (idx < len) A[dst + idx] = v[j];
    }
}

// one chained move, t, by the whole workgroup: load, publish, wait, store.
// slot: the vEB slots of a fused move's lp-key blocks (LDS)
template <typename T, int PER, int MODE>
__device__ inline void chained_move(T *__restrict__ A, const MoveSpec &ms, uint64_t t,
                                    uint32_t *flags, uint32_t *err, uint64_t limit, T *stash,
                                    const uint32_t *slot = nullptr) {
    {
        uint64_t src, dst, len;
        if constexpr (MODE != CH_FORWARD)
            move_at_spread(ms, t, src, dst, len);
        else
            move_at(ms, t, src, dst, len);
        const bool fused = MODE == CH_FORWARD     ? fused_slide_move(ms, t)
                           : MODE == CH_SPARSE_SPREAD ? fused_spread_move(ms, t)
                                                      : false;
        T v[PER];
        if constexpr (MODE == CH_ROW_SPREAD) {
            load_row_spread<T, PER>(A, stash, ms, src, len, v);
        } else if (MODE == CH_SPARSE_SPREAD && fused) {
            load_fused_spread<T, PER>(A, stash, ms, src, dst, len, slot, v);
        } else {
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const uint64_t idx = threadIdx.x + (uint64_t)j * blockDim.x;
                if (idx < len) v[j] = A[src + idx];
            }
        }
        // every wave's loads complete before the chunk is published as read
        loads_done();
        __shared__ uint32_t s_ok;
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_store(&flags[t], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bool ok;
            if constexpr (MODE != CH_FORWARD)
                ok = wait_deps_spread(ms, t, dst, len, flags, err, limit);
            else   // a fused move writes its leaf keys only
                ok = wait_deps(ms, t, dst, fused ? fused_leaves(ms, len) : len, flags, err, limit);
            s_ok = ok ? 1u : 0u;
        }
        __syncthreads();
        // a move whose wait gave up (the call then reports TL_ETIMEOUT) keeps
        // its loads to itself: its destination may hold keys not yet read
        if (s_ok) store_move<T, PER, MODE>(A, stash, ms, src, dst, len, v, fused, slot);
    }
}

// vEB slots of an lp-key block in LDS, for the fused sparse moves
__device__ inline void fill_fused_slots(const MoveSpec &ms, uint32_t *slot) {
    if (!ms.fuse) return;
    for (uint32_t x = threadIdx.x; x < (uint32_t)ms.lp; x += blockDim.x)
        slot[x] = mveb_pos32(x, (uint32_t)ms.lp, ms.fuse == 2);
    __syncthreads();
}

// largest lp of a fused sparse move (its slot table in LDS)
constexpr uint64_t FUSED_LP_MAX = 2047;

template <typename T, int PER, int MODE = CH_FORWARD>
__global__ __launch_bounds__(256) void k_chained(T *__restrict__ A, MoveSpec ms, uint64_t nmoves,
                                                 uint32_t *flags, unsigned long long *counter,
                                                 uint32_t *err, uint64_t limit, T *stash) {
    __shared__ uint64_t s_t;
    __shared__ uint32_t s_slot[FUSED_LP_MAX];
    if constexpr (MODE != CH_ROW_SPREAD) fill_fused_slots(ms, s_slot);
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t t = s_t;
        __syncthreads();
        if (t >= nmoves) return;
        chained_move<T, PER, MODE>(A, ms, t, flags, err, limit, stash, s_slot);
    }
}

// ---------------------------------------------------------------------------
// one-pass rotation of A[0, L) right by c (key at p goes to p + c mod L) by
// parallel cycle leaders.  The first w keys are stashed beforehand; lane j
// (one thread) writes j, j - c, j - 2c, ... (mod L), each from the position c
// before it, until that position is a lane start (< w), whose key comes from
// the stash.  Every position is written by exactly one lane (the first start
// reached walking forward from it, which exists when w >= gcd(L, c)) after that
// lane read it, so lanes never wait for each other.  Neighbouring lanes touch
// neighbouring keys, so a wave's accesses are contiguous; each lane keeps K
// loads in flight.
template <typename T, int K>
__device__ inline void walk_lane(T *__restrict__ A, uint64_t L, uint64_t c, uint64_t w,
                                 const T *__restrict__ stash, uint64_t j) {
    const uint64_t back = L - c;   // p - c (mod L) = p + back when p < c
    uint64_t dst = j;
    for (;;) {
        uint64_t s[K];
        T v[K];
        int m = 0;
        bool stop = false;
        uint64_t d = dst;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!stop) {
                const uint64_t src = d >= c ? d - c : d + back;
                s[k] = src;
                m = k + 1;
                if (src < w)
                    stop = true;
                else
                    d = src;
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (k < m) v[k] = (stop && k == m - 1) ? stash[s[k]] : A[s[k]];
#pragma unroll
        for (int k = 0; k < K; k++)
            if (k < m) A[k == 0 ? dst : s[k - 1]] = v[k];
        if (stop) return;
        dst = s[K - 1];
    }
}

template <typename T, int K, int BS = 256>
__global__ __launch_bounds__(BS) void k_rotate_walk(T *__restrict__ A, uint64_t L, uint64_t c,
                                                    uint64_t w, const T *__restrict__ stash) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < w) walk_lane<T, K>(A, L, c, w, stash, j);
}

// ---------------------------------------------------------------------------
// the rp roots of `count` blocks (block i at base + i*stride) from the stash
// (stash[soff + i*rp + a], put there by a fused sparse slide) to their vEB
// slots at the block's front; inv: from there back to the stash
template <typename T>
__global__ __launch_bounds__(256) void k_part_roots(T *__restrict__ A, T *__restrict__ stash,
                                                    uint64_t base, uint64_t stride, uint64_t count,
                                                    uint32_t rp, uint64_t soff, uint32_t classic,
                                                    uint32_t inv) {
    const uint64_t total = count * rp;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = g / rp;
        const uint32_t a = (uint32_t)(g - i * rp);
        T *slot = A + base + i * stride + mveb_pos32(a, rp, classic != 0);
        if (inv)
            stash[soff + g] = *slot;
        else
            *slot = stash[soff + g];
    }
}

// ---------------------------------------------------------------------------
// strided copy (stash / unstash of displaced keys)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_strided(T *__restrict__ dst, uint64_t dstride,
                                                      const T *__restrict__ src, uint64_t sstride,
                                                      uint64_t count) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i * dstride] = src[i * sstride];
}

}   // namespace tl
