// tl_permute.hip -- host planner + executor of the in-place layout
// permutations (BST, B-tree, modified vEB) on gfx950.
//
// A call builds a Plan: a sequence of kernel launches whose per-launch item
// lists ("runs" of equally sized, equally spaced sub-arrays) are packed into
// one host blob, uploaded once, and then launched back to back on the stream.
//
//  * sub-problems that fit a workgroup's LDS are finished in one read + one
//    write by k_lds_layout with the rank->slot closed forms;
//  * the dense gathers of the B-tree / BST ("every (b+1)-th key to the
//    front", extended_equidistant_gather(2) in gpu/src/cycles.cuh) are split
//    into LDS tiles plus one or two unit-level cycle merges (k_unit_merge);
//  * the perfect BST is two tiled bit-reversal involution passes
//    (gpu/src/bst-hybrid.cuh:31-69);
//  * the sparse gathers of the mvEB (r roots among leaf blocks of l >= r)
//    stream each partition through one workgroup, or, for a few huge
//    partitions, stash the r roots and slide the leaf blocks with decoupled
//    chained moves (gpu/src/mveb-cycles.cuh:230-414);
//  * rotations (shift_right, gpu/src/common.cuh:62-110) use chained moves
//    when one side is small, three reversals otherwise.
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "tl_internal.hpp"
#include "tl_kernels.hpp"
#include "tl_plan.hpp"

namespace tl {

// graph-safe clear (no memset nodes in captured work)
__global__ void k_fill_u32(uint32_t *p, uint64_t count, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// in-LDS layout over runs of items, `per` items per workgroup.  With several
// items per workgroup the slot of every in-item index is computed once into
// an LDS table behind the keys; i / n uses a multiply by ceil(2^32 / n),
// exact because i * n < 2^32.
__device__ inline uint32_t lds_slot(uint32_t x, uint32_t n, uint32_t kind, uint32_t b,
                                    uint32_t rows, const BtreeShape &sh) {
    if (kind == LDS_BTREE) return btree_pos32(x, sh);
    if (kind == LDS_MVEB) return mveb_pos32(x, n);
    if (kind == LDS_VEB) return mveb_pos32(x, n, true);
    if (kind == LDS_GATHER) return gather_pos32(x, rows, b);
    const uint32_t y = x + 1, t = 31u - __clz(y);
    return ((1u << t) | rev_low_bits32(y & ((1u << t) - 1), t)) - 1;
}

// inv = 1 applies the inverse permutation (slot -> rank instead of rank ->
// slot): the keys go into LDS in place and come out gathered through the slots.
// shift = 1: the item's keys are read from one place right (inverse: written
// one place right).
// one workgroup's share: items [group*per, group*per + per) of the runs
template <typename T>
__device__ inline void lds_body(T *__restrict__ A, const uint64_t *__restrict__ runs,
                                uint32_t nruns, uint64_t nitems, uint32_t per, uint32_t n,
                                uint32_t kind, uint32_t b, uint32_t rows, uint32_t magic,
                                uint32_t inv, uint64_t group, uint32_t shift = 0) {
    const uint32_t rshift = inv ? 0u : shift, wshift = inv ? shift : 0u;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *s = reinterpret_cast<T *>(smem_raw);
    uint32_t *tbl = reinterpret_cast<uint32_t *>(smem_raw + (size_t)per * n * sizeof(T));
    __shared__ uint64_t s_off[64];
    BtreeShape sh;
    if (kind == LDS_BTREE) sh = btree_shape(n, b);
    const uint64_t first = group * per;
    const uint32_t cnt = (uint32_t)min((uint64_t)per, nitems - first);
    const bool table = per > 1;
    for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) s_off[t] = run_offset(runs, nruns, first + t);
    if (table)
        for (uint32_t x = threadIdx.x; x < n; x += blockDim.x) tbl[x] = lds_slot(x, n, kind, b, rows, sh);
    __syncthreads();
    const uint32_t tot = cnt * n;
    constexpr int U = 8;
    for (uint32_t i0 = threadIdx.x; i0 < tot; i0 += blockDim.x * U) {
        T v[U];
        uint32_t it[U], ii[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * blockDim.x;
            if (i < tot) {
                it[u] = table ? __umulhi(i, magic) : 0;
                ii[u] = i - it[u] * n;
                v[u] = A[s_off[it[u]] + ii[u] + rshift];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * blockDim.x;
            if (i < tot) {
                const uint32_t p =
                    inv ? ii[u] : (table ? tbl[ii[u]] : lds_slot(ii[u], n, kind, b, rows, sh));
                s[it[u] * n + p] = v[u];
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tot; i += blockDim.x) {
        const uint32_t t = table ? __umulhi(i, magic) : 0, x = i - t * n;
        const uint32_t p = inv ? (table ? tbl[x] : lds_slot(x, n, kind, b, rows, sh)) : x;
        A[s_off[t] + x + wshift] = s[t * n + p];
    }
}

template <typename T>
__global__ __launch_bounds__(512) void k_lds_runs(T *__restrict__ A, const uint64_t *__restrict__ runs,
                                                  uint32_t nruns, uint64_t nitems, uint32_t per,
                                                  uint32_t n, uint32_t kind, uint32_t b,
                                                  uint32_t rows, uint32_t magic, uint32_t inv,
                                                  uint32_t shift) {
    lds_body<T>(A, runs, nruns, nitems, per, n, kind, b, rows, magic, inv, blockIdx.x, shift);
}

// reverse every segment [off, off+len): REV_U mirrored pairs per thread per
// step, all loads issued before any store (the pairs are disjoint)
constexpr int REV_U = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_reverse_runs(T *__restrict__ A, const uint64_t *__restrict__ runs,
                                                      uint32_t nruns, uint64_t nitems, uint64_t len) {
    const uint64_t half = len / 2;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * REV_U;
    for (uint64_t seg = blockIdx.y; seg < nitems; seg += gridDim.y) {
        T *a = A + run_offset(runs, nruns, seg);
        T *z = a + len - 1;
        for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * REV_U + threadIdx.x; base < half;
             base += step) {
            T x[REV_U], y[REV_U];
#pragma unroll
            for (int j = 0; j < REV_U; j++) {
                const uint64_t i = base + (uint64_t)j * blockDim.x;
                if (i < half) {
                    x[j] = a[i];
                    y[j] = *(z - i);
                }
            }
#pragma unroll
            for (int j = 0; j < REV_U; j++) {
                const uint64_t i = base + (uint64_t)j * blockDim.x;
                if (i < half) {
                    a[i] = y[j];
                    *(z - i) = x[j];
                }
            }
        }
    }
}

// unit u[k] of a cycle moves to u[k+1] (the last to u[0]); inv = 1 moves
// them the other way round
template <typename T>
__global__ __launch_bounds__(256) void k_unit_merge_runs(T *__restrict__ A, const uint64_t *__restrict__ runs,
                                                         uint32_t nruns, uint64_t nparents, uint64_t unit,
                                                         const uint64_t *__restrict__ cstart,
                                                         const uint64_t *__restrict__ cunits,
                                                         uint32_t nchunks, uint32_t inv) {
    const uint32_t chunk = blockIdx.x % nchunks;
    const uint32_t cyc = blockIdx.x / nchunks;
    const uint64_t c = (uint64_t)chunk * blockDim.x + threadIdx.x;
    if (c >= unit) return;
    const uint64_t s0 = cstart[cyc];
    const int64_t L = (int64_t)(cstart[cyc + 1] - s0);
    const uint64_t *u = cunits + s0;
    for (uint64_t p = blockIdx.y; p < nparents; p += gridDim.y) {
        T *a = A + run_offset(runs, nruns, p) + c;
        if (inv) {
            T hold = a[u[0] * unit];
            int64_t k = 0;
            while (k < L - 1) {
                T v[8];
                const int nb = L - 1 - k < 8 ? (int)(L - 1 - k) : 8;
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (j < nb) v[j] = a[u[k + j + 1] * unit];
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (j < nb) a[u[k + j] * unit] = v[j];
                k += nb;
            }
            a[u[L - 1] * unit] = hold;
            continue;
        }
        T hold = a[u[L - 1] * unit];
        int64_t k = L - 2;
        while (k >= 0) {
            T v[8];
            const int nb = k + 1 < 8 ? (int)(k + 1) : 8;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (j < nb) v[j] = a[u[k - j] * unit];
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (j < nb) a[u[k - j + 1] * unit] = v[j];
            k -= nb;
        }
        a[u[0] * unit] = hold;
    }
}

// one workgroup per partition of r rows of (l leaf keys, 1 root): the roots
// collect in LDS while the leaf keys stream right from the end, then the
// roots are written to the front.  inv = 1 undoes it: the front r roots go to
// LDS, then the leaf keys stream left from the front of the leaf area, each
// batch read before any of it is written, and each root lands after its
// row's last key.
template <typename T>
__device__ inline void stream_body(T *__restrict__ a, uint64_t r, uint64_t l, uint32_t inv) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *roots = reinterpret_cast<T *>(smem_raw);
    const uint64_t L1 = l + 1;
    const uint64_t total = r * L1;
    constexpr int U = 8;
    const uint64_t batch = (uint64_t)blockDim.x * U;
    if (inv) {
        // leaf keys stream left in source order: key y of the leaf area
        // (row y / l, column y % l) lands at row*(l+1) + column, never to the
        // right of where it was read; a row's root follows its last key
        for (uint64_t i = threadIdx.x; i < r; i += blockDim.x) roots[i] = a[i];
        __syncthreads();
        const uint64_t nleaf = r * l;
        for (uint64_t lo = 0; lo < nleaf; lo += batch) {
            T v[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint64_t y = lo + threadIdx.x + (uint64_t)k * blockDim.x;
                if (y < nleaf) v[k] = a[r + y];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint64_t y = lo + threadIdx.x + (uint64_t)k * blockDim.x;
                if (y < nleaf) {
                    const uint64_t row =
                        nleaf < (1ull << 32) ? (uint64_t)((uint32_t)y / (uint32_t)l) : y / l;
                    const uint64_t cc = y - row * l, d = row * L1 + cc;
                    a[d] = v[k];
                    if (cc == l - 1) a[d + 1] = roots[row];
                }
            }
        }
        return;
    }
    uint64_t hi = total;
    while (hi > 0) {
        const uint64_t lo = hi > batch ? hi - batch : 0;
        T v[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint64_t x = lo + threadIdx.x + (uint64_t)k * blockDim.x;
            if (x < hi) v[k] = a[x];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint64_t x = lo + threadIdx.x + (uint64_t)k * blockDim.x;
            if (x < hi) {
                uint64_t row, cc;
                if (total < (1ull << 32)) {
                    row = (uint32_t)x / (uint32_t)L1;
                    cc = x - row * L1;
                } else {
                    row = x / L1;
                    cc = x - row * L1;
                }
                if (cc == l)
                    roots[row] = v[k];
                else
                    a[r + row * l + cc] = v[k];
            }
        }
        hi = lo;
    }
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < r; i += blockDim.x) a[i] = roots[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_sparse_stream_runs(T *__restrict__ A, const uint64_t *__restrict__ runs,
                                                            uint32_t nruns, uint64_t r, uint64_t l,
                                                            uint32_t inv) {
    stream_body<T>(A + run_offset(runs, nruns, blockIdx.x), r, l, inv);
}

// The same gather with the layouts of its kids fused in (fuse = 1 mvEB,
// 2 classic vEB) over a whole perfect subtree, r rows of (l, 1) followed by
// the last leaf block: the r roots become a vEB of r keys at the front and
// every leaf block (the last one in place) a vEB of l keys, so the in-LDS
// pass over the kids disappears.
// Batches hold whole rows, so every row's destination block (inside which
// its keys are scattered) lies right of the batch start; the inverse walks
// the same way from the left with the slot map inverted in LDS.  Needs
// l + 1 <= 256 * U.
template <typename T, uint32_t NT = 256>
__device__ inline void stream_fused_body(T *__restrict__ a, uint32_t r, uint32_t l, uint32_t inv,
                                         uint32_t fuse) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *roots = reinterpret_cast<T *>(smem_raw);
    uint32_t *slot = reinterpret_cast<uint32_t *>(smem_raw + (((size_t)r * sizeof(T) + 15) & ~15ull));
    const bool classic = fuse == 2;
    const uint32_t L1 = l + 1;
    constexpr int U = 8;
    const uint32_t rows_per = (NT * U) / L1;
    // slot[x]: where rank x of a leaf block goes (its inverse when inv)
    for (uint32_t x = threadIdx.x; x < l; x += blockDim.x) {
        const uint32_t q = mveb_pos32(x, l, classic);
        if (inv)
            slot[q] = x;
        else
            slot[x] = q;
    }
    if (inv)
        for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) roots[i] = a[mveb_pos32(i, r, classic)];
    __syncthreads();
    if (inv) {
        {
            T v[U];
            const uint64_t b = (uint64_t)r * L1;
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint32_t y = threadIdx.x + (uint32_t)k * NT;
                if (y < l) v[k] = a[b + y];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint32_t y = threadIdx.x + (uint32_t)k * NT;
                if (y < l) a[b + slot[y]] = v[k];
            }
        }
        for (uint32_t lo = 0; lo < r; lo += rows_per) {
            const uint32_t hi = min(r, lo + rows_per);
            const uint32_t n_el = (hi - lo) * l;
            T v[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint32_t y = threadIdx.x + (uint32_t)k * NT;
                if (y < n_el) v[k] = a[r + (uint64_t)lo * l + y];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < U; k++) {
                const uint32_t y = threadIdx.x + (uint32_t)k * NT;
                if (y < n_el) {
                    const uint32_t rr = y / l, q = y - rr * l, row = lo + rr;
                    a[(uint64_t)row * L1 + slot[q]] = v[k];
                    if (q == 0) a[(uint64_t)row * L1 + l] = roots[row];
                }
            }
        }
        return;
    }
    {
        // the last leaf block, which the gather leaves in place: a[b + slot[y]]
        // takes the key at b + y (forward: rank y to its slot; inverse: the
        // key at slot y back to its rank)
        T v[U];
        const uint64_t b = (uint64_t)r * L1;
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint32_t y = threadIdx.x + (uint32_t)k * NT;
            if (y < l) v[k] = a[b + y];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint32_t y = threadIdx.x + (uint32_t)k * NT;
            if (y < l) a[b + slot[y]] = v[k];
        }
    }
    uint32_t hi = r;
    while (hi > 0) {
        const uint32_t lo = hi > rows_per ? hi - rows_per : 0;
        const uint32_t n_el = (hi - lo) * L1;
        T v[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint32_t x = threadIdx.x + (uint32_t)k * NT;
            if (x < n_el) v[k] = a[(uint64_t)lo * L1 + x];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < U; k++) {
            const uint32_t x = threadIdx.x + (uint32_t)k * NT;
            if (x < n_el) {
                const uint32_t rr = x / L1, cc = x - rr * L1, row = lo + rr;
                if (cc == l)
                    roots[row] = v[k];
                else
                    a[r + (uint64_t)row * l + slot[cc]] = v[k];
            }
        }
        hi = lo;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) a[mveb_pos32(i, r, classic)] = roots[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_sparse_stream_fused(T *__restrict__ A,
                                                             const uint64_t *__restrict__ runs,
                                                             uint32_t nruns, uint32_t r, uint32_t l,
                                                             uint32_t inv, uint32_t fuse) {
    stream_fused_body<T>(A + run_offset(runs, nruns, blockIdx.x), r, l, inv, fuse);
}

// ---------------------------------------------------------------------------
// executor

// steps of a rotation lane walk with their loads in flight together:
// $TL_WALK_DEPTH (4 or 8), else 4
static int walk_depth() {
    const char *e = getenv("TL_WALK_DEPTH");
    return (e && atoi(e) == 8) ? 8 : 4;
}

// lanes per workgroup of a rotation lane walk: $TL_WALK_BLOCK (256 or 1024),
// else 1024 (a workgroup's lanes cover 8 KiB of neighbouring keys, so fewer
// cache lines at workgroup edges are written from two XCDs: 2% faster at
// 2^30 than 256 lanes)
static int walk_block() {
    const char *e = getenv("TL_WALK_BLOCK");
    return (e && atoi(e) == 256) ? 256 : 1024;
}

// workgroups per compute unit of a chained-move launch (moves are taken
// from a ticket, so more workgroups only add moves in flight):
// $TL_CHAIN_WG_PER_CU (1..64), else 8
static uint64_t chained_wg_per_cu() {
    const char *e = getenv("TL_CHAIN_WG_PER_CU");
    const int v = e ? atoi(e) : 8;
    return (v >= 1 && v <= 64) ? (uint64_t)v : 8u;
}

// bit-reversal tiles: one workgroup per tile number (0, the default), or
// $TL_BITREV_GRID workgroups per compute unit (1..64) looping over them
static uint64_t bitrev_grid_per_cu() {
    const char *e = getenv("TL_BITREV_GRID");
    const int v = e ? atoi(e) : 0;
    return (v >= 1 && v <= 64) ? (uint64_t)v : 0u;
}

// bit-reversal tiles numbered XCD by XCD ($TL_BITREV_XCD=1), else in
// workgroup order
static uint32_t bitrev_xcd() {
    const char *e = getenv("TL_BITREV_XCD");
    return (e && atoi(e) == 1) ? 1u : 0u;
}

// a D-bit reversal at base (tile numbers < tiles): one workgroup per tile
// number, or `per` workgroups per compute unit looping over them
template <typename T, int FOLD>
static tl_status bitrev_launch(Ctx *c, T *A, int64_t base, uint32_t D, uint32_t K, uint64_t tiles,
                               uint32_t per, T *S, uint32_t t0, hipStream_t s) {
    if (per) {
        const uint32_t g = (uint32_t)std::min<uint64_t>(tiles, (uint64_t)c->num_cus * per);
        if (K == 6)
            hipLaunchKernelGGL((k_bitrev_tiles_loop<T, 6, FOLD>), dim3(g), dim3(256), 0, s, A,
                               base, D, tiles, S, t0);
        else
            hipLaunchKernelGGL((k_bitrev_tiles_loop<T, 5, FOLD>), dim3(g), dim3(256), 0, s, A,
                               base, D, tiles, S, t0);
    } else {
        const uint32_t xcd = bitrev_xcd();
        const uint64_t padded = xcd ? (tiles + 7) & ~7ull : tiles;   // tiles: a power of 2
        const uint32_t gx = (uint32_t)std::min<uint64_t>(padded, 1u << 20);
        const uint32_t gy = (uint32_t)((padded + gx - 1) / gx);
        if (K == 6)
            hipLaunchKernelGGL((k_bitrev_tiles<T, 6, FOLD>), dim3(gx, gy), dim3(256), 0, s, A,
                               base, D, S, t0, xcd);
        else
            hipLaunchKernelGGL((k_bitrev_tiles<T, 5, FOLD>), dim3(gx, gy), dim3(256), 0, s, A,
                               base, D, S, t0, xcd);
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// the per-level reversals of levels t0 .. d-1 in one launch
template <typename T, int FOLD>
static tl_status levels_launch(Ctx *c, T *A, uint32_t t0, uint32_t d, uint32_t K, uint32_t per,
                               T *S, hipStream_t s) {
    const uint64_t tiles = (1ull << (d - t0)) - 1;
    if (per) {
        const uint32_t g = (uint32_t)std::min<uint64_t>(tiles, (uint64_t)c->num_cus * per);
        if (K == 6)
            hipLaunchKernelGGL((k_bitrev_levels_loop<T, 6, FOLD>), dim3(g), dim3(256), 0, s, A,
                               t0, d, S);
        else
            hipLaunchKernelGGL((k_bitrev_levels_loop<T, 5, FOLD>), dim3(g), dim3(256), 0, s, A,
                               t0, d, S);
    } else {
        // tiles + 1 is a power of 2: padded, the grid is gx * gy exactly
        const uint32_t xcd = bitrev_xcd();
        const uint64_t padded = xcd ? (tiles + 8) & ~7ull : tiles;
        const uint32_t gx = (uint32_t)std::min<uint64_t>(padded, 1u << 20);
        const uint32_t gy = (uint32_t)((padded + gx - 1) / gx);
        if (K == 6)
            hipLaunchKernelGGL((k_bitrev_levels<T, 6, FOLD>), dim3(gx, gy), dim3(256), 0, s, A,
                               t0, d, xcd, S);
        else
            hipLaunchKernelGGL((k_bitrev_levels<T, 5, FOLD>), dim3(gx, gy), dim3(256), 0, s, A,
                               t0, d, xcd, S);
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

template <typename T>
static tl_status launch_plan(Ctx *c, const Plan &pl, T *A, hipStream_t s, const uint64_t *blob,
                             T *stash, uint32_t *flags, bool preclear) {
    // preclear: every chained op gets its own flag region (plan_flag_bytes_all),
    // zeroed by the caller beforehand; otherwise one region, cleared per op
    uint32_t *op_flags = flags;
    for (const Op &op : pl.ops) {
        const uint64_t *runs = blob ? blob + op.runs_at : nullptr;
        switch (op.type) {
        case OP_LDS: {
            const uint64_t n = op.p[0], per = op.p[4];
            const uint64_t groups = (op.nitems + per - 1) / per;
            const size_t bytes = per * n * sizeof(T) + (per > 1 ? n * 4 : 0);
            const int threads = (per * n >= 8192) ? 512 : 256;
            // ceil(2^32 / n); per * n < 2^16 whenever several items share a workgroup
            const uint32_t magic = (uint32_t)((0xFFFFFFFFull + n) / n);
            hipLaunchKernelGGL(k_lds_runs<T>, dim3((uint32_t)groups), dim3(threads), bytes, s, A,
                               runs, (uint32_t)op.nruns, op.nitems, (uint32_t)per, (uint32_t)n,
                               (uint32_t)op.p[1], (uint32_t)op.p[2], (uint32_t)op.p[3], magic,
                               (uint32_t)op.p[5], (uint32_t)op.p[6]);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_REVERSE: {
            const uint64_t len = op.p[0];
            if (len < 2) break;
            const uint64_t half = len / 2;
            uint64_t bx = (half + 256 * REV_U - 1) / (256 * REV_U);
            bx = std::min<uint64_t>(bx, (uint64_t)c->num_cus * 16);
            const uint64_t by = std::min<uint64_t>(op.nitems, 65535);
            hipLaunchKernelGGL(k_reverse_runs<T>, dim3((uint32_t)bx, (uint32_t)by), dim3(256), 0,
                               s, A, runs, (uint32_t)op.nruns, op.nitems, len);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_BITREV: {
            const int64_t base = (int64_t)op.p[0];
            const uint32_t D = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            const uint64_t tiles = 1ull << (D - 2 * K);
            const uint32_t per = (uint32_t)bitrev_grid_per_cu(), t0 = (uint32_t)op.p[4];
            if (op.p[3] == BITREV_ABSORB)
                TL_TRY((bitrev_launch<T, 4>(c, A, base, D, K, tiles, per, stash, t0, s)));
            else
                TL_TRY((bitrev_launch<T, 0>(c, A, base, D, K, tiles, per, nullptr, 0, s)));
            break;
        }
        case OP_BITREV_LEVELS: {
            const uint32_t t0 = (uint32_t)op.p[0], d = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            const uint32_t per = (uint32_t)bitrev_grid_per_cu();
            if (op.p[3] == BITREV_SHIFT)
                TL_TRY((levels_launch<T, 1>(c, A, t0, d, K, per, stash, s)));
            else if (op.p[3] == BITREV_UNFOLD)
                TL_TRY((levels_launch<T, 2>(c, A, t0, d, K, per, stash, s)));
            else
                TL_TRY((levels_launch<T, 0>(c, A, t0, d, K, per, nullptr, s)));
            break;
        }
        case OP_PART_ROOTS: {
            const uint64_t total = op.p[2] * op.p[3];
            const uint32_t g = (uint32_t)std::min<uint64_t>((total + 255) / 256, c->num_cus * 64ull);
            if (total)
                hipLaunchKernelGGL(k_part_roots<T>, dim3(g), dim3(256), 0, s, A, stash, op.p[0],
                                   op.p[1], op.p[2], (uint32_t)op.p[3], op.p[4],
                                   (uint32_t)(op.p[5] == 2), (uint32_t)op.p[6]);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_HEADS: {
            const uint32_t t0 = (uint32_t)op.p[0], d = (uint32_t)op.p[1], K = (uint32_t)op.p[2];
            const uint64_t total = rowend_slots(d, K, t0);
            const uint32_t g = (uint32_t)std::min<uint64_t>((total + 255) / 256, c->num_cus * 64ull);
            if (K == 6)
                hipLaunchKernelGGL((k_level_heads<T, 6>), dim3(g), dim3(256), 0, s, (const T *)A,
                                   stash, t0, d);
            else
                hipLaunchKernelGGL((k_level_heads<T, 5>), dim3(g), dim3(256), 0, s, (const T *)A,
                                   stash, t0, d);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_ROTATE: {
            const uint64_t P = op.p[0], L = op.p[1], k = op.p[2], w = op.p[3];
            const uint32_t g = (uint32_t)std::min<uint64_t>((w + 255) / 256, 65535);
            hipLaunchKernelGGL(k_copy_strided<T>, dim3(g), dim3(256), 0, s, stash, 1,
                               (const T *)(A + P), 1, w);
            TL_LAUNCH_CHECK();
            if (walk_depth() == 8)
                hipLaunchKernelGGL((k_rotate_walk<T, 8>), dim3((uint32_t)((w + 255) / 256)),
                                   dim3(256), 0, s, A + P, L, k, w, (const T *)stash);
            else if (walk_block() == 1024)
                hipLaunchKernelGGL((k_rotate_walk<T, 4, 1024>), dim3((uint32_t)((w + 1023) / 1024)),
                                   dim3(1024), 0, s, A + P, L, k, w, (const T *)stash);
            else
                hipLaunchKernelGGL((k_rotate_walk<T, 4>), dim3((uint32_t)((w + 255) / 256)),
                                   dim3(256), 0, s, A + P, L, k, w, (const T *)stash);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_UNIT_MERGE: {
            const uint64_t unit = op.p[0];
            const uint32_t nchunks = (uint32_t)((unit + 255) / 256);
            const uint64_t ncyc = op.p[2];
            const uint32_t gy = (uint32_t)std::min<uint64_t>(op.nitems, 65535);
            hipLaunchKernelGGL(k_unit_merge_runs<T>, dim3((uint32_t)(nchunks * ncyc), gy), dim3(256),
                               0, s, A, runs, (uint32_t)op.nruns, op.nitems, unit, blob + op.p[1],
                               blob + op.p[3], nchunks, (uint32_t)op.p[4]);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_SPARSE_STREAM: {
            const uint64_t r = op.p[0], l = op.p[1];
            if (op.p[3])
                hipLaunchKernelGGL(k_sparse_stream_fused<T>, dim3((uint32_t)op.nitems), dim3(256),
                                   ((r * sizeof(T) + 15) & ~15ull) + l * 4, s, A, runs,
                                   (uint32_t)op.nruns, (uint32_t)r, (uint32_t)l,
                                   (uint32_t)op.p[2], (uint32_t)op.p[3]);
            else
                hipLaunchKernelGGL(k_sparse_stream_runs<T>, dim3((uint32_t)op.nitems), dim3(256),
                                   r * sizeof(T), s, A, runs, (uint32_t)op.nruns, r, l,
                                   (uint32_t)op.p[2]);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_COPY: {
            const uint64_t P = op.p[0], stride = op.p[1], count = op.p[2];
            const uint32_t g = (uint32_t)std::min<uint64_t>((count + 255) / 256, 65535);
            if (op.p[3])
                hipLaunchKernelGGL(k_copy_strided<T>, dim3(g), dim3(256), 0, s, stash + op.p[4], 1,
                                   (const T *)(A + P), stride, count);
            else
                hipLaunchKernelGGL(k_copy_strided<T>, dim3(g), dim3(256), 0, s, A + P, stride,
                                   (const T *)(stash + op.p[4]), 1, count);
            TL_LAUNCH_CHECK();
            break;
        }
        case OP_CHAINED: {
            const uint64_t nmoves = op.p[0];
            const uint64_t words = ((nmoves + 15) / 16) * 16;
            if (!preclear) TL_HIP(hipMemsetAsync(op_flags, 0, (words + 16) * 4, s));
            unsigned long long *counter = (unsigned long long *)(op_flags + words);
            const uint32_t grid =
                (uint32_t)std::min<uint64_t>(nmoves, (uint64_t)c->num_cus * chained_wg_per_cu());
            if (op.ms.kind == MV_ROW_SPREAD)
                hipLaunchKernelGGL((k_chained<T, CHAINED_PER, CH_ROW_SPREAD>), dim3(grid), dim3(256),
                                   0, s, A, op.ms, nmoves, op_flags, counter, c->err_flag,
                                   c->wait_ticks, stash);
            else if (op.ms.kind == MV_SPARSE_SPREAD)
                hipLaunchKernelGGL((k_chained<T, CHAINED_PER, CH_SPARSE_SPREAD>), dim3(grid),
                                   dim3(256), 0, s, A, op.ms, nmoves, op_flags, counter,
                                   c->err_flag, c->wait_ticks, stash);
            else
                hipLaunchKernelGGL((k_chained<T, CHAINED_PER, CH_FORWARD>), dim3(grid), dim3(256),
                                   0, s, A, op.ms, nmoves, op_flags, counter, c->err_flag,
                                   c->wait_ticks, stash);
            TL_LAUNCH_CHECK();
            if (preclear) op_flags += words + 16;
            break;
        }
        }
    }
    return TL_OK;
}

template <typename T>
static tl_status run_plan(Ctx *c, Plan &pl, T *A, hipStream_t s) {
    uint64_t *blob = nullptr;
    if (!pl.blob.empty())
        TL_TRY(copy_to_device(c, 2, pl.blob.data(), pl.blob.size() * 8, s, (void **)&blob));
    T *stash = nullptr;
    if (pl.max_stash) TL_TRY(scratch(c, 0, pl.max_stash * sizeof(T), (void **)&stash));
    uint32_t *flags = nullptr;
    if (pl.max_moves) TL_TRY(scratch(c, 1, plan_flag_bytes(pl), (void **)&flags));
    return launch_plan<T>(c, pl, A, s, blob, stash, flags, false);
}

tl_status fill_u32_launch(uint32_t *p, uint64_t count, uint32_t v, hipStream_t s) {
    const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((count + 255) / 256, 1024));
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, p, count, v);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

tl_status plan_launch(Ctx *c, const Plan &pl, void *A, uint32_t W, hipStream_t s,
                      const uint64_t *dev_blob, void *stash, uint32_t *flags) {
    return W == 4 ? launch_plan<uint32_t>(c, pl, (uint32_t *)A, s, dev_blob, (uint32_t *)stash,
                                          flags, true)
                  : launch_plan<uint64_t>(c, pl, (uint64_t *)A, s, dev_blob, (uint64_t *)stash,
                                          flags, true);
}

// whole small arrays in LDS, one workgroup each: desc holds (address, n)
// pairs of arrays of one layout kind, b and key width, n <= max_n
template <typename T>
__global__ __launch_bounds__(512) void k_lds_items(const uint64_t *__restrict__ desc, uint32_t kind,
                                                   uint32_t b) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *s = reinterpret_cast<T *>(smem_raw);
    T *A = reinterpret_cast<T *>(desc[2 * blockIdx.x]);
    const uint32_t n = (uint32_t)desc[2 * blockIdx.x + 1];
    BtreeShape sh;
    if (kind == LDS_BTREE) sh = btree_shape(n, b);
    constexpr int U = 8;
    for (uint32_t i0 = threadIdx.x; i0 < n; i0 += blockDim.x * U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * blockDim.x;
            if (i < n) v[u] = A[i];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t i = i0 + u * blockDim.x;
            if (i < n) s[lds_slot(i, n, kind, b, 0, sh)] = v[u];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) A[i] = s[i];
}

tl_status lds_items_launch(const uint64_t *dev_desc, uint32_t count, uint32_t W, uint32_t kind,
                           uint32_t b, uint32_t max_n, hipStream_t s) {
    if (!count) return TL_OK;
    const size_t bytes = (size_t)max_n * W;
    const int threads = max_n >= 4096 ? 512 : 256;
    if (W == 4)
        hipLaunchKernelGGL(k_lds_items<uint32_t>, dim3(count), dim3(threads), bytes, s, dev_desc,
                           kind, b);
    else
        hipLaunchKernelGGL(k_lds_items<uint64_t>, dim3(count), dim3(threads), bytes, s, dev_desc,
                           kind, b);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// ---------------------------------------------------------------------------
// multi-array launches (descriptors: tl_plan.hpp)

// the descriptor owning task t: the last i with pre[i] <= t
__device__ inline uint32_t find_desc(const uint64_t *__restrict__ pre, uint32_t nd, uint64_t t) {
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

constexpr uint64_t MCOPY_CHUNK = 2048;

template <typename T>
__device__ inline void multi_copy_body(const MCopy &m, uint64_t part) {
    const uint64_t lo = part * MCOPY_CHUNK;
    const uint64_t hi = min(m.count, lo + MCOPY_CHUNK);
    T *dst = (T *)m.dst;
    const T *src = (const T *)m.src;
    // source and destination never overlap (array <-> stash), so every load
    // of the thread is issued before its first store
    constexpr int PER = (int)(MCOPY_CHUNK / 256);
    T v[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint64_t k = lo + threadIdx.x + (uint64_t)j * 256;
        if (k < hi) v[j] = src[k * m.sstride];
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint64_t k = lo + threadIdx.x + (uint64_t)j * 256;
        if (k < hi) dst[k * m.dstride] = v[j];
    }
}

// the copies of arrays of both key widths in one launch (the width is per
// descriptor)
__global__ __launch_bounds__(256) void k_multi_copy(const MCopy *__restrict__ d,
                                                    const uint64_t *__restrict__ pre, uint32_t nd) {
    const uint32_t i = find_desc(pre, nd, blockIdx.x);
    const MCopy m = d[i];
    if (m.w == 4)
        multi_copy_body<uint32_t>(m, blockIdx.x - pre[i]);
    else
        multi_copy_body<uint64_t>(m, blockIdx.x - pre[i]);
}

// in-LDS layouts of arrays of both key widths in one launch (the width is
// per descriptor; the dynamic LDS is sized for the largest)
__global__ __launch_bounds__(512) void k_multi_lds(const MLds *__restrict__ d,
                                                   const uint64_t *__restrict__ pre, uint32_t nd) {
    const uint32_t i = find_desc(pre, nd, blockIdx.x);
    const MLds m = d[i];
    if (m.W == 4)
        lds_body<uint32_t>((uint32_t *)m.A, (const uint64_t *)m.runs, m.nruns, m.nitems, m.per,
                           m.n, m.kind, m.b, m.rows, m.magic, m.inv, blockIdx.x - pre[i],
                           m.shift);
    else
        lds_body<uint64_t>((uint64_t *)m.A, (const uint64_t *)m.runs, m.nruns, m.nitems, m.per,
                           m.n, m.kind, m.b, m.rows, m.magic, m.inv, blockIdx.x - pre[i],
                           m.shift);
}

template <typename T, uint32_t NT>
__device__ inline void multi_stream_body(const MStream &m, uint64_t part) {
    T *a = (T *)m.A + run_offset((const uint64_t *)m.runs, m.nruns, part);
    if (m.fuse)
        stream_fused_body<T, NT>(a, m.r, m.l, m.inv, m.fuse);
    else
        stream_body<T>(a, m.r, m.l, m.inv);
}

// NT threads per partition (the stream's batches are NT * 8 keys; any batch
// size keeps every batch read before it is written); either key width, per
// descriptor
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_multi_stream(const MStream *__restrict__ d,
                                                     const uint64_t *__restrict__ pre,
                                                     uint32_t nd) {
    const uint32_t i = find_desc(pre, nd, blockIdx.x);
    const MStream m = d[i];
    if (m.W == 4)
        multi_stream_body<uint32_t, NT>(m, blockIdx.x - pre[i]);
    else
        multi_stream_body<uint64_t, NT>(m, blockIdx.x - pre[i]);
}

// fused streamed gathers of many arrays, NT threads per partition: a
// partition's workgroup finishes NT / 256 times sooner than with 256 threads,
// so a launch whose partitions fill the device a little over a whole number
// of times loses less to its last, partial round
template <uint32_t NT>
__global__ __launch_bounds__(NT) void k_multi_stream_fused(const MStream *__restrict__ d,
                                                           const uint64_t *__restrict__ pre,
                                                           uint32_t nd) {
    const uint32_t i = find_desc(pre, nd, blockIdx.x);
    const MStream m = d[i];
    const uint64_t part = blockIdx.x - pre[i];
    if (m.W == 4)
        stream_fused_body<uint32_t, NT>(
            (uint32_t *)m.A + run_offset((const uint64_t *)m.runs, m.nruns, part), m.r, m.l,
            m.inv, m.fuse);
    else
        stream_fused_body<uint64_t, NT>(
            (uint64_t *)m.A + run_offset((const uint64_t *)m.runs, m.nruns, part), m.r, m.l,
            m.inv, m.fuse);
}

// moves of many chained ops, numbered op after op: a move waits only for
// lower-numbered moves of its own op, which hold lower tickets.  The ops'
// arrays may have either key width (per descriptor; a workgroup's move
// belongs to one op, so its branch is uniform)
template <int PER, int MODE>
__global__ __launch_bounds__(256) void k_multi_chained(const MChain *__restrict__ d,
                                                       const uint64_t *__restrict__ pre, uint32_t nd,
                                                       uint64_t total, unsigned long long *counter,
                                                       uint32_t *err, uint64_t limit) {
    __shared__ uint64_t s_t;
    __shared__ uint32_t s_slot[FUSED_LP_MAX];
    uint64_t have = 0;   // (lp, fuse) of the slot table in s_slot; the same in every lane
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t t = s_t;
        __syncthreads();
        if (t >= total) return;
        const uint32_t i = find_desc(pre, nd, t);
        const MChain m = d[i];
        if (MODE != CH_ROW_SPREAD && m.ms.fuse) {
            // a fused sparse move: the vEB slots of its array's lp-key blocks
            // (every lane passed the barriers above since the last table use)
            const uint64_t want = (m.ms.lp << 2) | m.ms.fuse;
            if (want != have) {
                fill_fused_slots(m.ms, s_slot);
                have = want;
            }
        }
        if (m.W == 4)
            chained_move<uint32_t, PER, MODE>((uint32_t *)m.A, m.ms, t - pre[i],
                                              (uint32_t *)m.flags, err, limit,
                                              (uint32_t *)m.stash, s_slot);
        else
            chained_move<uint64_t, PER, MODE>((uint64_t *)m.A, m.ms, t - pre[i],
                                              (uint32_t *)m.flags, err, limit,
                                              (uint64_t *)m.stash, s_slot);
    }
}

// OP_PART_ROOTS of many arrays: 256 roots per workgroup
template <typename T>
__device__ inline void multi_proots_body(const MPRoots &m, uint64_t g) {
    const uint64_t blk = g / m.rp;
    const uint32_t a = (uint32_t)(g - blk * m.rp);
    T *slot = (T *)m.A + m.base + blk * m.stride + mveb_pos32(a, m.rp, m.classic != 0);
    T *st = (T *)m.stash + g;
    if (m.inv)
        *st = *slot;
    else
        *slot = *st;
}

// either key width, per descriptor
__global__ __launch_bounds__(256) void k_multi_proots(const MPRoots *__restrict__ d,
                                                      const uint64_t *__restrict__ pre,
                                                      uint32_t nd) {
    const uint64_t t = blockIdx.x;
    const uint32_t i = find_desc(pre, nd, t);
    const MPRoots m = d[i];
    const uint64_t g = (t - pre[i]) * 256 + threadIdx.x;
    if (g >= m.count * m.rp) return;
    if (m.W == 4)
        multi_proots_body<uint32_t>(m, g);
    else
        multi_proots_body<uint64_t>(m, g);
}

// SHIFT = 1: every descriptor is the levels of a BST kept one place right,
// written one slot left (BITREV_SHIFT)
template <typename T, int K, int SHIFT>
__device__ inline void multi_bitrev_body(const MBitrev &m, uint64_t g) {
    if constexpr (SHIFT == 1) {
        level_tile<T, K, 1>((T *)m.A, m.t0, m.d, g, (T *)m.stash);
    } else if (m.levels) {
        const uint32_t j = 63u - (uint32_t)__clzll(g + 1);
        const uint32_t lev = m.t0 + j;
        bitrev_tile<T, K>((T *)m.A, (int64_t)((1ull << lev) - 1), lev, g - ((1ull << j) - 1));
    } else {
        bitrev_tile<T, K>((T *)m.A, m.base, m.D, g);
    }
}

// either key width, per descriptor (the tile buffer is the 8-byte one)
template <int K, int SHIFT = 0>
__global__ __launch_bounds__(256) void k_multi_bitrev(const MBitrev *__restrict__ d,
                                                      const uint64_t *__restrict__ pre,
                                                      uint32_t nd) {
    const uint64_t t = (uint64_t)blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
    if (t >= pre[nd]) return;
    const uint32_t i = find_desc(pre, nd, t);
    const MBitrev m = d[i];
    if (m.W == 4)
        multi_bitrev_body<uint32_t, K, SHIFT>(m, t - pre[i]);
    else
        multi_bitrev_body<uint64_t, K, SHIFT>(m, t - pre[i]);
}

// the walks of arrays of both key widths in one launch (the width is per
// descriptor)
template <int BS = 256, int K = 4>
__global__ __launch_bounds__(BS) void k_multi_walk(const MWalk *__restrict__ d,
                                                   const uint64_t *__restrict__ pre, uint32_t nd) {
    const uint32_t i = find_desc(pre, nd, blockIdx.x);
    const MWalk m = d[i];
    const uint64_t j = (blockIdx.x - pre[i]) * BS + threadIdx.x;
    if (j >= m.w) return;
    if (m.W == 4)
        walk_lane<uint32_t, K>((uint32_t *)m.A, m.L, m.c, m.w, (const uint32_t *)m.stash, j);
    else
        walk_lane<uint64_t, K>((uint64_t *)m.A, m.L, m.c, m.w, (const uint64_t *)m.stash, j);
}

// steps of a multi-array rotation lane walk with their loads in flight
// together: $TL_MULTI_WALK_DEPTH (4, 8 or 16), else 16.  The batch's arrays
// are small, so its walks are latency-bound: 16 loads in flight cover most
// walks (about 16 steps each) in one round trip (the 4096-array batch's
// permutes: 12.92 ms with 4, 12.43 ms with 16).
static int multi_walk_depth() {
    const char *e = getenv("TL_MULTI_WALK_DEPTH");
    const int v = e ? atoi(e) : 16;
    return (v == 4 || v == 8) ? v : 16;
}

static tl_status multi_launch_body(Ctx *c, const MultiLaunch &m, hipStream_t s) {
    if (!m.nd || !m.tasks) return TL_OK;
    const uint64_t *pre = (const uint64_t *)m.pre;
    switch (m.kind) {
    case MK_COPY:
        hipLaunchKernelGGL(k_multi_copy, dim3((uint32_t)m.tasks), dim3(256), 0, s,
                           (const MCopy *)m.desc, pre, m.nd);
        break;
    case MK_LDS:
        hipLaunchKernelGGL(k_multi_lds, dim3((uint32_t)m.tasks), dim3(m.threads), m.lds, s,
                           (const MLds *)m.desc, pre, m.nd);
        break;
    case MK_STREAM:
        if (m.threads == 1024)
            hipLaunchKernelGGL((k_multi_stream<1024>), dim3((uint32_t)m.tasks), dim3(1024),
                               m.lds, s, (const MStream *)m.desc, pre, m.nd);
        else if (m.threads == 512)
            hipLaunchKernelGGL((k_multi_stream<512>), dim3((uint32_t)m.tasks), dim3(512), m.lds,
                               s, (const MStream *)m.desc, pre, m.nd);
        else
            hipLaunchKernelGGL((k_multi_stream<256>), dim3((uint32_t)m.tasks), dim3(256), m.lds,
                               s, (const MStream *)m.desc, pre, m.nd);
        break;
    case MK_FUSED:
        if (m.threads == 1024)
            hipLaunchKernelGGL((k_multi_stream_fused<1024>), dim3((uint32_t)m.tasks), dim3(1024),
                               m.lds, s, (const MStream *)m.desc, pre, m.nd);
        else if (m.threads == 512)
            hipLaunchKernelGGL((k_multi_stream_fused<512>), dim3((uint32_t)m.tasks), dim3(512),
                               m.lds, s, (const MStream *)m.desc, pre, m.nd);
        else
            hipLaunchKernelGGL((k_multi_stream_fused<256>), dim3((uint32_t)m.tasks), dim3(256),
                               m.lds, s, (const MStream *)m.desc, pre, m.nd);
        break;
    case MK_CHAIN: {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(m.tasks, (uint64_t)c->num_cus * 8);
        unsigned long long *ctr = (unsigned long long *)m.counter;
        if (m.param == CH_ROW_SPREAD)
            hipLaunchKernelGGL((k_multi_chained<CHAINED_PER, CH_ROW_SPREAD>), dim3(grid),
                               dim3(256), 0, s, (const MChain *)m.desc, pre, m.nd, m.tasks, ctr,
                               c->err_flag, c->wait_ticks);
        else if (m.param == CH_SPARSE_SPREAD)
            hipLaunchKernelGGL((k_multi_chained<CHAINED_PER, CH_SPARSE_SPREAD>), dim3(grid),
                               dim3(256), 0, s, (const MChain *)m.desc, pre, m.nd, m.tasks, ctr,
                               c->err_flag, c->wait_ticks);
        else
            hipLaunchKernelGGL((k_multi_chained<CHAINED_PER, CH_FORWARD>), dim3(grid),
                               dim3(256), 0, s, (const MChain *)m.desc, pre, m.nd, m.tasks, ctr,
                               c->err_flag, c->wait_ticks);
        break;
    }
    case MK_BITREV: {
        const uint32_t gx = (uint32_t)std::min<uint64_t>(m.tasks, 1u << 20);
        const uint32_t gy = (uint32_t)((m.tasks + gx - 1) / gx);
        // param: K, and 1 << 8 for the shifted levels of kept trees
        const bool shift = (m.param >> 8) == 1;
        if ((m.param & 0xff) == 6) {
            if (shift)
                hipLaunchKernelGGL((k_multi_bitrev<6, 1>), dim3(gx, gy), dim3(256), 0, s,
                                   (const MBitrev *)m.desc, pre, m.nd);
            else
                hipLaunchKernelGGL((k_multi_bitrev<6>), dim3(gx, gy), dim3(256), 0, s,
                                   (const MBitrev *)m.desc, pre, m.nd);
        } else {
            if (shift)
                hipLaunchKernelGGL((k_multi_bitrev<5, 1>), dim3(gx, gy), dim3(256), 0, s,
                                   (const MBitrev *)m.desc, pre, m.nd);
            else
                hipLaunchKernelGGL((k_multi_bitrev<5>), dim3(gx, gy), dim3(256), 0, s,
                                   (const MBitrev *)m.desc, pre, m.nd);
        }
        break;
    }
    case MK_WALK:
        if (m.threads == 1024 && multi_walk_depth() == 16)
            hipLaunchKernelGGL((k_multi_walk<1024, 16>), dim3((uint32_t)m.tasks), dim3(1024), 0,
                               s, (const MWalk *)m.desc, pre, m.nd);
        else if (m.threads == 1024 && multi_walk_depth() == 8)
            hipLaunchKernelGGL((k_multi_walk<1024, 8>), dim3((uint32_t)m.tasks), dim3(1024), 0,
                               s, (const MWalk *)m.desc, pre, m.nd);
        else if (m.threads == 1024)
            hipLaunchKernelGGL((k_multi_walk<1024>), dim3((uint32_t)m.tasks), dim3(1024), 0, s,
                               (const MWalk *)m.desc, pre, m.nd);
        else if (multi_walk_depth() == 16)
            hipLaunchKernelGGL((k_multi_walk<256, 16>), dim3((uint32_t)m.tasks), dim3(256), 0, s,
                               (const MWalk *)m.desc, pre, m.nd);
        else
            hipLaunchKernelGGL(k_multi_walk<>, dim3((uint32_t)m.tasks), dim3(256), 0, s,
                               (const MWalk *)m.desc, pre, m.nd);
        break;
    case MK_PROOTS:
        hipLaunchKernelGGL(k_multi_proots, dim3((uint32_t)m.tasks), dim3(256), 0, s,
                           (const MPRoots *)m.desc, pre, m.nd);
        break;
    default: set_error("multi_launch: unknown kind"); return TL_EINVAL;
    }
    TL_LAUNCH_CHECK();
    return TL_OK;
}

// every multi-array kernel takes the key width from its descriptors, so
// arrays of both widths share a launch
tl_status multi_launch(Ctx *c, const MultiLaunch &m, hipStream_t s) {
    return multi_launch_body(c, m, s);
}

// one workgroup waits for a flag that nothing sets, bounded by `limit`
// wall-clock ticks: the bounded wait must give up and raise the error word
__global__ void k_selftest_wait(uint32_t *never, uint32_t *err, uint64_t limit) {
    if (threadIdx.x == 0) (void)wait_flag(never, err, limit);
}

static tl_status check_args(Layout lay, const void *dev_A, uint64_t n, uint64_t b, tl_dtype dt,
                            uint32_t &W) {
    if (!width_of(dt, W)) {
        set_error("unknown dtype");
        return TL_EINVAL;
    }
    if (lay == L_BTREE && (b == 0 || b > 4096)) {
        set_error("b must be in [1, 4096]");
        return TL_EINVAL;
    }
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

static tl_status permute(Layout lay, void *dev_A, uint64_t n, uint64_t b, tl_dtype dt,
                         void *stream, float *ms, bool inverse = false) {
    uint32_t W;
    TL_TRY(check_args(lay, dev_A, n, b, dt, W));
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Plan pl;
    build_plan(pl, tuning_for(c), lay, n, b, W);
    if (inverse) {
        Plan fwd;
        fwd.ops.swap(pl.ops);
        fwd.blob.swap(pl.blob);
        fwd.max_stash = pl.max_stash;
        fwd.max_moves = pl.max_moves;
        invert_plan(fwd, pl);
    }
    std::string why;
    if (!validate_plan(pl, n, why)) {
        set_error(why);
        return TL_EINVAL;
    }
    Timer tm;
    TL_TRY(tm.start(s, ms));
    tl_status st = (W == 4) ? run_plan<uint32_t>(c, pl, (uint32_t *)dev_A, s)
                            : run_plan<uint64_t>(c, pl, (uint64_t *)dev_A, s);
    if (st != TL_OK) return st;
    TL_TRY(tm.stop());
    bool chained = false;
    for (auto &op : pl.ops)
        if (op.type == OP_CHAINED) chained = true;
    if (chained && ms) TL_TRY(check_err_flag(c, s));
    return TL_OK;
}

// tl_permute_btree with all scratch in the device's static array (Ctx::noalloc):
// the plan is built with every stash capped at the array's stash part, its
// tables are copied into the array, its flags cleared by a fill kernel (every
// chained op has a region of its own), and the launches take the array's
// addresses.  No allocation, no memset.
static tl_status permute_btree_noalloc(void *dev_A, uint64_t n, uint64_t b, tl_dtype dt,
                                       void *stream) {
    uint32_t W;
    TL_TRY(check_args(L_BTREE, dev_A, n, b, dt, W));
    if (n > TL_NOALLOC_MAX_KEYS) {
        set_error("tl_permute_btree_noalloc: n = " + std::to_string(n) +
                  " exceeds TL_NOALLOC_MAX_KEYS = " + std::to_string(TL_NOALLOC_MAX_KEYS));
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    Plan pl;
    build_plan(pl, tuning_for(c), L_BTREE, n, b, W, NOALLOC_STASH_BYTES / W);
    std::string why;
    if (!validate_plan(pl, n, why)) {
        set_error(why);
        return TL_EINVAL;
    }
    const size_t blob_bytes = pl.blob.size() * 8;
    const size_t flag_bytes = plan_flag_bytes_all(pl);
    const size_t stash_bytes = pl.max_stash * W;
    if (blob_bytes > NOALLOC_BLOB_BYTES || flag_bytes > NOALLOC_FLAG_BYTES ||
        stash_bytes > NOALLOC_STASH_BYTES) {
        set_error("tl_permute_btree_noalloc: the plan for n = " + std::to_string(n) +
                  " needs " + std::to_string(blob_bytes) + " table, " +
                  std::to_string(flag_bytes) + " flag and " + std::to_string(stash_bytes) +
                  " stash bytes; the static scratch holds " +
                  std::to_string(NOALLOC_BLOB_BYTES) + ", " + std::to_string(NOALLOC_FLAG_BYTES) +
                  " and " + std::to_string(NOALLOC_STASH_BYTES));
        return TL_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    uint64_t *blob = (uint64_t *)c->noalloc;
    uint32_t *flags = (uint32_t *)(c->noalloc + NOALLOC_BLOB_BYTES);
    void *stash = c->noalloc + NOALLOC_BLOB_BYTES + NOALLOC_FLAG_BYTES;
    if (blob_bytes) {
        // stream-ordered after earlier calls on `s`; the host vector dies here
        TL_HIP(hipMemcpyAsync(blob, pl.blob.data(), blob_bytes, hipMemcpyHostToDevice, s));
        TL_HIP(hipStreamSynchronize(s));
    }
    if (flag_bytes) TL_TRY(fill_u32_launch(flags, flag_bytes / 4, 0u, s));
    return plan_launch(c, pl, dev_A, W, s, blob, stash, flags);
}

// Serialized plan: per op 32 words
//   [type, nruns, nitems, p0..p7, ms.kind, ms.P, ms.L, ms.k, ms.r, ms.l, ms.CH, ms.nchunk,
//    runs_at, ms.fuse, ms.rp, ms.lp, ms.soff, 0...] followed by the blob.
static tl_status export_plan(Layout lay, uint64_t n, uint64_t b, tl_dtype dt, uint64_t lds_bytes,
                             uint64_t *buf, uint64_t cap, uint64_t *used, bool inverse) {
    uint32_t W;
    TL_TRY(check_args(lay, (void *)1, n, b, dt, W));
    Tuning tu;
    tuning_env(tu);
    if (lds_bytes) tu.lds_bytes = lds_bytes;
    Plan pl;
    build_plan(pl, tu, lay, n, b, W);
    if (inverse) {
        Plan fwd = pl;
        invert_plan(fwd, pl);
    }
    std::string why;
    if (!validate_plan(pl, n, why)) {
        set_error(why);
        return TL_EINVAL;
    }
    const uint64_t need = 2 + pl.ops.size() * 32 + pl.blob.size();
    if (used) *used = need;
    if (!buf || cap < need) return TL_OK;
    buf[0] = pl.ops.size();
    buf[1] = pl.blob.size();
    uint64_t *o = buf + 2;
    for (const Op &op : pl.ops) {
        uint64_t w[32] = {0};
        w[0] = op.type;
        w[1] = op.nruns;
        w[2] = op.nitems;
        for (int i = 0; i < 8; i++) w[3 + i] = op.p[i];
        w[11] = op.ms.kind;
        w[12] = op.ms.P;
        w[13] = op.ms.L;
        w[14] = op.ms.k;
        w[15] = op.ms.r;
        w[16] = op.ms.l;
        w[17] = op.ms.CH;
        w[18] = op.ms.nchunk;
        w[19] = op.runs_at;
        w[20] = op.ms.fuse;
        w[21] = op.ms.rp;
        w[22] = op.ms.lp;
        w[23] = op.ms.soff;
        for (int i = 0; i < 32; i++) *o++ = w[i];
    }
    for (uint64_t v : pl.blob) *o++ = v;
    return TL_OK;
}

}   // namespace tl

extern "C" tl_status tl_permute_bst(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                    float *ms) {
    return tl::permute(tl::L_BST, dev_A, n, 1, dtype, stream, ms);
}

extern "C" tl_status tl_permute_btree(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                      void *stream, float *ms) {
    return tl::permute(tl::L_BTREE, dev_A, n, b, dtype, stream, ms);
}

extern "C" tl_status tl_permute_mveb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                     float *ms) {
    return tl::permute(tl::L_MVEB, dev_A, n, 0, dtype, stream, ms);
}

extern "C" tl_status tl_permute_veb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                    float *ms) {
    return tl::permute(tl::L_VEB, dev_A, n, 0, dtype, stream, ms);
}

extern "C" tl_status tl_unpermute_veb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                      float *ms) {
    return tl::permute(tl::L_VEB, dev_A, n, 0, dtype, stream, ms, true);
}

extern "C" tl_status tl_unpermute_bst(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                      float *ms) {
    return tl::permute(tl::L_BST, dev_A, n, 1, dtype, stream, ms, true);
}

extern "C" tl_status tl_unpermute_btree(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                        void *stream, float *ms) {
    return tl::permute(tl::L_BTREE, dev_A, n, b, dtype, stream, ms, true);
}

extern "C" tl_status tl_unpermute_mveb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                       float *ms) {
    return tl::permute(tl::L_MVEB, dev_A, n, 0, dtype, stream, ms, true);
}

extern "C" tl_status tl_permute_btree_noalloc(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                              void *stream) {
    return tl::permute_btree_noalloc(dev_A, n, b, dtype, stream);
}

extern "C" tl_status tl_selftest_timeout(uint32_t wait_ms, void *stream, float *ms) {
    using namespace tl;
    if (wait_ms == 0 || wait_ms > 60000) {
        set_error("wait_ms must be in [1, 60000]");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    void *never = nullptr;
    TL_TRY(scratch(c, 3, 256, &never));
    TL_TRY(fill_u32_launch((uint32_t *)never, 64, 0u, s));
    Timer tm;
    float t = 0.f;
    TL_TRY(tm.start(s, &t));
    hipLaunchKernelGGL(k_selftest_wait, dim3(1), dim3(64), 0, s, (uint32_t *)never, c->err_flag,
                       (uint64_t)wait_ms * c->wall_khz);
    TL_LAUNCH_CHECK();
    TL_TRY(tm.stop());
    if (ms) *ms = t;
    return check_err_flag(c, s);
}

extern "C" tl_status tl_plan_export(uint32_t layout, uint64_t n, uint64_t b, tl_dtype dtype,
                                    uint64_t lds_bytes, uint64_t *buf, uint64_t cap,
                                    uint64_t *used) {
    const bool inverse = (layout & tl::PLAN_INVERSE) != 0;
    layout &= ~tl::PLAN_INVERSE;
    if (layout > 3) {
        tl::set_error("layout must be 0 (BST), 1 (B-tree), 2 (mvEB) or 3 (classic vEB), "
                      "optionally | 0x100");
        return TL_EINVAL;
    }
    return tl::export_plan((tl::Layout)layout, n, b, dtype, lds_bytes, buf, cap, used, inverse);
}
