// tl_sort.hip -- the step before the layouts: sort raw keys and drop
// duplicates on the device, so that an unsorted key set with repeats becomes
// the sorted, distinct input every layout assumes (the reference's drivers
// build such input themselves, cpu/src/common.h:33-40).
//
// LSD radix sort, 8-bit digits, stable, ping-ponging between the caller's
// array and one library-owned buffer of n keys.  Signed keys sort by their
// bits with the sign bit flipped.  A pass whose digit is the same for every
// key (found by one AND/OR reduction over all keys) is skipped: keys built
// from small gaps leave their top bytes constant.  Per pass:
//   k_radix_hist     per tile of 8192 keys, a 256-bucket histogram, stored
//                    digit-major (counts[d * tiles + tile]);
//   scan             exclusive scan of those counts (the global offset of
//                    every (digit, tile) bucket);
//   k_radix_scatter  every tile again: each wave ranks its keys by digit with
//                    8 ballots per key slot (lanes with equal digits find each
//                    other), waves are offset by the earlier waves' counts,
//                    and every key goes to its bucket's next slot.
// Keys are processed in (wave, slot, lane) order, which is their order in
// memory, so equal digits keep their order (the sort is stable).
// Duplicates: every sorted key that differs from its predecessor is kept;
// a count per tile, a scan, and a compaction into the other buffer.
#include <algorithm>
#include <vector>

#include "tl_internal.hpp"
#include "tl_plan.hpp"

namespace tl {

constexpr uint32_t SORT_THREADS = 256;
constexpr uint32_t SORT_ITEMS = 32;
constexpr uint32_t SORT_TILE = SORT_THREADS * SORT_ITEMS;   // keys per tile
constexpr uint32_t SORT_WAVES = SORT_THREADS / 64;
constexpr uint32_t SCAN_CHUNK = 4096;                        // counts per scan workgroup

// AND and OR of every key's order bits (bits ^ flip)
template <typename U>
__global__ __launch_bounds__(256) void k_and_or(const U *__restrict__ a, uint64_t n, U flip,
                                                unsigned long long *andor) {
    U va = ~(U)0, vo = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const U u = a[i] ^ flip;
        va &= u;
        vo |= u;
    }
    __shared__ unsigned long long sa[256], so[256];
    sa[threadIdx.x] = (unsigned long long)va | (sizeof(U) == 4 ? 0xFFFFFFFF00000000ull : 0ull);
    so[threadIdx.x] = (unsigned long long)vo;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] &= sa[threadIdx.x + s];
            so[threadIdx.x] |= so[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAnd(&andor[0], sa[0]);
        atomicOr(&andor[1], so[0]);
    }
}

template <typename U>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const U *__restrict__ a, uint64_t n,
                                                             uint32_t shift, U flip,
                                                             uint32_t *__restrict__ counts,
                                                             uint32_t tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    const uint64_t end = min(n, base + SORT_TILE);
    for (uint64_t i = base + threadIdx.x; i < end; i += SORT_THREADS)
        atomicAdd(&h[(uint32_t)(((a[i] ^ flip) >> shift) & 0xFF)], 1u);
    __syncthreads();
    counts[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of v[0..len) in chunks of SCAN_CHUNK: local scan, chunk sums
__global__ __launch_bounds__(256) void k_scan_chunks(uint32_t *__restrict__ v, uint64_t len,
                                                     uint32_t *__restrict__ sums) {
    constexpr uint32_t PER = SCAN_CHUNK / 256;
    __shared__ uint32_t part[256];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK + (uint64_t)threadIdx.x * PER;
    uint32_t x[PER], t = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        x[k] = base + k < len ? v[base + k] : 0u;
        const uint32_t y = x[k];
        x[k] = t;
        t += y;
    }
    part[threadIdx.x] = t;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 256 thread totals
    for (uint32_t s = 1; s < 256; s <<= 1) {
        const uint32_t y = threadIdx.x >= s ? part[threadIdx.x - s] : 0u;
        __syncthreads();
        part[threadIdx.x] += y;
        __syncthreads();
    }
    const uint32_t off = threadIdx.x ? part[threadIdx.x - 1] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++)
        if (base + k < len) v[base + k] = x[k] + off;
    if (threadIdx.x == 255) sums[blockIdx.x] = part[255];
}

// exclusive scan of the chunk sums in one workgroup (any count), then the
// grand total at sums[len]
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *__restrict__ sums, uint64_t len) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < len; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t x = i < len ? sums[i] : 0u;
        part[threadIdx.x] = x;
        __syncthreads();
        for (uint32_t s = 1; s < 1024; s <<= 1) {
            const uint32_t y = threadIdx.x >= s ? part[threadIdx.x - s] : 0u;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < len) sums[i] = carry + part[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[len] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t *__restrict__ v, uint64_t len,
                                                  const uint32_t *__restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    const uint32_t add = sums[blockIdx.x];
    for (uint64_t i = base + threadIdx.x; i < min(len, base + SCAN_CHUNK); i += 256) v[i] += add;
}

// (wave, slot, lane) order of a tile: element e of wave w, slot k, lane l
__device__ inline uint64_t tile_elem(uint64_t base, uint32_t w, uint32_t k, uint32_t lane) {
    return base + (uint64_t)w * (64 * SORT_ITEMS) + (uint64_t)k * 64 + lane;
}

template <typename U>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const U *__restrict__ a,
                                                                U *__restrict__ out, uint64_t n,
                                                                uint32_t shift, U flip,
                                                                const uint32_t *__restrict__ offs,
                                                                uint32_t tiles) {
    __shared__ uint32_t cnt[SORT_WAVES][256];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < SORT_WAVES * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    const uint64_t lt = (1ull << lane) - 1;
    U key[SORT_ITEMS];
    uint32_t rank[SORT_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < SORT_ITEMS; k++) {
        const uint64_t e = tile_elem(base, w, k, lane);
        const bool valid = e < n;
        key[k] = valid ? a[e] : (U)0;
        const uint32_t d = (uint32_t)(((key[k] ^ flip) >> shift) & 0xFF);
        // lanes whose digit equals mine: 8 ballots, one per digit bit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (uint32_t bit = 0; bit < 8; bit++) {
            const bool on = (d >> bit) & 1u;
            const uint64_t m = __ballot(on);
            peers &= on ? m : ~m;
        }
        uint32_t r = 0;
        if (valid) {
            // every peer reads the bucket before the lowest peer advances it
            // (one wave: the reads of this slot issue before its write)
            r = cnt[w][d] + (uint32_t)__popcll(peers & lt);
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & lt) == 0) cnt[w][d] += (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[k] = r;
    }
    __syncthreads();
    // bucket d of wave w starts at the tile's global offset of d plus the
    // counts of d in the earlier waves
    {
        const uint32_t d = threadIdx.x;   // SORT_THREADS == 256 digits
        uint32_t run = offs[(uint64_t)d * tiles + blockIdx.x];
        for (uint32_t ww = 0; ww < SORT_WAVES; ww++) {
            const uint32_t c = cnt[ww][d];
            cnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < SORT_ITEMS; k++) {
        const uint64_t e = tile_elem(base, w, k, lane);
        if (e < n) {
            const uint32_t d = (uint32_t)(((key[k] ^ flip) >> shift) & 0xFF);
            out[cnt[w][d] + rank[k]] = key[k];
        }
    }
}

// keys of a sorted tile that differ from their predecessor: count per tile
template <typename U>
__global__ __launch_bounds__(SORT_THREADS) void k_unique_count(const U *__restrict__ a, uint64_t n,
                                                               uint32_t *__restrict__ counts) {
    __shared__ uint32_t c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    const uint64_t end = min(n, base + SORT_TILE);
    uint32_t mine = 0;
    for (uint64_t i = base + threadIdx.x; i < end; i += SORT_THREADS)
        mine += (i == 0 || a[i] != a[i - 1]) ? 1u : 0u;
    atomicAdd(&c, mine);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

template <typename U>
__global__ __launch_bounds__(SORT_THREADS) void k_unique_scatter(const U *__restrict__ a,
                                                                 U *__restrict__ out, uint64_t n,
                                                                 const uint32_t *__restrict__ offs) {
    __shared__ uint32_t wsum[SORT_WAVES];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t before = 0;   // kept keys of this wave in earlier slots
    uint32_t pos[SORT_ITEMS];
    bool keep[SORT_ITEMS];
    U key[SORT_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < SORT_ITEMS; k++) {
        const uint64_t e = tile_elem(base, w, k, lane);
        keep[k] = false;
        key[k] = 0;
        if (e < n) {
            key[k] = a[e];
            keep[k] = (e == 0 || key[k] != a[e - 1]);
        }
        const uint64_t m = __ballot(keep[k]);
        pos[k] = before + (uint32_t)__popcll(m & lt);
        before += (uint32_t)__popcll(m);
    }
    if (lane == 0) wsum[w] = before;
    __syncthreads();
    uint32_t woff = offs[blockIdx.x];
    for (uint32_t ww = 0; ww < w; ww++) woff += wsum[ww];
#pragma unroll
    for (uint32_t k = 0; k < SORT_ITEMS; k++)
        if (keep[k]) out[woff + pos[k]] = key[k];
}

// exclusive scan of v[0..len) in place; the total lands in *total_dev
// (sums needs len / SCAN_CHUNK + 2 words)
static tl_status scan_u32(uint32_t *v, uint64_t len, uint32_t *sums, hipStream_t s) {
    const uint64_t chunks = (len + SCAN_CHUNK - 1) / SCAN_CHUNK;
    hipLaunchKernelGGL(k_scan_chunks, dim3((uint32_t)chunks), dim3(256), 0, s, v, len, sums);
    TL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, chunks);
    TL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)chunks), dim3(256), 0, s, v, len,
                       (const uint32_t *)sums);
    TL_LAUNCH_CHECK();
    return TL_OK;
}

template <typename U>
static tl_status sort_dedupe_t(Ctx *c, U *keys, uint64_t n, U flip, uint64_t *n_out,
                               hipStream_t s) {
    if (n <= 1) {
        *n_out = n;
        return TL_OK;
    }
    const uint64_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint64_t clen = 256 * tiles;
    const uint64_t csums = (clen + SCAN_CHUNK - 1) / SCAN_CHUNK + 2;
    U *tmp = nullptr;
    TL_TRY(scratch(c, 4, n * sizeof(U), (void **)&tmp));
    uint32_t *counts = nullptr;
    // counts, scan sums, and the AND/OR words (16-byte aligned at the end)
    const size_t cbytes = (clen + csums) * 4;
    TL_TRY(scratch(c, 5, cbytes + 64, (void **)&counts));
    uint32_t *sums = counts + clen;
    unsigned long long *andor =
        (unsigned long long *)(((uintptr_t)(counts + clen + csums) + 15) & ~(uintptr_t)15);
    const unsigned long long init[2] = {~0ull, 0ull};
    TL_HIP(hipMemcpyAsync(andor, init, sizeof(init), hipMemcpyHostToDevice, s));
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)c->num_cus * 8);
    hipLaunchKernelGGL(k_and_or<U>, dim3(g), dim3(256), 0, s, (const U *)keys, n, flip, andor);
    TL_LAUNCH_CHECK();
    unsigned long long ao[2];
    TL_HIP(hipMemcpyAsync(ao, andor, sizeof(ao), hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    const uint64_t varies = (uint64_t)(ao[0] ^ ao[1]);   // bits not equal across all keys
    U *src = keys, *dst = tmp;
    for (uint32_t shift = 0; shift < 8 * sizeof(U); shift += 8) {
        if (((varies >> shift) & 0xFF) == 0) continue;   // one digit value: nothing to move
        hipLaunchKernelGGL(k_radix_hist<U>, dim3((uint32_t)tiles), dim3(SORT_THREADS), 0, s,
                           (const U *)src, n, shift, flip, counts, (uint32_t)tiles);
        TL_LAUNCH_CHECK();
        TL_TRY(scan_u32(counts, clen, sums, s));
        hipLaunchKernelGGL(k_radix_scatter<U>, dim3((uint32_t)tiles), dim3(SORT_THREADS), 0, s,
                           (const U *)src, dst, n, shift, flip, (const uint32_t *)counts,
                           (uint32_t)tiles);
        TL_LAUNCH_CHECK();
        std::swap(src, dst);
    }
    // drop repeats: sorted src -> dst
    hipLaunchKernelGGL(k_unique_count<U>, dim3((uint32_t)tiles), dim3(SORT_THREADS), 0, s,
                       (const U *)src, n, counts);
    TL_LAUNCH_CHECK();
    TL_TRY(scan_u32(counts, tiles, sums, s));
    hipLaunchKernelGGL(k_unique_scatter<U>, dim3((uint32_t)tiles), dim3(SORT_THREADS), 0, s,
                       (const U *)src, dst, n, (const uint32_t *)counts);
    TL_LAUNCH_CHECK();
    uint32_t m = 0;
    const uint64_t chunks = (tiles + SCAN_CHUNK - 1) / SCAN_CHUNK;
    TL_HIP(hipMemcpyAsync(&m, sums + chunks, 4, hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    if (dst != keys && m)
        TL_HIP(hipMemcpyAsync(keys, dst, (size_t)m * sizeof(U), hipMemcpyDeviceToDevice, s));
    *n_out = m;
    return TL_OK;
}

}   // namespace tl

extern "C" tl_status tl_sort_dedupe(void *dev_keys, uint64_t n, tl_dtype dtype, uint64_t *n_out,
                                    void *stream, float *ms) {
    using namespace tl;
    uint32_t W;
    if (!width_of(dtype, W)) {
        set_error("unknown dtype");
        return TL_EINVAL;
    }
    if (!n_out || (n > 1 && !dev_keys)) {
        set_error("tl_sort_dedupe: null pointer");
        return TL_EINVAL;
    }
    if (n >= (1ull << 32)) {
        set_error("tl_sort_dedupe: n must be below 2^32");
        return TL_EINVAL;
    }
    Ctx *c;
    TL_TRY(get_ctx(&c));
    hipStream_t s = (hipStream_t)stream;
    Timer tm;
    TL_TRY(tm.start(s, ms));
    tl_status st;
    if (W == 4) {
        const uint32_t flip = dtype == TL_I32 ? 0x80000000u : 0u;
        st = sort_dedupe_t<uint32_t>(c, (uint32_t *)dev_keys, n, flip, n_out, s);
    } else {
        const uint64_t flip = dtype == TL_I64 ? 0x8000000000000000ull : 0ull;
        st = sort_dedupe_t<uint64_t>(c, (uint64_t *)dev_keys, n, flip, n_out, s);
    }
    if (st != TL_OK) return st;
    TL_TRY(tm.stop());
    return TL_OK;
}
