// tl_internal.hpp -- shared host/device helpers of the gfx950 Tree-Layouts
// library: status plumbing, the per-device context (LDS budget, CU count,
// grow-only scratch), and the rank -> slot closed forms used by the in-LDS
// layout kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tree_layouts.h"

namespace tl {

// ---------------------------------------------------------------------------
// status plumbing

void set_error(const std::string &msg);

#define TL_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::tl::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return TL_EHIP;                                                            \
        }                                                                              \
    } while (0)

#define TL_TRY(expr)                                                                   \
    do {                                                                               \
        tl_status s_ = (expr);                                                         \
        if (s_ != TL_OK) return s_;                                                    \
    } while (0)

#define TL_LAUNCH_CHECK()                                                              \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) {                                                        \
            ::tl::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));   \
            return TL_EHIP;                                                            \
        }                                                                              \
    } while (0)

// ---------------------------------------------------------------------------
// per-device context

struct Ctx {
    int device = -1;
    int num_cus = 256;
    size_t lds_max = 65536;   // dynamic LDS a workgroup may request
    // grow-only scratch areas (device memory), one per purpose
    void *scratch[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t scratch_bytes[6] = {0, 0, 0, 0, 0, 0};
    uint32_t *err_flag = nullptr;   // device word set by bounded waits on timeout
    uint64_t wall_khz = 100000;     // rate of the device wall clock
    uint64_t wait_ticks = 0;        // bound on one in-kernel wait, wall-clock ticks
    // this device's copy of the static scratch array of the allocation-free
    // entry point (TL_NOALLOC_SCRATCH_BYTES), as a device address
    unsigned char *noalloc = nullptr;
};

// split of the allocation-free scratch: plan tables, then move flags (every
// chained op its own region), then the stash
constexpr size_t NOALLOC_BLOB_BYTES = 8u << 20;
constexpr size_t NOALLOC_FLAG_BYTES = 4u << 20;
constexpr size_t NOALLOC_STASH_BYTES =
    (size_t)TL_NOALLOC_SCRATCH_BYTES - NOALLOC_BLOB_BYTES - NOALLOC_FLAG_BYTES;

tl_status get_ctx(Ctx **out);
// slot 0: stash of displaced keys, 1: move flags / counters, 2: plan tables,
// 3: miscellaneous, 4: the sort's second key buffer, 5: the sort's counters
tl_status scratch(Ctx *c, int slot, size_t bytes, void **out);
tl_status copy_to_device(Ctx *c, int slot, const void *host, size_t bytes, hipStream_t s,
                         void **dev_out);
tl_status check_err_flag(Ctx *c, hipStream_t s);

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s = nullptr;
    float *ms = nullptr;
    tl_status start(hipStream_t s_, float *ms_);
    tl_status stop();
    ~Timer();
};

// ---------------------------------------------------------------------------
// integer helpers (host and device)

__host__ __device__ inline uint32_t ilog_base(uint64_t n, uint64_t base) {
    uint32_t h = 0;
    uint64_t p = 1;
    while (p <= n / base) {
        p *= base;
        h++;
    }
    return h;
}

__host__ __device__ inline uint64_t ipow64(uint64_t base, uint32_t e) {
    uint64_t r = 1;
    while (e--) r *= base;
    return r;
}

__host__ __device__ inline uint32_t floor_log2_64(uint64_t n) {
#ifdef __HIP_DEVICE_COMPILE__
    return 63u - (uint32_t)__clzll((long long)n);
#else
    return 63u - (uint32_t)__builtin_clzll(n);
#endif
}

// Split of a depth-d tree into a top tree of depth root_d and bottom trees of
// depth leaf_d, leaf_d rounded up to a power of two
// (gpu/src/mveb-cycles.cuh:240-247).
// classic split of the CPU vEB (cpu/src/veb-cycles.h:114-115,
// cpu/src/query-veb.cpp:28-29): top tree floor(d/2) levels, bottom trees the
// rest; `classic` = false is the modified vEB's split below
__host__ __device__ inline void mveb_split(uint32_t d, uint32_t &rd, uint32_t &ld);
__host__ __device__ inline void veb_split(uint32_t d, uint32_t &rd, uint32_t &ld, bool classic) {
    if (classic) {
        rd = (d - 2) / 2 + 1;
        ld = d - rd;
    } else {
        mveb_split(d, rd, ld);
    }
}

__host__ __device__ inline void mveb_split(uint32_t d, uint32_t &rd, uint32_t &ld) {
    uint32_t r = (d - 2) / 2 + 1;
    uint32_t l = d - r;
    if (l & (l - 1)) {
        uint32_t p = 1;
        while (p < l) p <<= 1;
        l = p;
        r = d - l;
    }
    rd = r;
    ld = l;
}

// ---------------------------------------------------------------------------
// rank -> slot closed forms on 32-bit sizes (used for n <= LDS capacity)

// Constants of a level-order B-tree of n keys (b keys per node) reduced to its
// perfect internal part: the first `span` ranks alternate b leaf keys with a
// separator, the next `rem` ranks fill the partial leaf, the rest are the
// remaining internal keys.  Mirrors cpu/src/btree-cycles.h:29-36,71-91.
struct BtreeShape {
    uint32_t n, b, k1;
    uint32_t perfect;
    uint32_t nint, span, rem, nsep;
};

__host__ __device__ inline BtreeShape btree_shape(uint32_t n, uint32_t b) {
    BtreeShape s;
    s.n = n;
    s.b = b;
    s.k1 = b + 1;
    uint32_t h = ilog_base(n, s.k1);
    uint64_t P = ipow64(s.k1, h);
    s.perfect = ((uint64_t)n == P * s.k1 - 1);
    if (s.perfect) {
        s.nint = n;
        s.span = s.rem = s.nsep = 0;
    } else {
        uint32_t nint = (uint32_t)(P - 1), nleaf = n - nint;
        uint32_t rem = nleaf % b, l = nleaf - rem, i = l / b;
        s.nint = nint;
        s.span = l + i;
        s.rem = rem;
        s.nsep = i;
    }
    return s;
}

// slot of rank x; for a perfect tree of (b+1)^d - 1 keys the 1-based rank
// (u, a, 0^s) in base b+1 sits on level d-1-s at (b+1)^(d-1-s) - 1 + u*b + a - 1
__host__ __device__ inline uint32_t btree_pos32(uint32_t x, const BtreeShape &s) {
    uint32_t n = s.n;
    if (!s.perfect) {
        if (x < s.span) {
            uint32_t row = x / s.k1, c = x - row * s.k1;
            if (c < s.b) return s.nint + row * s.b + c;
            x = row;
        } else if (x < s.span + s.rem) {
            return s.nint + (s.span - s.nsep) + (x - s.span);
        } else {
            x = s.nsep + (x - s.span - s.rem);
        }
        n = s.nint;
    }
    uint32_t r = x + 1, div = 1;
    uint32_t q = r / s.k1;
    while (r == q * s.k1) {
        r = q;
        div *= s.k1;
        q = r / s.k1;
    }
    uint32_t a = r - q * s.k1;
    uint32_t level_start = (uint32_t)(((uint64_t)n + 1) / ((uint64_t)div * s.k1)) - 1;
    return level_start + q * s.b + a - 1;
}

// slot of rank x in the modified vEB of n keys (structural recursion of
// gpu/src/mveb-cycles.cuh:230-414)
__host__ __device__ inline uint32_t mveb_pos32(uint32_t x, uint32_t n, bool classic = false) {
    uint32_t base = 0;
    for (;;) {
        if (n <= 1) return base;
        uint32_t d = floor_log2_64(n) + 1, rd, ld;
        veb_split(d, rd, ld, classic);
        uint32_t r = (1u << rd) - 1, l = (1u << ld) - 1;
        uint64_t full = (1ull << d) - 1;
        uint32_t nf, inc, span;
        if ((uint64_t)n == full) {
            span = n;
            nf = r + 1;
            inc = 0;
        } else {
            nf = (n - r) / l;
            inc = n - r - nf * l;
            span = nf << ld;   // nf * (l + 1)
        }
        if (x < span) {
            uint32_t k = x >> ld, c = x & l;
            if (c == l) {
                n = r;
                x = k;
                continue;
            }
            base += r + k * l;
            n = l;
            x = c;
            continue;
        }
        if (x < span + inc) {
            base += r + nf * l;
            x -= span;
            n = inc;
            continue;
        }
        x = nf + (x - span - inc);
        n = r;
    }
}

// slot of rank x for one equidistant gather step ("G"): `rows` rows of b
// leaf keys followed by one separator become [separators][leaf keys]
__host__ __device__ inline uint32_t gather_pos32(uint32_t x, uint32_t rows, uint32_t b) {
    uint32_t k1 = b + 1;
    uint32_t row = x / k1, c = x - row * k1;
    if (row >= rows) return x;   // trailing keys stay
    return (c == b) ? row : rows + row * b + c;
}

// in-LDS layout kinds (3 is the BST top levels' per-level bit reversal)
enum LdsKind : uint32_t { LDS_BTREE = 0, LDS_MVEB = 1, LDS_GATHER = 2, LDS_VEB = 4 };

}   // namespace tl
