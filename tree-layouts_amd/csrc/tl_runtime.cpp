// tl_runtime.cpp -- status string, per-device context, grow-only scratch and
// event timing for the gfx950 Tree-Layouts library.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <mutex>

#include "tl_internal.hpp"

namespace tl {

static thread_local std::string g_last_error;
static std::mutex g_ctx_mu;
static Ctx g_ctx[64];

// Device memory the library holds without allocating, one copy per device
// (zero at load): the error word of the bounded in-kernel waits, and the
// scratch of the allocation-free permutation.  Host code uses these only
// through the addresses hipGetSymbolAddress returns for the current device
// (Ctx::err_flag, Ctx::noalloc); the names themselves are host stand-ins.
static __device__ __align__(256) uint32_t err_words[64];
static __device__ __align__(256) unsigned char noalloc_scratch[TL_NOALLOC_SCRATCH_BYTES];

void set_error(const std::string &msg) { g_last_error = msg; }

tl_status get_ctx(Ctx **out) {
    int dev = 0;
    TL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index out of range");
        return TL_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx[dev];
    if (c.device < 0) {
        int cus = 0, lds = 0;
        TL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        TL_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        c.num_cus = cus > 0 ? cus : 256;
        c.lds_max = lds > 0 ? (size_t)lds : 65536;
        int khz = 0;
        TL_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        c.wall_khz = khz > 0 ? (uint64_t)khz : 100000;
        const char *w = getenv("TL_WAIT_MS");
        const uint64_t wait_ms = w ? std::max<uint64_t>(1, strtoull(w, nullptr, 10)) : 4000;
        c.wait_ticks = wait_ms * c.wall_khz;
        // set-up allocates nothing: the device's own copies of the static
        // arrays, looked up once with this device current
        void *p = nullptr;
        TL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(err_words)));
        c.err_flag = (uint32_t *)p;
        void *q = nullptr;
        TL_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(noalloc_scratch)));
        c.noalloc = (unsigned char *)q;
        c.device = dev;
    }
    *out = &c;
    return TL_OK;
}

tl_status scratch(Ctx *c, int slot, size_t bytes, void **out) {
    if (bytes == 0) bytes = 256;
    if (c->scratch_bytes[slot] < bytes) {
        if (c->scratch[slot]) {
            // the old buffer may still be read by enqueued work
            TL_HIP(hipDeviceSynchronize());
            TL_HIP(hipFree(c->scratch[slot]));
            c->scratch[slot] = nullptr;
            c->scratch_bytes[slot] = 0;
        }
        size_t want = bytes + bytes / 4;
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipMalloc of " + std::to_string(want) + " scratch bytes failed");
            return TL_ENOMEM;
        }
        c->scratch[slot] = p;
        c->scratch_bytes[slot] = want;
    }
    *out = c->scratch[slot];
    return TL_OK;
}

tl_status copy_to_device(Ctx *c, int slot, const void *host, size_t bytes, hipStream_t s,
                         void **dev_out) {
    TL_TRY(scratch(c, slot, bytes, dev_out));
    if (bytes) {
        // synchronous w.r.t. the host buffer: the plan vectors die after the call
        TL_HIP(hipMemcpyAsync(*dev_out, host, bytes, hipMemcpyHostToDevice, s));
        TL_HIP(hipStreamSynchronize(s));
    }
    return TL_OK;
}

tl_status check_err_flag(Ctx *c, hipStream_t s) {
    uint32_t h = 0;
    TL_HIP(hipMemcpyAsync(&h, c->err_flag, sizeof(h), hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    if (h) {
        TL_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(uint32_t), s));
        TL_HIP(hipStreamSynchronize(s));
        set_error("an in-kernel bounded wait timed out");
        return TL_ETIMEOUT;
    }
    return TL_OK;
}

tl_status Timer::start(hipStream_t s_, float *ms_) {
    s = s_;
    ms = ms_;
    if (!ms) return TL_OK;
    TL_HIP(hipEventCreate(&a));
    TL_HIP(hipEventCreate(&b));
    TL_HIP(hipEventRecord(a, s));
    return TL_OK;
}

tl_status Timer::stop() {
    if (!ms) return TL_OK;
    TL_HIP(hipEventRecord(b, s));
    TL_HIP(hipEventSynchronize(b));
    float t = 0.f;
    TL_HIP(hipEventElapsedTime(&t, a, b));
    *ms = t;
    return TL_OK;
}

Timer::~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
}

}   // namespace tl

extern "C" const char *tl_last_error(void) { return tl::g_last_error.c_str(); }

// Waits for `stream` and reports (and clears) a give-up of any in-kernel
// bounded wait since the last report: the way an untimed call (ms == NULL),
// or a captured graph of such calls, learns of TL_ETIMEOUT.
extern "C" tl_status tl_sync(void *stream) {
    tl::Ctx *c;
    TL_TRY(tl::get_ctx(&c));
    return tl::check_err_flag(c, (hipStream_t)stream);
}
extern "C" const char *tl_version(void) { return "tree-layouts-mi355x 0.1 (gfx950)"; }
