// The pieces of the library's other translation units that the emulated
// tl_batch.hip links against: a host context, grow-only host scratch, and
// stubs that abort for the entry points the persistent launch never calls.
#include <stdio.h>
#include "src/tl_internal.hpp"
#include "src/tl_plan.hpp"

namespace tl {
static Ctx g;
static uint32_t g_err[64];
void set_error(const std::string &m) { fprintf(stderr, "[emu] set_error: %s\n", m.c_str()); }
tl_status get_ctx(Ctx **out) {
    if (g.device < 0) {
        g.device = 0;
        g.num_cus = 256;
        g.lds_max = 163840;
        g.wall_khz = 100000;
        g.wait_ticks = 100000ull * 2000;   // 2 s
        g.err_flag = g_err;
    }
    *out = &g;
    return TL_OK;
}
tl_status scratch(Ctx *c, int slot, size_t bytes, void **out) {
    if (c->scratch_bytes[slot] < bytes) {
        free(c->scratch[slot]);
        c->scratch[slot] = malloc(bytes);
        c->scratch_bytes[slot] = bytes;
    }
    *out = c->scratch[slot];
    return TL_OK;
}
tl_status copy_to_device(Ctx *c, int slot, const void *host, size_t bytes, hipStream_t, void **dev_out) {
    TL_TRY(scratch(c, slot, bytes, dev_out));
    memcpy(*dev_out, host, bytes);
    return TL_OK;
}
tl_status check_err_flag(Ctx *c, hipStream_t) {
    if (*c->err_flag) { *c->err_flag = 0; return TL_ETIMEOUT; }
    return TL_OK;
}
tl_status Timer::start(hipStream_t, float *m) { ms = m; return TL_OK; }
tl_status Timer::stop() { if (ms) *ms = 0; return TL_OK; }
Timer::~Timer() {}
#ifndef EMU_WITH_PERMUTE   // the emulated tl_permute.hip defines the launches itself
tl_status fill_u32_launch(uint32_t *p, uint64_t count, uint32_t v, hipStream_t) {
    for (uint64_t i = 0; i < count; i++) p[i] = v;
    return TL_OK;
}
tl_status plan_launch(Ctx *, const Plan &, void *, uint32_t, hipStream_t, const uint64_t *, void *, uint32_t *) {
    fprintf(stderr, "[emu] plan_launch not expected\n"); abort();
}
tl_status lds_items_launch(const uint64_t *, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, hipStream_t) {
    fprintf(stderr, "[emu] lds_items_launch not expected\n"); abort();
}
tl_status multi_launch(Ctx *, const MultiLaunch &, hipStream_t) {
    fprintf(stderr, "[emu] multi_launch not expected\n"); abort();
}
#endif
}   // namespace tl
#define NOTEXP(name) { fprintf(stderr, "[emu] " #name " not expected\n"); abort(); }
extern "C" {
const char *tl_last_error(void) { return "see [emu] set_error lines"; }
tl_status tl_query_bst(const void *, uint64_t, tl_dtype, const void *, uint64_t, uint64_t *, void *, float *) NOTEXP(tl_query_bst)
tl_status tl_query_btree(const void *, uint64_t, uint64_t, tl_dtype, const void *, uint64_t, uint64_t *, void *, float *) NOTEXP(tl_query_btree)
tl_status tl_query_mveb_nonperfect(const void *, const tl_veb_table *, const uint64_t *, const uint32_t *, uint64_t, uint32_t, uint32_t, tl_dtype, const void *, uint64_t, uint64_t *, void *, float *) NOTEXP(tl_query_mveb_nonperfect)
#ifndef EMU_WITH_PERMUTE
tl_status tl_permute_bst(void *, uint64_t, tl_dtype, void *, float *) NOTEXP(tl_permute_bst)
tl_status tl_permute_btree(void *, uint64_t, uint64_t, tl_dtype, void *, float *) NOTEXP(tl_permute_btree)
tl_status tl_permute_mveb(void *, uint64_t, tl_dtype, void *, float *) NOTEXP(tl_permute_mveb)
#endif
}

#include "src/tl_queue.hpp"
namespace tl {
// the emulated persistent launch never uses the query queues
bool queue_takes(const tl_batch_item &) { return false; }
QueueItem queue_item(const tl_batch_item &, uint64_t, uint64_t, uint64_t, uint32_t, uint32_t,
                     size_t *lds) {
    *lds = 0;
    return QueueItem{};
}
uint64_t queue_chunk() { return 4096; }
tl_status queue_launch(Ctx *, uint32_t, const QueueItem *, const uint64_t *, uint32_t, uint64_t,
                       uint64_t, unsigned long long *, size_t, hipStream_t) {
    fprintf(stderr, "[emu] queue_launch not expected\n");
    abort();
}
int &veb_qpl_fixed() {
    static thread_local int v = 0;
    return v;
}
}   // namespace tl
