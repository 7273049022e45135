// tl_queue.hpp -- the uneven batch's query queue: one persistent launch that
// answers the queries of many arrays (tl_query.hip), driven by tl_batch.hip.
#pragma once

#include "tl_internal.hpp"

namespace tl {

// one array's query batch as the queue sees it (device addresses)
struct QueueItem {
    uint64_t A, Q, ans, n, nq;
    uint64_t tables, idx, tidx;     // mvEB table set
    uint32_t dtype, layout, b, G;   // G: lanes per node of the wide B-tree descent
    uint32_t ntables, table_size, ntop, pad;
};

// queries of one array per ticket: $TL_QQ_CHUNK (2^10 .. 2^20), else 4096
// (the in-flight chunks of all workgroups are the queue's working set)
uint64_t queue_chunk();

// whether the queue answers this array's queries (B-tree arrays need the wide
// descent with 32-bit node numbers and at most 16 lanes per node; the others
// are left to the one-array kernels)
bool queue_takes(const tl_batch_item &it);
// the array's queue record and the LDS bytes its chunks need
QueueItem queue_item(const tl_batch_item &it, uint64_t tables, uint64_t idx, uint64_t tidx,
                     uint32_t ntables, uint32_t table_size, size_t *lds);
// the persistent launch over nitems records of one layout whose chunk counts
// (of `chunk` queries) have the prefix sums pre[0..nitems) (chunks in all);
// *ticket must be zero.  Workgroups per compute unit: $TL_QQ_GRID (1..64),
// else 16.
tl_status queue_launch(Ctx *c, uint32_t layout, const QueueItem *items, const uint64_t *pre,
                       uint32_t nitems, uint64_t chunks, uint64_t chunk,
                       unsigned long long *ticket, size_t lds, hipStream_t s);

}   // namespace tl
