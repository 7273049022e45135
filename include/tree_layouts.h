/*
 * tree_layouts.h -- C ABI of the MI355X (gfx950) Tree-Layouts library.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are
 * hipMalloc'ed by the caller; `stream` is a hipStream_t passed as void*
 * (NULL = the default stream).  When `ms` is non-NULL the call waits for its
 * work and stores the elapsed GPU time in milliseconds (the reference's
 * functions are synchronous and return a time); when `ms` is NULL the call
 * only enqueues work on `stream`.
 *
 * Permutations that use chained moves wait inside their kernels, and every
 * such wait is bounded in time.  A wait that gives up sets a per-device error
 * word; a timed call reports it as TL_ETIMEOUT when it returns, an untimed
 * call leaves it for tl_sync().  The word is sticky: it holds any give-up
 * since the last report, and reading it clears it.
 *
 * Scratch (stash, move flags, plan tables, error word) is per device and
 * owned by the library.  Calls that overlap on one device -- on two streams,
 * or from two host threads -- share it and race: serialise them.
 *
 * Query answers are always uint64: the slot of the query key in the layout,
 * or n when the key is absent (the reference's own miss value,
 * cpu/src/query-bst.h:42, cpu/src/query-btree.h:53).
 */
#ifndef TREE_LAYOUTS_H
#define TREE_LAYOUTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TL_OK = 0,
    TL_EINVAL = 1,      /* bad size / width / pointer combination */
    TL_ENOMEM = 2,      /* device scratch could not be allocated */
    TL_EHIP = 3,        /* a HIP call failed; see tl_last_error() */
    TL_ETIMEOUT = 4,    /* a bounded in-kernel wait gave up (never expected) */
    TL_EUNSUPPORTED = 5, /* combination this build does not implement */
    TL_EIO = 6           /* a layout file could not be read or written, or is corrupt */
} tl_status;

typedef enum { TL_U32 = 0, TL_U64 = 1, TL_I32 = 2, TL_I64 = 3 } tl_dtype;

/* Layout-compatible with the reference's struct vEB_table
 * (gpu/src/query-mveb.cuh:27-31): {u64 L; u64 R; u32 D;}, 24-byte stride. */
typedef struct {
    uint64_t L;
    uint64_t R;
    uint32_t D;
} tl_veb_table;

const char *tl_last_error(void);
const char *tl_version(void);

/* Waits for `stream`, then returns TL_ETIMEOUT (and clears the error word)
 * if an in-kernel bounded wait gave up since the last report, else TL_OK.
 * No counterpart in the reference, which has no error path. */
tl_status tl_sync(void *stream);

/* Diagnostic: runs one workgroup whose bounded wait can never be satisfied,
 * bounded at wait_ms (1..60000), and returns what tl_sync returns afterwards
 * -- TL_ETIMEOUT when the bound works.  *ms (optional) receives the elapsed
 * GPU time.  Lets a test check the wall-clock bound on the device; no
 * counterpart in the reference. */
tl_status tl_selftest_timeout(uint32_t wait_ms, void *stream, float *ms);

/* ---- in-place permutations of a sorted array -------------------------- */

/* Level-order BST (Eytzinger).  Replaces
 * timePermuteBST<T>(T *dev_A, uint64_t n)  (gpu/src/bst-hybrid.cuh:120). */
tl_status tl_permute_bst(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);

/* Level-order B-tree with b keys per node.  Replaces
 * timePermuteBtree<T>(T *dev_A, uint64_t n, uint64_t b)
 * (gpu/src/btree-cycles.cuh:77). */
tl_status tl_permute_btree(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype, void *stream,
                           float *ms);

/* Modified van Emde Boas layout.  Replaces
 * timePermutevEB<T>(T *dev_A, uint64_t n)  (gpu/src/mveb-cycles.cuh:420). */
tl_status tl_permute_mveb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);

/* Classic van Emde Boas layout, the reference's CPU-only layout (top tree of
 * floor(d/2) levels, no power-of-two rounding).  Replaces
 * timePermutevEB<T>(T *A, uint64_t n, uint32_t p)  (cpu/src/veb-cycles.h:381). */
tl_status tl_permute_veb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);

/* Allocation-free tl_permute_btree: the same layout, enqueued on `stream`,
 * with every piece of scratch (plan tables, move flags, stash) taken from a
 * fixed static device array of TL_NOALLOC_SCRATCH_BYTES per device instead
 * of library allocations.  No counterpart in the reference; it is for
 * callers that cannot allocate (code around graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the call makes no hipMalloc,
 * hipFree or other allocating call.  Its plan keeps every stash within the
 * array, so large n may take slower passes than tl_permute_btree.  Returns
 * TL_EINVAL, before any launch, for n > TL_NOALLOC_MAX_KEYS or a plan whose
 * scratch does not fit the array.  Every call on a device shares that
 * device's array: calls that overlap on one device (two streams, two host
 * threads) race and must be serialised.  A bounded wait that gives up is
 * reported by the next tl_sync(). */
#define TL_NOALLOC_SCRATCH_BYTES (64u << 20)
#define TL_NOALLOC_MAX_KEYS (1ull << 31)
tl_status tl_permute_btree_noalloc(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                   void *stream);

/* ---- the involution family ---------------------------------------------- */

/* The reference's second set of permutation programs, in which every pass
 * swaps each key with a closed-form partner (an involution).  Same layouts as
 * tl_permute_{bst,btree,mveb}; more passes and scattered partners, so slower
 * (the gather / cycle-leader path above is the fast one).  n < 2^40.  No
 * n-sized scratch: the library's context holds the pass list and the partner
 * tables of pass shapes that repeat over partitions or run in LDS (at most
 * 2^17 four-byte entries per shape).  Replace
 *   timePermuteBST<T>(dev_A, n)       (gpu/src/bst-involutions.cuh:167),
 *   timePermuteBtree<T>(dev_A, n, b)  (gpu/src/btree-involutions.cuh:151),
 *   timePermutevEB<T>(dev_A, n)       (gpu/src/mveb-involutions.cuh:503).
 * The BST's perfect part (phaseOne, phaseTwo) runs as tl_permute_bst's tiled
 * bit reversals. */
tl_status tl_permute_bst_involutions(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                     float *ms);
tl_status tl_permute_btree_involutions(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                       void *stream, float *ms);
tl_status tl_permute_mveb_involutions(void *dev_A, uint64_t n, tl_dtype dtype, void *stream,
                                      float *ms);

/* Their primitives, each two involution passes over dev_A[0..n):
 * k-way shuffle of n = d*k keys, key i*d + j -> j*k + i
 *   (shuffle_dk_phaseOne/Two, gpu/src/involutions.cuh:84,106);
 * its inverse, every k-th key gathered: key j*k + i -> i*d + j
 *   (unshuffle_dk, cpu/src/involutions.h:161);
 * k-way unshuffle of n = k^d - 1 keys: a stable partition by (i+1) mod k,
 *   residue 0 first, then 1, ..., k-1 (unshuffle_phaseOne/Two,
 *   gpu/src/involutions.cuh:38,61);
 * rotation right by k <= n by two reversals (shift_right_phaseOne/Two,
 *   gpu/src/common.cuh:62-110). */
tl_status tl_shuffle_dk(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype, void *stream,
                        float *ms);
tl_status tl_unshuffle_dk(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype, void *stream,
                          float *ms);
tl_status tl_unshuffle(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype, void *stream,
                       float *ms);
tl_status tl_shift_right(void *dev_A, uint64_t n, uint64_t k, tl_dtype dtype, void *stream,
                         float *ms);

/* ---- inverse permutations: layout -> sorted order, in place ------------- */

/* The exact inverses of tl_permute_{bst,btree,mveb} for the same (n, b): the
 * key at each slot of the layout goes back to its rank, whatever the values.
 * No counterpart in the reference, whose README (README.md:2) names the
 * permutations "between sorted order and" the layouts but implements only
 * the forward direction.  Same scratch bound as the forward call. */
tl_status tl_unpermute_bst(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);
tl_status tl_unpermute_btree(void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype, void *stream,
                             float *ms);
tl_status tl_unpermute_mveb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);
tl_status tl_unpermute_veb(void *dev_A, uint64_t n, tl_dtype dtype, void *stream, float *ms);

/* ---- batched queries --------------------------------------------------- */

/* Replaces searchAll / timeQueryBST (gpu/src/query-bst.cuh:47,57). */
tl_status tl_query_bst(const void *dev_A, uint64_t n, tl_dtype dtype, const void *dev_q,
                       uint64_t nq, uint64_t *dev_ans, void *stream, float *ms);

/* Replaces searchAll / timeQueryBtree (gpu/src/query-btree.cuh:64,76); any
 * b >= 1 (the reference kernel assumes b = 32). */
tl_status tl_query_btree(const void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                         const void *dev_q, uint64_t nq, uint64_t *dev_ans, void *stream,
                         float *ms);

/* Replaces searchAll / timeQueryvEB for a perfect mvEB
 * (gpu/src/query-mveb.cuh:73,86). dev_table holds d entries. */
tl_status tl_query_mveb(const void *dev_A, const tl_veb_table *dev_table, uint64_t n, uint32_t d,
                        tl_dtype dtype, const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                        void *stream, float *ms);

/* Replaces searchAll_nonperfect / timeQueryvEB_nonperfect
 * (gpu/src/query-mveb.cuh:218,231). */
tl_status tl_query_mveb_nonperfect(const void *dev_A, const tl_veb_table *dev_tables,
                                   const uint64_t *dev_idx, const uint32_t *dev_table_idx,
                                   uint64_t n, uint32_t d, uint32_t num_tables, tl_dtype dtype,
                                   const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                                   void *stream, float *ms);

/* Binary search of the sorted dev_A[0..n), the layouts' comparator.  Replaces
 * searchAll / timeQuery of gpu/src/binary-search.cuh:52,63 (CPU twin
 * cpu/src/binary-search.h:55,64).  For distinct keys dev_ans[i] is the index
 * of q[i], or of its predecessor when q[i] is absent, or n when every key is
 * larger than q[i] (there the reference's loop does not end). */
tl_status tl_query_binary(const void *dev_A, uint64_t n, tl_dtype dtype, const void *dev_q,
                          uint64_t nq, uint64_t *dev_ans, void *stream, float *ms);

/* Classic vEB queries (cpu/src/query-veb.h:31-52, 103-137): the same search
 * with tables of the classic rule (tl_veb_classic_build_table,
 * tl_veb_classic_tableset_device).  Replace searchAll / timeQueryvEB and
 * searchAll_nonperfect / timeQueryvEB_nonperfect of cpu/src/query-veb.h:57,71,
 * 141,155; the table set is packed as for the mvEB instead of the
 * reference's array of per-structure pointers (cpu/src/query-veb.h:128). */
tl_status tl_query_veb(const void *dev_A, const tl_veb_table *dev_table, uint64_t n, uint32_t d,
                       tl_dtype dtype, const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                       void *stream, float *ms);
tl_status tl_query_veb_nonperfect(const void *dev_A, const tl_veb_table *dev_tables,
                                  const uint64_t *dev_idx, const uint32_t *dev_table_idx,
                                  uint64_t n, uint32_t d, uint32_t num_tables, tl_dtype dtype,
                                  const void *dev_q, uint64_t nq, uint64_t *dev_ans,
                                  void *stream, float *ms);

/* ---- mvEB query tables -------------------------------------------------- */

/* Host table build.  Replaces buildTable(vEB_table*, n, d, root_depth)
 * (gpu/src/query-mveb.cu:25); table must hold d entries, zeroed. */
tl_status tl_veb_build_table(tl_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth);

/* Device build of the whole table set (the table-set walk of
 * gpu/test/query-mveb-fixedN.cu:42-165 plus one buildTable per table), done
 * by a kernel.  Outputs live in device memory:
 *   dev_tables    : >= 64*64 entries,  dev_idx, dev_table_idx : >= 64 entries,
 *   dev_meta[0..2]: d, num_tables, table_size.                               */
tl_status tl_veb_tableset_device(uint64_t n, tl_veb_table *dev_tables, uint64_t *dev_idx,
                                 uint32_t *dev_table_idx, uint32_t *dev_meta, void *stream);

/* The same two for the classic vEB: buildTable of cpu/src/query-veb.cpp:25
 * and the table-set walk of cpu/test/query-veb-fixedN.cpp:37-158. */
tl_status tl_veb_classic_build_table(tl_veb_table *table, uint64_t n, uint32_t d,
                                     uint32_t root_depth);
tl_status tl_veb_classic_tableset_device(uint64_t n, tl_veb_table *dev_tables, uint64_t *dev_idx,
                                         uint32_t *dev_table_idx, uint32_t *dev_meta,
                                         void *stream);

/* ---- learned start node (no counterpart in the reference) -------------- */

/* Bytes of a model with num_leaves leaf models (64-byte header + 32 per leaf). */
uint64_t tl_learned_model_bytes(uint32_t num_leaves);

/* Fits the two-layer rank model (root linear map -> num_leaves linear leaf
 * maps, FP32 inference, per-leaf max rank error) on the GPU from the sorted
 * keys dev_sorted[0..n) into dev_model (tl_learned_model_bytes bytes of
 * device memory).  The largest leaf error is the uint32 at byte offset 32 of
 * the model.  Addition on top of the reference; its queries answer exactly
 * like tl_query_btree. */
tl_status tl_learned_fit(const void *dev_sorted, uint64_t n, tl_dtype dtype, uint32_t num_leaves,
                         void *dev_model, void *stream, float *ms);

/* Per key of dev_keys[0..n): the model's predicted rank (dev_pred) and the
 * error bound of the leaf the key maps to (dev_err), both int64, computed
 * with the query path's arithmetic; -1 in both for a key whose leaf is empty.
 * For the sorted keys the model was fitted on, |pred - rank| <= err holds for
 * every key.  Lets a caller (or a test) check the bound; asynchronous. */
tl_status tl_learned_predict(const void *dev_model, const void *dev_keys, uint64_t n,
                             tl_dtype dtype, int64_t *dev_pred, int64_t *dev_err, void *stream);

/* tl_query_btree with the descent started at the deepest node whose subtree
 * holds the model's rank window; queries the window does not settle are
 * answered by the plain root descent.  dev_stats (optional, one uint64,
 * accumulated) counts those fallbacks.  2 <= b <= 64. */
tl_status tl_query_btree_learned(const void *dev_A, uint64_t n, uint64_t b, tl_dtype dtype,
                                 const void *dev_model, const void *dev_q, uint64_t nq,
                                 uint64_t *dev_ans, uint64_t *dev_stats, void *stream, float *ms);

/* ---- before the layouts: raw keys -> sorted distinct keys --------------- */

/* Sorts dev_keys[0..n) ascending in the key type's order (signed for TL_I32 /
 * TL_I64) and drops repeated keys; the distinct keys then fill
 * dev_keys[0..*n_out) (host).  Waits for its work.  n < 2^32; scratch is one
 * library-owned buffer of n keys plus counters.  No counterpart in the
 * reference: its drivers build sorted distinct input themselves
 * (cpu/src/common.h:33-40) and every layout assumes it. */
tl_status tl_sort_dedupe(void *dev_keys, uint64_t n, tl_dtype dtype, uint64_t *n_out,
                         void *stream, float *ms);

/* ---- layout files (no counterpart: the reference has no serialisation) --- */

/* layout codes of a file: 0 BST, 1 B-tree, 2 mvEB, 3 classic vEB, or sorted */
#define TL_FILE_SORTED 255u

typedef struct {
    uint32_t layout;      /* 0..3 or TL_FILE_SORTED */
    tl_dtype dtype;
    uint64_t n;
    uint64_t b;           /* keys per node (B-tree), 1 for a BST, else 0 */
    uint32_t d;           /* levels of the binary tree (vEB layouts), floor(log2 n)+1 otherwise */
    uint32_t num_tables;  /* vEB query table set: structures ... */
    uint32_t table_size;  /* ... and table entries in all; 0 for the other layouts */
    uint32_t version;
    uint64_t checksum;    /* position-sensitive checksum of the keys */
} tl_layout_info;

/* Writes the n keys of dev_A (laid out as `layout`, with node width b for a
 * B-tree) to `path`: a 256-byte header {magic "TLLAYOUT", version, layout,
 * dtype, n, b, d, table-set shape, offsets, checksum}, the vEB query table
 * set for layouts 2 and 3 (tl_veb_table entries, idx, table_idx), and the raw
 * keys at a 4096-aligned offset.  Waits for its work. */
tl_status tl_layout_save(const char *path, const void *dev_A, uint64_t n, uint32_t layout,
                         uint64_t b, tl_dtype dtype, void *stream);
/* Reads and checks the header only (to size the buffers of a load). */
tl_status tl_layout_read_info(const char *path, tl_layout_info *info);
/* Loads the keys into dev_A (capacity keys, >= n) and, when the pointers are
 * non-NULL, the table set into dev_tables (>= table_size entries), dev_idx
 * and dev_table_idx (>= num_tables entries): ready for tl_query_mveb_nonperfect
 * / tl_query_veb_nonperfect.  Recomputes the checksum on the device and
 * returns TL_EIO on a mismatch.  Waits for its work. */
tl_status tl_layout_load(const char *path, void *dev_A, uint64_t capacity,
                         tl_veb_table *dev_tables, uint64_t *dev_idx, uint32_t *dev_table_idx,
                         tl_layout_info *info, void *stream);

/* ---- uneven batch (no counterpart in the reference) --------------------- */

/* One independent array of a batch: n sorted keys at A (device) to be laid
 * out in place as a BST (layout 0), a B-tree with b keys per node (layout 1,
 * 1 <= b <= 64) or an mvEB (layout 2), then nq queries Q answered into ans
 * (device).  n < 2^32. */
typedef struct {
    void *A;
    uint64_t n;
    tl_dtype dtype;
    uint32_t layout;
    uint64_t b;
    const void *Q;
    uint64_t nq;
    uint64_t *ans;
} tl_batch_item;

/* lds_bytes: LDS budget of the in-LDS sub-layouts (0 = the device's whole
 * budget, 160 KiB less a 2 KiB reserve on gfx950);
 * wg_per_cu, wait_ms: workgroups per compute unit and the bound on one
 * in-kernel wait of the single persistent launch, which is switched off in
 * this build (tl_batch_run runs the planned path below); both are ignored. */
typedef struct {
    uint32_t lds_bytes;
    uint32_t wg_per_cu;
    uint32_t wait_ms;
    uint32_t tile_k;          /* bit-reversal tiles of 2^k x 2^k keys: 5 or 6 (0 = 5) */
    uint32_t stream_threads;  /* threads per streamed-gather partition: 256, 512 or
                                 1024 (0 = $TL_STREAM_THREADS, else 1024) */
    uint32_t btree_qpg;       /* B-tree descents per lane group: 2 or 4 (0 = 4) */
    uint32_t veb_qpl;         /* mvEB queries per lane: 1, 2 or 4 (0 = 2) */
    uint32_t query_queue;     /* 1: every array's queries through one persistent
                                 work-queue launch per layout (workgroups take
                                 chunks of 4096 queries from an atomic ticket);
                                 0: one query launch per array */
} tl_batch_config;

/* Plans every array (the same plans as the single-array permutations) and
 * uploads the task list; cfg may be NULL. */
tl_status tl_batch_create(const tl_batch_item *items, uint32_t count, const tl_batch_config *cfg,
                          void **handle);
/* Permutes and queries every array of the batch by the planned path: the
 * arrays that fit in LDS in batched one-workgroup-per-array launches, every
 * other array's plan by waves of multi-array launches, then every array's
 * queries; use_graph != 0 captures all of it once in a HIP graph and replays
 * it.  (The single persistent work-queue launch is switched off.) */
tl_status tl_batch_run(void *handle, void *stream, int use_graph, float *ms);
/* The same work through one tl_permute_* and one tl_query_* call per array. */
tl_status tl_batch_run_separate(void *handle, void *stream, float *ms);
/* Launch plan size: number of ops, of sub-tasks and the LDS bytes per workgroup. */
tl_status tl_batch_stats(void *handle, uint64_t *ops, uint64_t *tasks, uint64_t *smem_bytes);
/* Autotuner: times configurations of the planned path on device copies of up
 * to `sample` items of the batch (restored before every run) and returns the
 * fastest in *best.  Coordinate search, one knob at a time from the defaults:
 * the LDS budget of the plans (32, 64, 96, 160 KiB), the bit-reversal tile
 * (32 or 64 keys a side), the threads per streamed-gather partition (256,
 * 512, 1024: the keys each thread moves), the B-tree descents per lane group
 * (2, 4), the mvEB queries per lane (1, 2, 4) and the query path (one launch
 * per array, or the persistent query queues).  B-tree arrays keep the b
 * they were given: b is part of the layout the caller asked for. */
tl_status tl_batch_autotune(const tl_batch_item *items, uint32_t count, uint32_t sample,
                            void *stream, tl_batch_config *best, float *best_ms);
tl_status tl_batch_destroy(void *handle);

/* ---- plan inspection (no device needed) -------------------------------- */

/* Dry run of tl_permute_{bst,btree,mveb,veb} (layout 0, 1, 2, 3; layout | 0x100
 * for tl_unpermute_*): builds and bounds-checks the launch plan for (n, b, dtype)
 * with an LDS budget of lds_bytes
 * (0 = 64 KiB) and serializes it into buf (cap words; *used receives the
 * words needed).  Word layout: [ops, blob_words], then 32 words per launch,
 * then the blob of item runs and cycle tables.  Lets a test replay the plan
 * on the host; it has no counterpart in the reference. */
tl_status tl_plan_export(uint32_t layout, uint64_t n, uint64_t b, tl_dtype dtype,
                         uint64_t lds_bytes, uint64_t *buf, uint64_t cap, uint64_t *used);

/* Dry run of tl_batch_create for a device with num_cus compute units and
 * lds_max LDS bytes per workgroup (0 = 256 and 160 KiB): builds the batch,
 * checks that every op's predecessor ends before the op's first sub-task,
 * and copies out the op records (ops_cap bytes), the sub-task list
 * (tasks_cap bytes) and the blob (blob_cap words), each only when its buffer
 * is large enough.  sizes[0..4] receive the op count, the sub-task count, the
 * blob words, the LDS bytes per workgroup and the size of one op record.
 * Lets a test replay the persistent launch on the host; no counterpart in
 * the reference. */
tl_status tl_batch_export(const tl_batch_item *items, uint32_t count, const tl_batch_config *cfg,
                          uint32_t num_cus, uint32_t lds_max, void *ops_out, uint64_t ops_cap,
                          void *tasks_out, uint64_t tasks_cap, uint64_t *blob_out,
                          uint64_t blob_cap, uint64_t *sizes);

#ifdef __cplusplus
}
#endif
#endif
