/*
 * tl_oracle.h -- CPU oracle for the Tree-Layouts hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load liboracle.so, and only as the checker or the timed CPU baseline.
 *
 * Two independent families live here:
 *   (1) a restatement, in plain C + OpenMP, of the reference's CPU algorithms
 *       (cycle-leader gathers, reversal shifts, BST involutions, queries, the
 *       involution programs and their shuffles) and of the reference's GPU
 *       mvEB recursions (cycles and involutions), which have no CPU twin;
 *   (2) constructions that do not share code with (1): the in-order DFS fill
 *       of the heap-indexed (b+1)-ary node tree (BST = b 1), the rank->slot
 *       closed forms, and the structural mvEB recursion.
 * The tests pin (1) against (2), against the compiled reference (oracle/_ref,
 * when it was built) and against committed golden fixtures.
 */
#ifndef TL_ORACLE_H
#define TL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout-compatible with the reference's struct vEB_table
 * (gpu/src/query-mveb.cuh:27-31): 24-byte stride. */
typedef struct {
    uint64_t L; /* size of each bottom tree */
    uint64_t R; /* size of the top tree (also a low-bit mask) */
    uint32_t D; /* depth of the top tree's root in the whole tree */
} or_veb_table;

/* ---- integer helpers ---------------------------------------------------- */
uint64_t or_ipow(uint64_t base, uint32_t e);
uint32_t or_ilog(uint64_t n, uint64_t base);      /* floor(log_base n), n >= 1 */
void     or_mveb_split(uint32_t d, uint32_t *root_d, uint32_t *leaf_d);
void     or_classic_split(uint32_t d, uint32_t *root_d, uint32_t *leaf_d);

/* ---- independent constructions (family 2) ------------------------------ */
/* pos_of_rank[r] = slot of the r-th smallest key, in-order DFS of the
 * complete heap-indexed (b+1)-ary node tree with ceil(n/b) nodes. */
int      or_inorder_btree(uint64_t n, uint64_t b, uint64_t *pos_of_rank);
/* closed form, O(depth) per rank: the same map as above for b >= 1 */
uint64_t or_btree_pos(uint64_t rank, uint64_t n, uint64_t b);
/* structural recursion of the modified vEB (split rule of
 * gpu/src/mveb-cycles.cuh:240-247) */
uint64_t or_mveb_pos(uint64_t rank, uint64_t n);
int      or_mveb_layout(uint64_t n, uint64_t *pos_of_rank);
/* the same recursion with the classic split (the CPU vEB layout) */
uint64_t or_classic_pos(uint64_t rank, uint64_t n);
int      or_classic_layout(uint64_t n, uint64_t *pos_of_rank);

/* ---- mvEB query tables ------------------------------------------------- */
void or_build_table(or_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth);
/* Table set of a non-perfect mvEB (gpu/test/query-mveb-fixedN.cu:42-165).
 * Returns 0 and fills the outputs; tables must hold >= 64*4 entries,
 * idx/table_idx >= 64 entries.  For a perfect n it returns num_tables = 1. */
int or_veb_tableset(uint64_t n, uint32_t *d_out, uint32_t *num_tables, uint32_t *table_size,
                    or_veb_table *tables, uint64_t *idx, uint32_t *table_idx);
/* classic vEB: table (cpu/src/query-veb.cpp:25-41) and table set */
void or_build_table_classic(or_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth);
int or_classic_tableset(uint64_t n, uint32_t *d_out, uint32_t *num_tables, uint32_t *table_size,
                        or_veb_table *tables, uint64_t *idx, uint32_t *table_idx);

/* ---- restated permutations (family 1) ---------------------------------- *
 * p = number of OpenMP threads (1 = the reference's serial path).
 * Return 0 on success, -1 on invalid arguments.                            */
int or_permute_bst_u32(uint32_t *A, uint64_t n, int p);
int or_permute_bst_u64(uint64_t *A, uint64_t n, int p);
int or_permute_btree_u32(uint32_t *A, uint64_t n, uint64_t b, int p);
int or_permute_btree_u64(uint64_t *A, uint64_t n, uint64_t b, int p);
int or_permute_mveb_u32(uint32_t *A, uint64_t n);
int or_permute_mveb_u64(uint64_t *A, uint64_t n);
int or_permute_classic_u32(uint32_t *A, uint64_t n, int p);
int or_permute_classic_u64(uint64_t *A, uint64_t n, int p);

/* ---- the involution family (family 1): cpu/src/involutions.{h,cpp},
 * cpu/src/bst-involutions.h, cpu/src/btree-involutions.h and, for the mvEB,
 * the host recursion of gpu/src/mveb-involutions.cuh (no CPU twin) ------- */
/* the low d base-`base` digits of x reversed (cpu/src/involutions.cpp:44-63),
 * the low b of d digits reversed (:66-80) */
uint64_t or_rev_base_d(uint64_t x, uint64_t base, uint64_t d);
uint64_t or_rev_base_b(uint64_t x, uint64_t base, uint64_t d, uint64_t b);
/* g * (r * inverse(x/g) mod (m/g)), g = gcd(x, m) (cpu/src/involutions.cpp:83-121) */
uint64_t or_involution(uint64_t r, uint64_t x, uint64_t m);
#define OR_INVDECL(SUF, T)                                                                   \
    void or_unshuffle_##SUF(T *A, uint64_t k, uint64_t n, uint64_t d);                      \
    void or_shuffle_dk_##SUF(T *A, uint64_t k, uint64_t n);                                 \
    void or_unshuffle_dk_##SUF(T *A, uint64_t k, uint64_t n);                               \
    void or_rotate_right_##SUF(T *A, uint64_t n, uint64_t k);                               \
    int or_permute_bst_inv_##SUF(T *A, uint64_t n);                                         \
    int or_permute_btree_inv_##SUF(T *A, uint64_t n, uint64_t b);                           \
    int or_permute_mveb_inv_##SUF(T *A, uint64_t n);
OR_INVDECL(u32, uint32_t)
OR_INVDECL(u64, uint64_t)
#undef OR_INVDECL

/* ---- restated queries: answers[i] = slot of queries[i] or n ------------ */
#define OR_QDECL(SUF, T)                                                                    \
    void or_search_bst_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,               \
                             uint64_t *ans, int p);                                         \
    void or_search_btree_##SUF(const T *A, uint64_t n, uint64_t b, const T *q, uint64_t nq, \
                               uint64_t *ans, int p);                                       \
    void or_search_mveb_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,              \
                              uint64_t *ans, int p);                                        \
    /* classic vEB layout A, tables of the classic rule */                                \
    void or_search_classic_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,           \
                                 uint64_t *ans, int p);                                     \
    /* sorted A: hit -> index, miss -> predecessor's index, none -> n */                    \
    void or_search_binary_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,            \
                                uint64_t *ans, int p);
OR_QDECL(u32, uint32_t)
OR_QDECL(u64, uint64_t)
OR_QDECL(i32, int32_t)
OR_QDECL(i64, int64_t)
#undef OR_QDECL

/* ---- seeded input generation (counter-based, splitmix64) --------------- */
uint64_t or_splitmix64(uint64_t seed, uint64_t i);

#ifdef __cplusplus
}
#endif
#endif
