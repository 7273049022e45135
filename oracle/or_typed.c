/*
 * or_typed.c -- instantiates the key-type generic oracle code.
 * TEST INFRASTRUCTURE ONLY (see tl_oracle.h).
 */
#include "tl_oracle.h"

#include <stdint.h>

#define T uint32_t
#define SUF u32
#include "or_typed.inc"
#undef T
#undef SUF

#define T uint64_t
#define SUF u64
#include "or_typed.inc"
#undef T
#undef SUF

/* ---- queries ------------------------------------------------------------
 * The three layout searches return the slot of the query key, or n when it
 * is absent; the binary search of the sorted array answers misses with the
 * predecessor's index.                                                     */

#define OR_QUERIES(SUF, T)                                                                     \
    /* cpu/src/query-bst.h:30-43 */                                                            \
    static uint64_t bst_one_##SUF(const T *A, uint64_t n, T q) {                               \
        uint64_t i = 0;                                                                        \
        while (i < n) {                                                                        \
            if (q == A[i]) return i;                                                           \
            i = 2 * i + 1 + (q > A[i]);                                                        \
        }                                                                                      \
        return n;                                                                              \
    }                                                                                          \
    void or_search_bst_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq, uint64_t *ans,   \
                             int p) {                                                          \
        _Pragma("omp parallel for schedule(guided) num_threads(p) if (p > 1)")                 \
        for (uint64_t i = 0; i < nq; i++) ans[i] = bst_one_##SUF(A, n, q[i]);                  \
    }                                                                                          \
    /* cpu/src/query-btree.h:33-54; a key slot at or past n compares as +inf,                  \
     * which only matters for misses in the partial last node */                               \
    static uint64_t btree_one_##SUF(const T *A, uint64_t n, uint64_t b, T q) {                 \
        uint64_t i = 0;                                                                        \
        while (i < n) {                                                                        \
            uint64_t j = 0;                                                                    \
            for (; j < b; j++) {                                                               \
                if (i + j >= n) break;                                                         \
                T cur = A[i + j];                                                              \
                if (q == cur) return i + j;                                                    \
                if (q < cur) break;                                                            \
            }                                                                                  \
            i = (i + 1) * (b + 1) + j * b - 1;                                                 \
        }                                                                                      \
        return n;                                                                              \
    }                                                                                          \
    void or_search_btree_##SUF(const T *A, uint64_t n, uint64_t b, const T *q, uint64_t nq,    \
                               uint64_t *ans, int p) {                                         \
        _Pragma("omp parallel for schedule(guided) num_threads(p) if (p > 1)")                 \
        for (uint64_t i = 0; i < nq; i++) ans[i] = btree_one_##SUF(A, n, b, q[i]);             \
    }                                                                                          \
    /* gpu/src/query-mveb.cuh:43-68 and :173-213.  The search of structure j                  \
     * stops at depth d_j (the reference reads one table entry past the end                    \
     * there); running out of depth means the key is absent. */                                \
    static uint64_t mveb_one_##SUF(const T *A, uint64_t n, const or_veb_table *tables,         \
                                   const uint64_t *idx, const uint32_t *tidx, uint32_t nt,     \
                                   T q) {                                                      \
        uint64_t pos[66];                                                                      \
        uint64_t offset = 0;                                                                   \
        for (uint32_t j = 0; j < nt; j++) {                                                    \
            const or_veb_table *tb = tables + tidx[j];                                         \
            uint32_t dj = 63u - (uint32_t)__builtin_clzll(n - offset) + 1;                     \
            uint64_t bfs = 1, index = 0;                                                       \
            uint32_t cd = 0;                                                                   \
            pos[0] = 1;                                                                        \
            while (index + offset < idx[j]) {                                                  \
                T cur = A[offset + index];                                                     \
                if (q == cur) return offset + index;                                           \
                bfs = 2 * bfs + (q > cur);                                                     \
                cd++;                                                                          \
                if (cd >= dj) return n;                                                        \
                pos[cd] = pos[tb[cd].D] + tb[cd].R + (bfs & tb[cd].R) * tb[cd].L;              \
                index = pos[cd] - 1;                                                           \
            }                                                                                  \
            offset = idx[j];                                                                   \
        }                                                                                      \
        return n;                                                                              \
    }                                                                                          \
    /* cpu/src/query-veb.h:103-137: the same search over the classic vEB      \
     * with the classic table set; running out of depth answers n */          \
    void or_search_classic_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,              \
                                 uint64_t *ans, int p) {                                       \
        if (n == 0) {                                                                          \
            for (uint64_t i = 0; i < nq; i++) ans[i] = 0;                                      \
            return;                                                                            \
        }                                                                                      \
        or_veb_table tables[64 * 64];                                                          \
        uint64_t idx[64];                                                                      \
        uint32_t tidx[64], d, nt, ts;                                                          \
        or_classic_tableset(n, &d, &nt, &ts, tables, idx, tidx);                               \
        _Pragma("omp parallel for schedule(guided) num_threads(p) if (p > 1)")                 \
        for (uint64_t i = 0; i < nq; i++)                                                      \
            ans[i] = mveb_one_##SUF(A, n, tables, idx, tidx, nt, q[i]);                        \
    }                                                                                          \
    /* cpu/src/binary-search.h:30-50 on the sorted array: a hit answers the    \
     * key's index, a miss the index of its predecessor (`right` at the end of \
     * the loop).  With no predecessor (q < A[0], or n = 0) the reference's    \
     * unsigned `right` wraps and its loop never ends; here that answers n. */ \
    static uint64_t binary_one_##SUF(const T *A, uint64_t n, T q) {                            \
        if (n == 0) return n;                                                                  \
        uint64_t left = 0, right = n - 1;                                                      \
        while (left <= right) {                                                                \
            const uint64_t mid = (left + right + 1) / 2;                                       \
            if (q == A[mid]) return mid;                                                       \
            if (q > A[mid]) {                                                                  \
                left = mid + 1;                                                                \
            } else {                                                                           \
                if (mid == 0) return n;                                                        \
                right = mid - 1;                                                               \
            }                                                                                  \
        }                                                                                      \
        return right;                                                                          \
    }                                                                                          \
    void or_search_binary_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq,               \
                                uint64_t *ans, int p) {                                        \
        _Pragma("omp parallel for schedule(guided) num_threads(p) if (p > 1)")                 \
        for (uint64_t i = 0; i < nq; i++) ans[i] = binary_one_##SUF(A, n, q[i]);              \
    }                                                                                          \
    void or_search_mveb_##SUF(const T *A, uint64_t n, const T *q, uint64_t nq, uint64_t *ans,  \
                              int p) {                                                         \
        if (n == 0) {                                                                          \
            for (uint64_t i = 0; i < nq; i++) ans[i] = 0;                                      \
            return;                                                                            \
        }                                                                                      \
        or_veb_table tables[64 * 64];                                                          \
        uint64_t idx[64];                                                                      \
        uint32_t tidx[64], d, nt, ts;                                                          \
        or_veb_tableset(n, &d, &nt, &ts, tables, idx, tidx);                                   \
        _Pragma("omp parallel for schedule(guided) num_threads(p) if (p > 1)")                 \
        for (uint64_t i = 0; i < nq; i++)                                                      \
            ans[i] = mveb_one_##SUF(A, n, tables, idx, tidx, nt, q[i]);                        \
    }

OR_QUERIES(u32, uint32_t)
OR_QUERIES(u64, uint64_t)
OR_QUERIES(i32, int32_t)
OR_QUERIES(i64, int64_t)
