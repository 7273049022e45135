// Binds the reference's CPU B-tree query loop (cpu/src/query-btree.h:59-64).
#include "query-btree.h"
extern "C" void ref_search_btree_u64(uint64_t *A, uint64_t n, uint64_t b, uint64_t *q, uint64_t nq, uint64_t *ans, int p) {
    searchAll<uint64_t>(A, n, b, q, ans, nq, p);
}
extern "C" void ref_search_btree_u32(uint32_t *A, uint64_t n, uint64_t b, uint32_t *q, uint64_t nq, uint64_t *ans, int p) {
    searchAll<uint32_t>(A, n, b, q, ans, nq, p);
}
