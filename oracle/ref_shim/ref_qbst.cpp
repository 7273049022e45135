// Binds the reference's CPU BST query loop (cpu/src/query-bst.h:47-53).
#include "query-bst.h"
extern "C" void ref_search_bst_u64(uint64_t *A, uint64_t n, uint64_t *q, uint64_t nq, uint64_t *ans, int p) {
    searchAll<uint64_t>(A, n, q, ans, nq, p);
}
extern "C" void ref_search_bst_u32(uint32_t *A, uint64_t n, uint32_t *q, uint64_t nq, uint64_t *ans, int p) {
    searchAll<uint32_t>(A, n, q, ans, nq, p);
}
