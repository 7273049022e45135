// Binds the reference's CPU binary search loop (cpu/src/binary-search.h:54-60).
// Callers pass only queries >= A[0]: below it the reference never returns.
#include "binary-search.h"
extern "C" void ref_search_binary_u64(uint64_t *A, uint64_t n, uint64_t *q, uint64_t nq, uint64_t *ans, int p) {
    searchAll<uint64_t>(A, n, q, ans, nq, (uint32_t)p);
}
