// Binds the reference's CPU B-tree permutation (cpu/src/btree-cycles.h:71).
#include "btree-cycles.h"
extern "C" void ref_permute_btree_u32(uint32_t *A, uint64_t n, uint64_t b, int p) { timePermuteBtree<uint32_t>(A, n, b, p); }
extern "C" void ref_permute_btree_u64(uint64_t *A, uint64_t n, uint64_t b, int p) { timePermuteBtree<uint64_t>(A, n, b, p); }
