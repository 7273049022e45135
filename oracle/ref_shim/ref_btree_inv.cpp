// Binds the reference's CPU B-tree involution permutation
// (cpu/src/btree-involutions.h:79).  Compiled only by `make ref`.
#include "btree-involutions.h"
extern "C" void ref_permute_btree_inv_u64(uint64_t *A, uint64_t n, uint64_t b, int p) { timePermuteBtree<uint64_t>(A, n, b, p); }
