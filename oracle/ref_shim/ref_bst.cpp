// Binds the reference's CPU BST permutation (cpu/src/bst-hybrid.h:120) for
// the oracle checks.  Compiled only by `make ref` from the reference tree.
#include "bst-hybrid.h"
extern "C" void ref_permute_bst_u32(uint32_t *A, uint64_t n, int p) { timePermuteBST<uint32_t>(A, n, p); }
extern "C" void ref_permute_bst_u64(uint64_t *A, uint64_t n, int p) { timePermuteBST<uint64_t>(A, n, p); }
