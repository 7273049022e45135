// Binds the reference's CPU BST involution permutation
// (cpu/src/bst-involutions.h:119) and its shuffle primitives
// (cpu/src/involutions.h:41,103,161) for the oracle checks.  Compiled only by
// `make ref` from the reference tree.
#include "bst-involutions.h"
extern "C" void ref_permute_bst_inv_u64(uint64_t *A, uint64_t n, int p) { timePermuteBST<uint64_t>(A, n, p); }
extern "C" void ref_unshuffle_u64(uint64_t *A, uint64_t k, uint64_t n, uint64_t d) { unshuffle<uint64_t>(A, k, n, d); }
extern "C" void ref_shuffle_dk_u64(uint64_t *A, uint64_t k, uint64_t n) { shuffle_dk<uint64_t>(A, k, n); }
extern "C" void ref_unshuffle_dk_u64(uint64_t *A, uint64_t k, uint64_t n) { unshuffle_dk<uint64_t>(A, k, n); }
