// Binds the reference's CPU classic vEB permutation (cpu/src/veb-cycles.h:381),
// used only where the classic and the modified split coincide.
#include "veb-cycles.h"
extern "C" void ref_permute_veb_u64(uint64_t *A, uint64_t n, int p) { timePermutevEB<uint64_t>(A, n, p); }
