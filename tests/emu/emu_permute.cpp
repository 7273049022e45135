// CPU emulation of the single-array permutations: the product's
// tl_permute.hip, compiled for the host over the HIP shim, runs every plan
// launch one workgroup at a time with every lane a thread, on exactly sized
// host buffers under AddressSanitizer.  Chained moves take their tickets in
// order (one workgroup runs to the end before the next starts), so every
// wait is already satisfied when it is reached.  Each case permutes 0..n-1,
// compares every slot with the CPU oracle, runs the inverse and checks the
// identity.  Arguments: layout (bst, btree, mveb, veb), width (4, 8), then
// sizes.
#include "hip/hip_runtime.h"
#include "tree_layouts.h"
#include "tl_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

template <typename T>
static int run(const char *layout, uint64_t n) {
    T *A = (T *)malloc(n * sizeof(T) ? n * sizeof(T) : 1);
    T *E = (T *)malloc(n * sizeof(T) ? n * sizeof(T) : 1);
    for (uint64_t i = 0; i < n; i++) A[i] = E[i] = (T)i;
    const tl_dtype dt = sizeof(T) == 4 ? TL_U32 : TL_U64;
    tl_status st;
    if (!strcmp(layout, "bst")) {
        st = tl_permute_bst(A, n, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_bst_u32((uint32_t *)E, n, 1);
        else or_permute_bst_u64((uint64_t *)E, n, 1);
    } else if (!strcmp(layout, "btree")) {
        st = tl_permute_btree(A, n, 16, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_btree_u32((uint32_t *)E, n, 16, 1);
        else or_permute_btree_u64((uint64_t *)E, n, 16, 1);
    } else if (!strcmp(layout, "mveb")) {
        st = tl_permute_mveb(A, n, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_mveb_u32((uint32_t *)E, n);
        else or_permute_mveb_u64((uint64_t *)E, n);
    } else {
        st = tl_permute_veb(A, n, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_classic_u32((uint32_t *)E, n, 1);
        else or_permute_classic_u64((uint64_t *)E, n, 1);
    }
    uint64_t bad = 0;
    for (uint64_t i = 0; i < n; i++) bad += A[i] != E[i];
    const tl_status st2 = !strcmp(layout, "bst")     ? tl_unpermute_bst(A, n, dt, nullptr, nullptr)
                          : !strcmp(layout, "btree") ? tl_unpermute_btree(A, n, 16, dt, nullptr, nullptr)
                          : !strcmp(layout, "mveb")  ? tl_unpermute_mveb(A, n, dt, nullptr, nullptr)
                                                     : tl_unpermute_veb(A, n, dt, nullptr, nullptr);
    uint64_t bad_inv = 0;
    for (uint64_t i = 0; i < n; i++) bad_inv += A[i] != (T)i;
    fprintf(stderr, "[emu-permute] %s W=%zu n=%llu: status %d/%d, %llu wrong, inverse %llu wrong\n",
            layout, sizeof(T), (unsigned long long)n, (int)st, (int)st2, (unsigned long long)bad,
            (unsigned long long)bad_inv);
    free(A);
    free(E);
    return (st != TL_OK || st2 != TL_OK || bad || bad_inv) ? 1 : 0;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: emu_permute layout width n...\n");
        _exit(2);
    }
    const int W = atoi(argv[2]);
    int bad = 0;
    for (int a = 3; a < argc; a++) {
        const uint64_t n = strtoull(argv[a], nullptr, 10);
        bad += W == 4 ? run<uint32_t>(argv[1], n) : run<uint64_t>(argv[1], n);
    }
    fprintf(stderr, "[emu-permute] %d cases wrong\n", bad);
    fflush(stderr);
    _exit(bad ? 1 : 0);
}
