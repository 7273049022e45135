// CPU emulation of the single-array permutations: the product's
// tl_permute.hip, compiled for the host over the HIP shim, runs every plan
// launch one workgroup at a time with every lane a thread, on exactly sized
// host buffers under AddressSanitizer.  Chained moves take their tickets in
// order (one workgroup runs to the end before the next starts), so every
// wait is already satisfied when it is reached.  Each case permutes 0..n-1,
// compares every slot with the CPU oracle, runs the inverse and checks the
// identity.  The involution family (bst_inv, btree_inv with b = 16, mveb_inv)
// is compared with the same oracle layouts, and `prims` checks the shuffle,
// unshuffle and rotation primitives against their closed forms.  Arguments:
// layout (bst, btree, mveb, veb, bst_inv, btree_inv, mveb_inv, prims),
// width (4, 8), then sizes.
#include "hip/hip_runtime.h"
#include "tree_layouts.h"
#include "tl_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

// the involution family: the layout of the gather / cycle-leader oracle
template <typename T>
static int run_inv(const char *layout, uint64_t n, uint64_t b) {
    T *A = (T *)malloc(n * sizeof(T) ? n * sizeof(T) : 1);
    T *E = (T *)malloc(n * sizeof(T) ? n * sizeof(T) : 1);
    for (uint64_t i = 0; i < n; i++) A[i] = E[i] = (T)i;
    const tl_dtype dt = sizeof(T) == 4 ? TL_U32 : TL_U64;
    tl_status st;
    if (!strcmp(layout, "bst_inv")) {
        st = tl_permute_bst_involutions(A, n, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_bst_u32((uint32_t *)E, n, 1);
        else or_permute_bst_u64((uint64_t *)E, n, 1);
    } else if (!strcmp(layout, "btree_inv")) {
        st = tl_permute_btree_involutions(A, n, b, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_btree_u32((uint32_t *)E, n, b, 1);
        else or_permute_btree_u64((uint64_t *)E, n, b, 1);
    } else {
        st = tl_permute_mveb_involutions(A, n, dt, nullptr, nullptr);
        if (sizeof(T) == 4) or_permute_mveb_u32((uint32_t *)E, n);
        else or_permute_mveb_u64((uint64_t *)E, n);
    }
    uint64_t bad = 0;
    for (uint64_t i = 0; i < n; i++) bad += A[i] != E[i];
    fprintf(stderr, "[emu-permute] %s W=%zu n=%llu b=%llu: status %d, %llu wrong\n", layout,
            sizeof(T), (unsigned long long)n, (unsigned long long)b, (int)st,
            (unsigned long long)bad);
    free(A);
    free(E);
    return (st != TL_OK || bad) ? 1 : 0;
}

// the primitives against their closed forms, for k in a few values dividing n
// (shuffles), n = k^d - 1 (unshuffle) and k <= n (rotation)
template <typename T>
static int run_prims(uint64_t n) {
    const tl_dtype dt = sizeof(T) == 4 ? TL_U32 : TL_U64;
    T *A = (T *)malloc((n + 1) * sizeof(T));
    int bad = 0;
    for (uint64_t k = 1; k <= 64 && k <= n; k++) {
        if (n % k) continue;
        const uint64_t d = n / k;
        for (int inv = 0; inv < 2; inv++) {
            for (uint64_t i = 0; i < n; i++) A[i] = (T)i;
            const tl_status st = inv ? tl_unshuffle_dk(A, n, k, dt, nullptr, nullptr)
                                     : tl_shuffle_dk(A, n, k, dt, nullptr, nullptr);
            uint64_t w = 0;
            for (uint64_t x = 0; x < n; x++) {
                // shuffle: slot j*k + i holds key i*d + j; unshuffle_dk the inverse
                const uint64_t i = x % k, j = x / k;
                const uint64_t want = inv ? (x % d) * k + x / d : i * d + j;
                w += A[x] != (T)want;
            }
            if (st != TL_OK || w) {
                fprintf(stderr, "[emu-permute] %s n=%llu k=%llu: status %d, %llu wrong\n",
                        inv ? "unshuffle_dk" : "shuffle_dk", (unsigned long long)n,
                        (unsigned long long)k, (int)st, (unsigned long long)w);
                bad++;
            }
        }
    }
    for (uint64_t k = 2; k <= 64; k++) {
        uint64_t m = k;
        while (m - 1 < n) m *= k;
        m /= k;   // largest k^d <= n
        if (m < k) continue;
        const uint64_t len = m - 1;
        for (uint64_t i = 0; i < len; i++) A[i] = (T)i;
        const tl_status st = tl_unshuffle(A, len, k, dt, nullptr, nullptr);
        // stable partition by (i + 1) mod k: residue 0, then 1 .. k-1
        uint64_t w = 0, at = 0;
        for (uint64_t r = 0; r < k; r++)
            for (uint64_t i = 0; i < len; i++)
                if ((i + 1) % k == r) w += A[at++] != (T)i;
        if (st != TL_OK || w) {
            fprintf(stderr, "[emu-permute] unshuffle n=%llu k=%llu: status %d, %llu wrong\n",
                    (unsigned long long)len, (unsigned long long)k, (int)st,
                    (unsigned long long)w);
            bad++;
        }
    }
    for (uint64_t k : {(uint64_t)0, (uint64_t)1, n / 3, n / 2, n - 1, n}) {
        if (k > n) continue;
        for (uint64_t i = 0; i < n; i++) A[i] = (T)i;
        const tl_status st = tl_shift_right(A, n, k, dt, nullptr, nullptr);
        uint64_t w = 0;
        for (uint64_t x = 0; x < n; x++) w += A[x] != (T)((x + n - k) % n);
        if (st != TL_OK || w) {
            fprintf(stderr, "[emu-permute] shift_right n=%llu k=%llu: status %d, %llu wrong\n",
                    (unsigned long long)n, (unsigned long long)k, (int)st, (unsigned long long)w);
            bad++;
        }
    }
    fprintf(stderr, "[emu-permute] prims W=%zu n=%llu: %d wrong\n", sizeof(T),
            (unsigned long long)n, bad);
    free(A);
    return bad ? 1 : 0;
}

template <typename T>
static int run(const char *layout, uint64_t n) {
    if (!strcmp(layout, "prims")) return run_prims<T>(n);
    if (!strcmp(layout, "bst_inv") || !strcmp(layout, "mveb_inv")) return run_inv<T>(layout, n, 1);
    if (!strcmp(layout, "btree_inv")) {
        int bad = 0;
        for (uint64_t b : {(uint64_t)1, (uint64_t)2, (uint64_t)3, (uint64_t)8, (uint64_t)16})
            bad += run_inv<T>(layout, n, b);
        return bad;
    }
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
