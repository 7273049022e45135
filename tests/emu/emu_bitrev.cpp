// CPU emulation of the BST of 2^D keys kept one place right: the
// bit-reversal tiles of tl_kernels.hpp in all their forms (plain; phase one
// absorbing row ends; the levels shifted in place / unfolded) and the level
// heads copy, built for the host over the HIP shim and run one workgroup at a
// time with every lane a thread, on exactly sized host buffers under
// AddressSanitizer.
//
// The forward sequence (phase one in place, the level heads and the leaf to
// the stash, the top levels read one place right, the levels shifted, the
// leaf last) must give what the plan without the fold gives after its leaf
// move: the plain phases of the keys after the leaf, then the leaf.  The
// inverse sequence (the leaf to the stash, the levels unfolded, the top
// levels written one place right, the leaf first, phase one absorbing the row
// ends) must restore the input.  The top levels (an in-LDS launch on the
// device) are done here on the host; every tile launch shape runs (one
// workgroup per tile number, a looping grid, tile numbers XCD by XCD).
#include "hip/hip_runtime.h"
#include "src/tl_kernels.hpp"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

using namespace tl;

template <typename T>
static T *buf(uint64_t n) {
    return (T *)malloc(n * sizeof(T) ? n * sizeof(T) : 1);
}

// phase one, base b: one workgroup per tile number (shape 0), a looping grid
// of 3 (1), or tile numbers XCD by XCD over a grid padded to a multiple of 8 (2)
template <typename T, int K, int FOLD>
static void phase_one(T *A, int64_t b, uint32_t D, T *S, uint32_t t0, int shape) {
    const uint64_t tiles = 1ull << (D - 2 * K);
    if (shape == 1) {
        const uint32_t g = (uint32_t)(tiles > 3 ? 3 : tiles);
        hipLaunchKernelGGL((k_bitrev_tiles_loop<T, K, FOLD>), dim3(g), dim3(256), 0, 0, A, b, D,
                           tiles, S, t0);
    } else {
        const uint32_t xcd = shape == 2 ? 1u : 0u;
        const uint64_t g = xcd ? (tiles + 7) & ~7ull : tiles;
        hipLaunchKernelGGL((k_bitrev_tiles<T, K, FOLD>), dim3((uint32_t)g), dim3(256), 0, 0, A, b,
                           D, S, t0, xcd);
    }
}

template <typename T, int K, int FOLD>
static void levels(T *A, uint32_t d, T *S, int shape) {
    const uint32_t t0 = 2 * K;
    const uint64_t tiles = (1ull << (d - t0)) - 1;
    if (shape == 1) {
        const uint32_t g = (uint32_t)(tiles > 3 ? 3 : tiles);
        hipLaunchKernelGGL((k_bitrev_levels_loop<T, K, FOLD>), dim3(g), dim3(256), 0, 0, A, t0, d,
                           S);
    } else {
        const uint32_t xcd = shape == 2 ? 1u : 0u;
        const uint64_t g = xcd ? (tiles + 8) & ~7ull : tiles;
        hipLaunchKernelGGL((k_bitrev_levels<T, K, FOLD>), dim3((uint32_t)g), dim3(256), 0, 0, A,
                           t0, d, xcd, S);
    }
}

// the per-level bit reversal of the top m = 2^t0 - 1 keys: slot x (0-based)
// goes to level t = floor(log2(x + 1)), place rev_t(x + 1 - 2^t)
static uint32_t top_pos(uint32_t x) {
    const uint32_t y = x + 1, t = 31u - (uint32_t)__builtin_clz(y);
    uint32_t r = 0;
    for (uint32_t i = 0; i < t; i++) r |= ((y >> i) & 1u) << (t - 1 - i);
    return (1u << t) + r - 1;
}

template <typename T>
static void top(T *A, uint32_t m, int shift, bool inv) {
    T *tmp = buf<T>(m);
    if (!inv) {
        for (uint32_t x = 0; x < m; x++) tmp[x] = A[x + shift];
        for (uint32_t x = 0; x < m; x++) A[top_pos(x)] = tmp[x];
    } else {
        for (uint32_t x = 0; x < m; x++) tmp[x] = A[top_pos(x)];
        for (uint32_t x = 0; x < m; x++) A[x + shift] = tmp[x];
    }
    free(tmp);
}

template <typename T, int K>
static int check(uint32_t D, int shape) {
    const uint32_t t0 = 2 * K;
    const uint64_t n = 1ull << D, rows = ((1ull << D) - (1ull << t0)) >> K;
    T *A = buf<T>(n), *E = buf<T>(n - 1), *F = buf<T>(n), *S = buf<T>(rows + 1);
    uint64_t x = 0x9E3779B97F4A7C15ull * (D + 17 * K + 5 * shape);
    for (uint64_t i = 0; i < n; i++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        A[i] = (T)x;
    }
    // expected: the plain phases of the keys after the leaf, then the leaf
    memcpy(E, A + 1, (n - 1) * sizeof(T));
    phase_one<T, K, 0>(E, -1, D, nullptr, 0, shape);
    levels<T, K, 0>(E, D, nullptr, shape);
    top<T>(E, (1u << t0) - 1, 0, false);
    // the forward sequence
    memcpy(F, A, n * sizeof(T));
    memset(S, 0x5A, (rows + 1) * sizeof(T));
    phase_one<T, K, 0>(F, 0, D, nullptr, 0, shape);
    hipLaunchKernelGGL((k_level_heads<T, K>), dim3(3), dim3(256), 0, 0, (const T *)F, S, t0, D);
    S[rows] = F[0];
    top<T>(F, (1u << t0) - 1, 1, false);
    levels<T, K, 1>(F, D, S, shape);
    F[n - 1] = S[rows];
    int bad = 0;
    for (uint64_t i = 0; i + 1 < n; i++) bad += F[i] != E[i];
    bad += F[n - 1] != A[0];
    // the inverse sequence
    memset(S, 0xA5, (rows + 1) * sizeof(T));
    S[rows] = F[n - 1];
    levels<T, K, 2>(F, D, S, shape);
    top<T>(F, (1u << t0) - 1, 1, true);
    F[0] = S[rows];
    phase_one<T, K, 4>(F, 0, D, S, t0, shape);
    int bad_inv = 0;
    for (uint64_t i = 0; i < n; i++) bad_inv += F[i] != A[i];
    fprintf(stderr, "[emu-bitrev] W=%zu K=%d D=%u shape=%d: fold %d wrong, unfold %d wrong\n",
            sizeof(T), K, D, shape, bad, bad_inv);
    free(A);
    free(E);
    free(F);
    free(S);
    return bad + bad_inv;
}

int main() {
    int bad = 0;
    for (int shape = 0; shape < 3; shape++) {
        for (uint32_t D : {12u, 13u, 15u}) {
            bad += check<uint32_t, 5>(D, shape);
            bad += check<uint64_t, 5>(D, shape);
        }
        for (uint32_t D : {14u, 15u}) {
            bad += check<uint32_t, 6>(D, shape);
            bad += check<uint64_t, 6>(D, shape);
        }
    }
    fprintf(stderr, "[emu-bitrev] %d wrong in all\n", bad);
    fflush(stderr);
    _exit(bad ? 1 : 0);
}
