// Serial CPU emulation of the uneven batch's single persistent launch (k_batch):
// the product's tl_batch.hip, compiled for the host over a shim of the HIP
// surface (tests/emu/hip/hip_runtime.h), runs its kernels one workgroup at a
// time with every lane a thread, on host memory, under AddressSanitizer.  This
// driver builds a mixed batch, runs the persistent launch through
// tl_batch_run, and checks every array against the oracle.
#include "hip/hip_runtime.h"
#include "tree_layouts.h"
#include "tl_oracle.h"
#include <math.h>
#include <stdio.h>
#include <random>
#include <vector>
#include <unistd.h>

int main(int argc, char **argv) {
    const int count = argc > 1 ? atoi(argv[1]) : 240;
    const double hi = argc > 2 ? atof(argv[2]) : 18.0;
    const uint32_t lds = argc > 3 ? (uint32_t)atoi(argv[3]) : 0;
    const uint64_t qmax = argc > 4 ? strtoull(argv[4], 0, 10) : 256;
    const uint64_t seed = argc > 5 ? strtoull(argv[5], 0, 10) : 5;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, hi);
    std::uniform_int_distribution<int> gap(1, 64);
    struct Arr { bool i32; int layout; uint64_t n, nq; std::vector<uint64_t> keys64; std::vector<int32_t> keys32;
                 void *A; void *Q; uint64_t *ans; std::vector<uint64_t> q64; std::vector<int32_t> q32; };
    std::vector<Arr> arrs(count);
    std::vector<tl_batch_item> items(count);
    uint64_t total = 0;
    const char *one = getenv("EMU_ONE");   // "layout kind n" for a one-array batch
    for (int i = 0; i < count; i++) {
        Arr &a = arrs[i];
        a.n = (uint64_t)floor(pow(2.0, U(rng)));
        a.i32 = (rng() & 1) == 0;
        a.layout = i % 3;
        if (one) {
            char kind[8];
            unsigned long long nn;
            sscanf(one, "%d %7s %llu", &a.layout, kind, &nn);
            a.n = nn;
            a.i32 = strcmp(kind, "i32") == 0;
        }
        total += a.n;
        uint64_t k = 0;
        std::vector<uint64_t> keys(a.n);
        for (uint64_t j = 0; j < a.n; j++) { k += gap(rng); keys[j] = k; }
        const uint64_t nhit = std::min<uint64_t>(4 * a.n, qmax), nmiss = nhit / 4;
        a.nq = nhit + nmiss;
        std::uniform_int_distribution<uint64_t> pick(0, a.n - 1);
        std::uniform_int_distribution<uint64_t> any(0, k + 3);
        if (a.i32) {
            const int64_t mid = (int64_t)keys[a.n / 2];
            a.keys32.resize(a.n);
            for (uint64_t j = 0; j < a.n; j++) a.keys32[j] = (int32_t)((int64_t)keys[j] - mid);
            for (uint64_t j = 0; j < nhit; j++) a.q32.push_back(a.keys32[pick(rng)]);
            for (uint64_t j = 0; j < nmiss; j++) a.q32.push_back((int32_t)((int64_t)any(rng) - mid));
            a.A = malloc(a.n * 4); memcpy(a.A, a.keys32.data(), a.n * 4);
            a.Q = malloc(a.nq * 4 + 4); if (a.nq) memcpy(a.Q, a.q32.data(), a.nq * 4);
        } else {
            a.keys64 = keys;
            for (uint64_t j = 0; j < nhit; j++) a.q64.push_back(keys[pick(rng)]);
            for (uint64_t j = 0; j < nmiss; j++) a.q64.push_back(any(rng));
            a.A = malloc(a.n * 8); memcpy(a.A, keys.data(), a.n * 8);
            a.Q = malloc(a.nq * 8 + 8); if (a.nq) memcpy(a.Q, a.q64.data(), a.nq * 8);
        }
        a.ans = (uint64_t *)malloc(a.nq * 8);
        for (uint64_t j = 0; j < a.nq; j++) a.ans[j] = ~0ull;
        items[i] = tl_batch_item{a.A, a.n, a.i32 ? TL_I32 : TL_U64, (uint32_t)a.layout,
                                 (uint64_t)(a.layout == 1 ? 8 : 1), a.Q, a.nq, a.ans};
    }
    fprintf(stderr, "[emu] %d arrays, %llu keys\n", count, (unsigned long long)total);
    tl_batch_config cfg{lds, 4, 0};
    void *h = nullptr;
    tl_status st = tl_batch_create(items.data(), count, &cfg, &h);
    if (st != TL_OK) { fprintf(stderr, "[emu] create: %d %s\n", st, tl_last_error()); return 2; }
    uint64_t ops = 0, tasks = 0, smem = 0;
    tl_batch_stats(h, &ops, &tasks, &smem);
    fprintf(stderr, "[emu] ops %llu tasks %llu smem %llu\n", (unsigned long long)ops, (unsigned long long)tasks, (unsigned long long)smem);
    st = tl_batch_run(h, nullptr, 0, nullptr);
    fprintf(stderr, "[emu] run status %d\n", st);
    extern volatile uint64_t emu_ticket;
    fprintf(stderr, "[emu] last ticket %llu\n", (unsigned long long)emu_ticket);
    int bad = 0;
    for (int i = 0; i < count; i++) {
        Arr &a = arrs[i];
        std::vector<uint64_t> exp_ans(a.nq);
        bool lay_ok, ans_ok;
        if (a.i32) {
            std::vector<int32_t> e = a.keys32;
            if (a.layout == 0) or_permute_bst_u32((uint32_t *)e.data(), a.n, 1);
            else if (a.layout == 1) or_permute_btree_u32((uint32_t *)e.data(), a.n, 8, 1);
            else or_permute_mveb_u32((uint32_t *)e.data(), a.n);
            lay_ok = memcmp(e.data(), a.A, a.n * 4) == 0;
            if (a.layout == 0) or_search_bst_i32(e.data(), a.n, a.q32.data(), a.nq, exp_ans.data(), 1);
            else if (a.layout == 1) or_search_btree_i32(e.data(), a.n, 8, a.q32.data(), a.nq, exp_ans.data(), 1);
            else or_search_mveb_i32(e.data(), a.n, a.q32.data(), a.nq, exp_ans.data(), 1);
        } else {
            std::vector<uint64_t> e = a.keys64;
            if (a.layout == 0) or_permute_bst_u64(e.data(), a.n, 1);
            else if (a.layout == 1) or_permute_btree_u64(e.data(), a.n, 8, 1);
            else or_permute_mveb_u64(e.data(), a.n);
            lay_ok = memcmp(e.data(), a.A, a.n * 8) == 0;
            if (a.layout == 0) or_search_bst_u64(e.data(), a.n, a.q64.data(), a.nq, exp_ans.data(), 1);
            else if (a.layout == 1) or_search_btree_u64(e.data(), a.n, 8, a.q64.data(), a.nq, exp_ans.data(), 1);
            else or_search_mveb_u64(e.data(), a.n, a.q64.data(), a.nq, exp_ans.data(), 1);
        }
        ans_ok = a.nq == 0 || memcmp(exp_ans.data(), a.ans, a.nq * 8) == 0;
        if (!lay_ok || !ans_ok) {
            if (bad < 20) fprintf(stderr, "[emu] array %d layout %d %s n=%llu: layout %s answers %s\n", i, a.layout,
                                  a.i32 ? "i32" : "u64", (unsigned long long)a.n, lay_ok ? "ok" : "BAD", ans_ok ? "ok" : "BAD");
            bad++;
        }
    }
    fprintf(stderr, "[emu] %d of %d arrays differ from the oracle\n", bad, count);
    fflush(stderr);
    _exit(bad ? 1 : 0);
}
