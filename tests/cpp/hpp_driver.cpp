// A reference-style driver written against include/tree_layouts.hpp only
// (compare gpu/test/build-btree-cycles.cu:19-53, gpu/test/query-btree-fixedN.cu,
// gpu/test/query-mveb-fixedN.cu and gpu/test/query-binarysearch-fixedN.cu):
// ITERS timed permutations of initSortedList input through the
// reference-named templates, printTimings / printQueryTimings lines, and
// checks of what the templates produced:
//   * BST and B-tree layouts slot for slot against the in-order fill of the
//     heap-numbered node tree (node k holds slots k*b .. k*b+b-1 below n, its
//     children are nodes k*(b+1)+1 .. k*(b+1)+b+1);
//   * the mvEB layout by a digest of its rank sequence, printed as
//     "digest <type> <n> <hex>" for the test to compare with the oracle;
//   * every query answer (all hits) equal to the slot that holds the key;
//   * the involution programs' permutations (namespace involutions) by the
//     same checks ("digest-inv" lines for the mvEB).
// Exit status 0 when every check passes.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../include/tree_layouts.hpp"

// in-order fill: rank of every slot of a level-order tree with b keys per node
static void inorder(uint64_t node, uint64_t b, uint64_t n, uint64_t &next,
                    std::vector<uint64_t> &rank_of_slot) {
    if (node * b >= n) return;
    for (uint64_t j = 0; j <= b; j++) {
        inorder(node * (b + 1) + 1 + j, b, n, next, rank_of_slot);
        if (j < b && node * b + j < n) rank_of_slot[node * b + j] = next++;
    }
}

// order-sensitive digest of a rank layout: the sum over slots s of a
// splitmix64 mix of (rank, s), wrapping (tests/oracle_lib.py computes the same)
static uint64_t digest(const std::vector<uint64_t> &v) {
    uint64_t h = 0;
    for (uint64_t s = 0; s < v.size(); s++) {
        uint64_t z = v[s] * 0x9E3779B97F4A7C15ull + (s + 1) * 0xD1B54A32D192ED03ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        h += z ^ (z >> 31);
    }
    return h;
}

template <typename T>
static int run(uint64_t n, const char *tname) {
    std::vector<T> A(n);
    T *dev_A = nullptr;
    if (hipMalloc(&dev_A, n * sizeof(T)) != hipSuccess) return 1;
    int bad = 0;
    const uint64_t nq = 100000;
    for (int layout = 0; layout < 3; layout++) {
        double time[ITERS];
        for (int i = 0; i < ITERS; ++i) {
            initSortedList<T>(A.data(), n);
            if (hipMemcpy(dev_A, A.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
                return 1;
            time[i] = layout == 0 ? timePermuteBST<T>(dev_A, n)
                    : layout == 1 ? timePermuteBtree<T>(dev_A, n, 16)
                                  : timePermutevEB<T>(dev_A, n);
            if (time[i] < 0) bad++;
            if (hipMemcpy(A.data(), dev_A, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
                return 1;
        }
        printf("layout %d %s permute: ", layout, tname);
        printTimings(n, time);
        // ranks of the slots (keys are rank + 1), and where every rank sits
        std::vector<uint64_t> rank(n), slot_of(n, n);
        for (uint64_t s = 0; s < n; s++) {
            rank[s] = (uint64_t)A[s] - 1;
            if (rank[s] >= n || slot_of[rank[s]] != n) {
                bad++;
                continue;
            }
            slot_of[rank[s]] = s;
        }
        if (layout < 2) {
            std::vector<uint64_t> expect(n, n);
            uint64_t next = 0;
            inorder(0, layout == 0 ? 1 : 16, n, next, expect);
            uint64_t wrong = 0;
            for (uint64_t s = 0; s < n; s++) wrong += expect[s] != rank[s];
            if (wrong || next != n) {
                printf("layout %d %s: %llu slots differ from the in-order fill\n", layout, tname,
                       (unsigned long long)wrong);
                bad++;
            }
        } else {
            printf("digest %s %llu %016llx\n", tname, (unsigned long long)n,
                   (unsigned long long)digest(rank));
        }
        // queries: every answer is the slot of its key
        std::vector<uint64_t> ans;
        tl_detail::answer_sink() = &ans;
        double qtime[ITERS];
        for (int i = 0; i < ITERS; ++i) {
            if (layout == 0)
                qtime[i] = timeQueryBST<T>(A.data(), dev_A, n, nq, 7 + i);
            else if (layout == 1)
                qtime[i] = timeQueryBtree<T>(A.data(), dev_A, n, 16, nq, 7 + i);
            else {
                vEB_table *dt;
                uint64_t *di;
                uint32_t *dti, *meta;
                if (hipMalloc(&dt, 64 * 64 * sizeof(vEB_table)) != hipSuccess ||
                    hipMalloc(&di, 64 * 8) != hipSuccess || hipMalloc(&dti, 64 * 4) != hipSuccess ||
                    hipMalloc(&meta, 16) != hipSuccess)
                    return 1;
                if (tl_veb_tableset_device(n, dt, di, dti, meta, nullptr) != TL_OK) return 1;
                uint32_t hm[3];
                if (hipMemcpy(hm, meta, 12, hipMemcpyDeviceToHost) != hipSuccess) return 1;
                if (hm[1] == 1) {   // a perfect tree: the single-table form
                    std::vector<vEB_table> tbl(hm[0]);
                    buildTable(tbl.data(), n, hm[0], 0);
                    if (hipMemcpy(dt, tbl.data(), hm[0] * sizeof(vEB_table),
                                  hipMemcpyHostToDevice) != hipSuccess)
                        return 1;
                    qtime[i] = timeQueryvEB<T>(A.data(), dev_A, dt, n, hm[0], nq, 7 + i);
                } else {
                    qtime[i] = timeQueryvEB_nonperfect<T>(A.data(), dev_A, dt, di, dti, n, hm[0],
                                                          hm[1], nq, 7 + i);
                }
                (void)hipFree(dt);
                (void)hipFree(di);
                (void)hipFree(dti);
                (void)hipFree(meta);
            }
            if (qtime[i] < 0) bad++;
            const std::vector<T> hits = tl_detail::random_hits(A.data(), n, nq, 7 + i);
            uint64_t wrong = ans.size() != nq;
            for (uint64_t k = 0; k < nq && k < ans.size(); k++)
                wrong += ans[k] != slot_of[(uint64_t)hits[k] - 1];
            if (wrong) {
                printf("layout %d %s: %llu query answers differ\n", layout, tname,
                       (unsigned long long)wrong);
                bad++;
            }
        }
        tl_detail::answer_sink() = nullptr;
        printf("layout %d %s query: ", layout, tname);
        printQueryTimings(n, nq, qtime);
    }
    // the involution programs' names (tree_layouts.hpp, namespace involutions):
    // the same layouts, the mvEB's digest printed as "digest-inv"
    for (int layout = 0; layout < 3; layout++) {
        initSortedList<T>(A.data(), n);
        if (hipMemcpy(dev_A, A.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return 1;
        const double t = layout == 0 ? involutions::timePermuteBST<T>(dev_A, n)
                       : layout == 1 ? involutions::timePermuteBtree<T>(dev_A, n, 16)
                                     : involutions::timePermutevEB<T>(dev_A, n);
        if (t < 0) bad++;
        if (hipMemcpy(A.data(), dev_A, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            return 1;
        std::vector<uint64_t> rank(n);
        for (uint64_t s = 0; s < n; s++) rank[s] = (uint64_t)A[s] - 1;
        if (layout < 2) {
            std::vector<uint64_t> expect(n, n);
            uint64_t next = 0;
            inorder(0, layout == 0 ? 1 : 16, n, next, expect);
            uint64_t wrong = 0;
            for (uint64_t s = 0; s < n; s++) wrong += expect[s] != rank[s];
            if (wrong) bad++;
            printf("involutions layout %d %s n=%llu: %.3f ms, %llu slots differ\n", layout, tname,
                   (unsigned long long)n, t, (unsigned long long)wrong);
        } else {
            printf("digest-inv %s %llu %016llx\n", tname, (unsigned long long)n,
                   (unsigned long long)digest(rank));
        }
    }
    // the comparator on the sorted list (gpu/test/query-binarysearch-fixedN.cu):
    // the answer of a hit is its rank
    initSortedList<T>(A.data(), n);
    if (hipMemcpy(dev_A, A.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return 1;
    std::vector<uint64_t> ans;
    tl_detail::answer_sink() = &ans;
    const float bms = timeQuery<T>(A.data(), dev_A, n, nq, 7);
    tl_detail::answer_sink() = nullptr;
    if (bms < 0) bad++;
    const std::vector<T> hits = tl_detail::random_hits(A.data(), n, nq, 7);
    uint64_t wrong = ans.size() != nq;
    for (uint64_t k = 0; k < nq && k < ans.size(); k++) wrong += ans[k] != (uint64_t)hits[k] - 1;
    if (wrong) bad++;
    printf("binary search %s n=%llu, %llu queries %.3f ms, %llu wrong\n", tname,
           (unsigned long long)n, (unsigned long long)nq, bms, (unsigned long long)wrong);
    (void)hipFree(dev_A);
    return bad;
}

int main(int argc, char **argv) {
    uint64_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1000003;
    int bad = run<uint32_t>(n, "u32") + run<uint64_t>(n, "u64") +
              run<uint64_t>((1u << 20) - 1, "u64");
    printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
