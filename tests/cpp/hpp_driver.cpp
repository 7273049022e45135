// A reference-style driver written against include/tree_layouts.hpp only
// (compare gpu/test/query-btree-fixedN.cu, gpu/test/query-mveb-fixedN.cu and
// gpu/test/query-binarysearch-fixedN.cu):
// build each layout on the GPU, copy it back, and query it through the
// reference-named templates.  Exit status 0 when every check passes.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../include/tree_layouts.hpp"

template <typename T>
static int run(uint64_t n) {
    std::vector<T> A(n);
    for (uint64_t i = 0; i < n; i++) A[i] = (T)(2 * i + 1);   // initSortedList, odd keys
    T *dev_A = nullptr;
    if (hipMalloc(&dev_A, n * sizeof(T)) != hipSuccess) return 1;
    int bad = 0;
    for (int layout = 0; layout < 3; layout++) {
        for (uint64_t i = 0; i < n; i++) A[i] = (T)(2 * i + 1);
        if (hipMemcpy(dev_A, A.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return 1;
        double ms = layout == 0 ? timePermuteBST<T>(dev_A, n)
                  : layout == 1 ? timePermuteBtree<T>(dev_A, n, 16)
                                : timePermutevEB<T>(dev_A, n);
        if (ms < 0) bad++;
        if (hipMemcpy(A.data(), dev_A, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return 1;
        // every key once: the layout is a permutation of the input
        std::vector<char> seen(n, 0);
        for (uint64_t i = 0; i < n; i++) {
            uint64_t r = ((uint64_t)A[i] - 1) / 2;
            if (r >= n || seen[r]) bad++;
            else seen[r] = 1;
        }
        float qms;
        if (layout == 0) qms = timeQueryBST<T>(A.data(), dev_A, n, 100000, 7);
        else if (layout == 1) qms = timeQueryBtree<T>(A.data(), dev_A, n, 16, 100000, 7);
        else {
            uint32_t d = 0;
            while ((1ull << d) - 1 < n) d++;
            std::vector<vEB_table> tbl(d);
            if ((1ull << d) - 1 == n) {
                buildTable(tbl.data(), n, d, 0);
                vEB_table *dt;
                if (hipMalloc(&dt, d * sizeof(vEB_table)) != hipSuccess) return 1;
                if (hipMemcpy(dt, tbl.data(), d * sizeof(vEB_table), hipMemcpyHostToDevice) != hipSuccess) return 1;
                qms = timeQueryvEB<T>(A.data(), dev_A, dt, n, d, 100000, 7);
                (void)hipFree(dt);
            } else {
                vEB_table *dt; uint64_t *di; uint32_t *dti, *meta;
                if (hipMalloc(&dt, 64 * 64 * sizeof(vEB_table)) != hipSuccess) return 1;
                if (hipMalloc(&di, 64 * 8) != hipSuccess || hipMalloc(&dti, 64 * 4) != hipSuccess ||
                    hipMalloc(&meta, 16) != hipSuccess) return 1;
                if (tl_veb_tableset_device(n, dt, di, dti, meta, nullptr) != TL_OK) return 1;
                uint32_t hm[3];
                if (hipMemcpy(hm, meta, 12, hipMemcpyDeviceToHost) != hipSuccess) return 1;
                qms = timeQueryvEB_nonperfect<T>(A.data(), dev_A, dt, di, dti, n, hm[0], hm[1], 100000, 7);
                (void)hipFree(dt); (void)hipFree(di); (void)hipFree(dti); (void)hipFree(meta);
            }
        }
        if (qms < 0) bad++;
        printf("layout %d n=%llu permute %.3f ms, 100000 queries %.3f ms\n", layout,
               (unsigned long long)n, ms, qms);
    }
    // the comparator on the sorted list (gpu/test/query-binarysearch-fixedN.cu)
    for (uint64_t i = 0; i < n; i++) A[i] = (T)(2 * i + 1);
    if (hipMemcpy(dev_A, A.data(), n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return 1;
    const float bms = timeQuery<T>(A.data(), dev_A, n, 100000, 7);
    if (bms < 0) bad++;
    printf("binary search n=%llu, 100000 queries %.3f ms\n", (unsigned long long)n, bms);
    (void)hipFree(dev_A);
    return bad;
}

int main(int argc, char **argv) {
    uint64_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1000003;
    int bad = run<uint32_t>(n) + run<uint64_t>(n) + run<uint64_t>((1u << 20) - 1);
    printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
}
