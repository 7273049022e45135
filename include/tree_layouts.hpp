// tree_layouts.hpp -- the reference's C++ surface over the gfx950 C ABI.
//
// Same names, argument order and meaning as the reference's GPU headers
// (README.md:31-67; gpu/src/bst-hybrid.cuh:120, gpu/src/btree-cycles.cuh:77,
// gpu/src/mveb-cycles.cuh:420, gpu/src/query-bst.cuh:57,
// gpu/src/query-btree.cuh:76, gpu/src/query-mveb.cuh:86,231,
// gpu/src/query-mveb.cu:25, gpu/src/binary-search.cuh:63), so a reference driver switches by changing its
// includes.  Differences, all deliberate:
//   * one header for every layout (the reference needs one translation unit
//     per layout header because their helper templates collide);
//   * the timeQuery* functions draw their hit queries from a seeded
//     counter-based generator (an extra `seed` argument, default 0) instead of
//     rand() seeded with time(NULL), and they also compute a miss answer (n)
//     for absent keys;
//   * failures are reported on stderr and the functions return -1 ms (the
//     reference has no error path at all).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "tree_layouts.h"

typedef tl_veb_table vEB_table;

// timed runs per measurement in the reference's drivers (gpu/src/params.cuh:21)
#ifndef ITERS
#define ITERS 10
#endif

namespace tl_detail {
template <typename T> struct dtype_of;
template <> struct dtype_of<uint32_t> { static constexpr tl_dtype v = TL_U32; };
template <> struct dtype_of<uint64_t> { static constexpr tl_dtype v = TL_U64; };
template <> struct dtype_of<int32_t> { static constexpr tl_dtype v = TL_I32; };
template <> struct dtype_of<int64_t> { static constexpr tl_dtype v = TL_I64; };

inline bool ok(tl_status s, const char *what) {
    if (s == TL_OK) return true;
    fprintf(stderr, "%s failed (%d): %s\n", what, (int)s, tl_last_error());
    return false;
}

inline uint64_t mix(uint64_t seed, uint64_t i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// the reference's createRandomQueries (gpu/src/common.cuh:44-60): hits drawn
// from the host copy of the layout
template <typename T>
std::vector<T> random_hits(const T *A, uint64_t n, uint64_t q, uint64_t seed) {
    std::vector<T> out(q);
    for (uint64_t i = 0; i < q; i++) out[i] = A[mix(seed, i) % n];
    return out;
}

// Optional answer sink for tests: when set, the next timeQuery* call copies
// its uint64 answers (slot of each query, n for a miss) into it.  The
// reference's functions keep no answers (they overwrite a buffer they free).
inline std::vector<uint64_t> *&answer_sink() {
    static thread_local std::vector<uint64_t> *sink = nullptr;
    return sink;
}

// mean and (population) standard deviation of x[0..k)
inline void mean_sd(const double *x, int k, double &mean, double &sd) {
    mean = 0.;
    for (int i = 0; i < k; i++) mean += x[i];
    mean /= k;
    sd = 0.;
    for (int i = 0; i < k; i++) sd += (x[i] - mean) * (x[i] - mean);
    sd = sqrt(sd / k);
}

// uploads A and the queries, runs `search`, returns its kernel time
template <typename T, typename F>
float run_queries(T *A, T *dev_A, uint64_t n, uint64_t q, uint64_t seed, F search) {
    std::vector<T> hq = random_hits(A, n, q, seed);
    T *dq = nullptr;
    uint64_t *dans = nullptr;
    if (hipMalloc(&dq, q * sizeof(T)) != hipSuccess) return -1.f;
    if (hipMalloc(&dans, q * sizeof(uint64_t)) != hipSuccess) {
        (void)hipFree(dq);
        return -1.f;
    }
    float ms = -1.f;
    if (hipMemcpy(dev_A, A, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(dq, hq.data(), q * sizeof(T), hipMemcpyHostToDevice) == hipSuccess) {
        float t = 0.f;
        if (search(dq, dans, &t)) {
            ms = t;
            if (std::vector<uint64_t> *sink = answer_sink()) {
                sink->resize(q);
                if (hipMemcpy(sink->data(), dans, q * sizeof(uint64_t), hipMemcpyDeviceToHost) !=
                    hipSuccess)
                    ms = -1.f;
            }
        }
    }
    (void)hipFree(dq);
    (void)hipFree(dans);
    return ms;
}
}   // namespace tl_detail

// ---- permutations (in place on dev_A, synchronous, return milliseconds) ----

template <typename T>
float timePermuteBST(T *dev_A, uint64_t n) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_bst(dev_A, n, tl_detail::dtype_of<T>::v, nullptr, &ms),
                       "timePermuteBST"))
        return -1.f;
    return ms;
}

template <typename T>
double timePermuteBtree(T *dev_A, uint64_t n, uint64_t b) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_btree(dev_A, n, b, tl_detail::dtype_of<T>::v, nullptr, &ms),
                       "timePermuteBtree"))
        return -1.0;
    return ms;
}

template <typename T>
double timePermutevEB(T *dev_A, uint64_t n) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_mveb(dev_A, n, tl_detail::dtype_of<T>::v, nullptr, &ms),
                       "timePermutevEB"))
        return -1.0;
    return ms;
}

// The involution programs define the same three names
// (gpu/src/bst-involutions.cuh:167, gpu/src/btree-involutions.cuh:151,
// gpu/src/mveb-involutions.cuh:503); a driver built on them calls these.
namespace involutions {
template <typename T>
float timePermuteBST(T *dev_A, uint64_t n) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_bst_involutions(dev_A, n, tl_detail::dtype_of<T>::v, nullptr,
                                                  &ms),
                       "involutions::timePermuteBST"))
        return -1.f;
    return ms;
}

template <typename T>
double timePermuteBtree(T *dev_A, uint64_t n, uint64_t b) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_btree_involutions(dev_A, n, b, tl_detail::dtype_of<T>::v,
                                                    nullptr, &ms),
                       "involutions::timePermuteBtree"))
        return -1.0;
    return ms;
}

template <typename T>
double timePermutevEB(T *dev_A, uint64_t n) {
    float ms = 0.f;
    if (!tl_detail::ok(tl_permute_mveb_involutions(dev_A, n, tl_detail::dtype_of<T>::v, nullptr,
                                                   &ms),
                       "involutions::timePermutevEB"))
        return -1.0;
    return ms;
}
}   // namespace involutions

// ---- queries (hits drawn from the host layout A; kernel time in ms) ---------

template <typename T>
float timeQueryBST(T *A, T *dev_A, uint64_t n, uint64_t q, uint64_t seed = 0) {
    return tl_detail::run_queries(A, dev_A, n, q, seed, [&](T *dq, uint64_t *dans, float *t) {
        return tl_detail::ok(
            tl_query_bst(dev_A, n, tl_detail::dtype_of<T>::v, dq, q, dans, nullptr, t),
            "timeQueryBST");
    });
}

template <typename T>
float timeQueryBtree(T *A, T *dev_A, uint64_t n, uint64_t b, uint64_t q, uint64_t seed = 0) {
    return tl_detail::run_queries(A, dev_A, n, q, seed, [&](T *dq, uint64_t *dans, float *t) {
        return tl_detail::ok(
            tl_query_btree(dev_A, n, b, tl_detail::dtype_of<T>::v, dq, q, dans, nullptr, t),
            "timeQueryBtree");
    });
}

// binary search of the sorted array (gpu/src/binary-search.cuh:63): here A
// and dev_A hold the sorted keys, not a layout
template <typename T>
float timeQuery(T *A, T *dev_A, uint64_t n, uint64_t q, uint64_t seed = 0) {
    return tl_detail::run_queries(A, dev_A, n, q, seed, [&](T *dq, uint64_t *dans, float *t) {
        return tl_detail::ok(
            tl_query_binary(dev_A, n, tl_detail::dtype_of<T>::v, dq, q, dans, nullptr, t),
            "timeQuery");
    });
}

template <typename T>
float timeQueryvEB(T *A, T *dev_A, vEB_table *dev_table, uint64_t n, uint32_t d, uint64_t q,
                   uint64_t seed = 0) {
    return tl_detail::run_queries(A, dev_A, n, q, seed, [&](T *dq, uint64_t *dans, float *t) {
        return tl_detail::ok(tl_query_mveb(dev_A, dev_table, n, d, tl_detail::dtype_of<T>::v, dq,
                                           q, dans, nullptr, t),
                             "timeQueryvEB");
    });
}

template <typename T>
float timeQueryvEB_nonperfect(T *A, T *dev_A, vEB_table *dev_tables, uint64_t *dev_idx,
                              uint32_t *dev_table_idx, uint64_t n, uint32_t d,
                              uint32_t num_tables, uint64_t q, uint64_t seed = 0) {
    return tl_detail::run_queries(A, dev_A, n, q, seed, [&](T *dq, uint64_t *dans, float *t) {
        return tl_detail::ok(
            tl_query_mveb_nonperfect(dev_A, dev_tables, dev_idx, dev_table_idx, n, d, num_tables,
                                     tl_detail::dtype_of<T>::v, dq, q, dans, nullptr, t),
            "timeQueryvEB_nonperfect");
    });
}

// ---- mvEB table (gpu/src/query-mveb.cu:25) ----------------------------------

inline void buildTable(vEB_table *table, uint64_t n, uint32_t d, uint32_t root_depth) {
    tl_detail::ok(tl_veb_build_table(table, n, d, root_depth), "buildTable");
}

// ---- driver helpers (gpu/src/common.cuh:33-40, gpu/src/common.cu:41-93) ------

// the reference's sorted input: A[i] = i + 1
template <typename T>
void initSortedList(T *A, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) A[i] = (T)(i + 1);
}

// "n mean_ms sd_ms mean_throughput sd_throughput t_1 .. t_ITERS", throughput
// in keys per microsecond, as the reference's printTimings prints it
inline void printTimings(uint64_t n, double time[ITERS]) {
    double tp[ITERS], avg, sd, tp_avg, tp_sd;
    for (int i = 0; i < ITERS; i++) tp[i] = n / (1000. * time[i]);
    tl_detail::mean_sd(time, ITERS, avg, sd);
    // the mean throughput is n over the mean time; its spread is taken about it
    tp_avg = n / (1000. * avg);
    tp_sd = 0.;
    for (int i = 0; i < ITERS; i++) tp_sd += (tp[i] - tp_avg) * (tp[i] - tp_avg);
    tp_sd = sqrt(tp_sd / ITERS);
    printf("%llu %lf %lf %lf %lf ", (unsigned long long)n, avg, sd, tp_avg, tp_sd);
    for (int i = 0; i < ITERS; i++) printf(" %lf", time[i]);
    printf("\n");
}

// "n q mean_ms sd_ms mean_throughput sd_throughput t_1 .. t_ITERS", queries
// per microsecond (printQueryTimings)
inline void printQueryTimings(uint64_t n, uint64_t q, double time[ITERS]) {
    double tp[ITERS], avg, sd, tp_avg, tp_sd;
    for (int i = 0; i < ITERS; i++) tp[i] = q / (1000. * time[i]);
    tl_detail::mean_sd(time, ITERS, avg, sd);
    // the mean throughput is q over the mean time; its spread is taken about it
    tp_avg = q / (1000. * avg);
    tp_sd = 0.;
    for (int i = 0; i < ITERS; i++) tp_sd += (tp[i] - tp_avg) * (tp[i] - tp_avg);
    tp_sd = sqrt(tp_sd / ITERS);
    printf("%llu %llu %lf %lf %lf %lf ", (unsigned long long)n, (unsigned long long)q, avg, sd,
           tp_avg, tp_sd);
    for (int i = 0; i < ITERS; i++) printf(" %lf", time[i]);
    printf("\n");
}
