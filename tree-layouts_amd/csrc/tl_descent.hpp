// tl_descent.hpp -- B-tree node descent by a group of G lanes (G a power of
// two >= b, at most 64), shared by the plain and the learned-start queries.
//
// Node at key index i holds keys A[i .. i+b-1]; lane j of the group loads key
// j, two ballots give (keys < q) and (keys == q) for the whole group, the
// popcount of the first is the branch and a set bit of the second at that
// position is the hit (gpu/src/query-btree.cuh:32-58; child of branch js is
// (i+1)(b+1) + js*b - 1).  Falling off the tree answers n.
//
// descent2 walks two independent queries per group in lock step, so every
// lane has two node loads in flight.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tl {

template <int G>
__device__ inline uint64_t group_bits(uint64_t ballot, uint32_t g) {
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1);
    return (ballot >> (g * G)) & gmask;
}

// one level of one query: branch, or record the hit; falling off ends it
__device__ inline void descent_step(uint64_t mlt, uint64_t meq, uint32_t b, uint64_t n,
                                    uint64_t &i, bool &active, uint64_t &res, bool &found) {
    const uint32_t js = (uint32_t)__popcll(mlt);
    if ((meq >> js) & 1ull) {
        res = i + js;
        found = true;
        active = false;
    } else {
        i = (i + 1) * (b + 1) + (uint64_t)js * b - 1;
        if (i >= n) active = false;
    }
}

template <typename K, int G>
__device__ inline void descent2(const K *__restrict__ A, uint64_t n, uint32_t b, uint32_t g,
                                uint32_t j, K q0, K q1, uint64_t i0, uint64_t i1, bool a0, bool a1,
                                uint64_t &r0, uint64_t &r1, bool &f0, bool &f1) {
    r0 = n;
    r1 = n;
    f0 = false;
    f1 = false;
    while (__any(a0 || a1)) {
        K k0{}, k1{};
        const bool l0 = a0 && j < b && i0 + j < n;
        const bool l1 = a1 && j < b && i1 + j < n;
        if (l0) k0 = A[i0 + j];
        if (l1) k1 = A[i1 + j];
        const uint64_t lt0 = group_bits<G>(__ballot(l0 && k0 < q0), g);
        const uint64_t eq0 = group_bits<G>(__ballot(l0 && k0 == q0), g);
        const uint64_t lt1 = group_bits<G>(__ballot(l1 && k1 < q1), g);
        const uint64_t eq1 = group_bits<G>(__ballot(l1 && k1 == q1), g);
        if (a0) descent_step(lt0, eq0, b, n, i0, a0, r0, f0);
        if (a1) descent_step(lt1, eq1, b, n, i1, a1, r1, f1);
    }
}

}   // namespace tl
