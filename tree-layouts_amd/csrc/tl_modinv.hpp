// tl_modinv.hpp -- the modular-inverse partner of the shuffle passes
// (`involution` of gpu/src/involutions.cu:106-113), kept apart from the
// kernels so that a host build can check it at moduli no test array reaches.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tl {

// quotient of a / b by 32-bit division when both fit
__device__ inline uint64_t udiv(uint64_t a, uint64_t b) {
    if (((a | b) >> 32) == 0) return (uint32_t)a / (uint32_t)b;
    return a / b;
}

// a * b mod m for a, b < m < 2^40, exact: below 2^32 the product fits 64 bits;
// above, b is split at bit 20 so that every partial product stays below 2^60
__device__ inline uint64_t mulmod40(uint64_t a, uint64_t b, uint64_t m) {
    if (m <= (1ull << 32)) {
        const uint64_t v = a * b;
        return v - udiv(v, m) * m;
    }
    uint64_t t = (a * (b >> 20)) % m;
    t = (t << 20) % m;
    return (t + (a * (b & 0xFFFFFull)) % m) % m;
}

// the partner of x in [1, m), m < 2^40, under the involution
// x -> g * (r * inv(x / g) mod (m / g)), g = gcd(x, m), inv the inverse modulo
// m / g.  One extended-Euclid run gives both g and the inverse (it keeps
// a = s_a * x mod m for both remainders), where the reference runs gcd and
// egcd separately; the product is taken modulo m / g exactly, where the
// reference's 64-bit r * inv would wrap once it reaches 2^64.
__device__ inline uint64_t mod_partner(uint64_t x, uint64_t m, uint64_t r) {
    if (m < (1ull << 32)) {
        // all in 32 bits: the coefficients of x are kept modulo 2^32; the one
        // returned, s with s * x = g (mod m), has |s| <= m / (2g) < 2^31, so
        // its 32-bit pattern read as signed is exact (only the coefficient of
        // the last step, +-m / g, can leave that range, and it is dropped)
        uint32_t a = (uint32_t)m, b = (uint32_t)x, sa = 0, sb = 1;
        while (b) {
            const uint32_t q = a / b;
            const uint32_t t = a - q * b;
            a = b;
            b = t;
            const uint32_t st = sa - q * sb;
            sa = sb;
            sb = st;
        }
        const uint32_t mg = (uint32_t)m / a;
        int64_t inv = (int64_t)(int32_t)sa;
        if (inv < 0) inv += (int64_t)mg;
        return (uint64_t)a * mulmod40((uint64_t)inv, r % mg, mg);
    }
    uint64_t a = m, b = x;
    int64_t sa = 0, sb = 1;
    while (b) {
        const uint64_t q = udiv(a, b);
        const uint64_t t = a - q * b;
        a = b;
        b = t;
        const int64_t st = sa - (int64_t)q * sb;
        sa = sb;
        sb = st;
    }
    // a = g; sa * (x / g) = 1 (mod m / g)
    const uint64_t mg = udiv(m, a);
    int64_t inv = sa % (int64_t)mg;
    if (inv < 0) inv += (int64_t)mg;
    return a * mulmod40((uint64_t)inv, r % mg, mg);
}

}  // namespace tl
