// Host build of the modular-inverse partner of the shuffle passes
// (tl_modinv.hpp) over the HIP shim: reads "x m r" triples, one a line, and
// prints each partner, so that a test can check moduli up to 2^40, far past
// any array a test permutes.
#include "hip/hip_runtime.h"
#include "tl_modinv.hpp"
#include <stdio.h>

int main() {
    unsigned long long x, m, r;
    while (scanf("%llu %llu %llu", &x, &m, &r) == 3)
        printf("%llu\n", (unsigned long long)tl::mod_partner(x, m, r));
    return 0;
}
