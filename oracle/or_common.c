/*
 * or_common.c -- oracle helpers that do not depend on the key type.
 * TEST INFRASTRUCTURE ONLY (see tl_oracle.h).
 */
#include "tl_oracle.h"

#include <stdlib.h>
#include <string.h>

uint64_t or_ipow(uint64_t base, uint32_t e) {
    uint64_t r = 1;
    while (e--) r *= base;
    return r;
}

/* floor(log_base n) in integer arithmetic.  The reference computes it in
 * double precision as log10(n)/log10(b+1) (cpu/src/btree-cycles.h:75) or
 * log2(n) (cpu/src/bst-hybrid.h:124). */
uint32_t or_ilog(uint64_t n, uint64_t base) {
    uint32_t h = 0;
    uint64_t p = 1;
    while (p <= n / base) {
        p *= base;
        h++;
    }
    return h;
}

static uint32_t floor_log2(uint64_t n) { return 63u - (uint32_t)__builtin_clzll(n); }

/* Split of a depth-d tree into a top tree of depth root_d and bottom trees of
 * depth leaf_d, with leaf_d rounded up to a power of two
 * (gpu/src/mveb-cycles.cuh:240-247, gpu/src/query-mveb.cu:28-35). */
void or_mveb_split(uint32_t d, uint32_t *root_d, uint32_t *leaf_d) {
    uint32_t r = (d - 2) / 2 + 1;
    uint32_t l = d - r;
    if (l & (l - 1)) {
        uint32_t p = 1;
        while (p < l) p <<= 1;
        l = p;
        r = d - l;
    }
    *root_d = r;
    *leaf_d = l;
}

/* The classic split of the CPU vEB: top tree of floor(d/2) levels, bottom
 * trees of ceil(d/2) (cpu/src/veb-cycles.h:36-50,114-115,
 * cpu/src/query-veb.cpp:28-29). */
void or_classic_split(uint32_t d, uint32_t *root_d, uint32_t *leaf_d) {
    *root_d = (d - 2) / 2 + 1;
    *leaf_d = d - *root_d;
}

typedef void (*split_fn)(uint32_t, uint32_t *, uint32_t *);

/* ------------------------------------------------------------------------ */
/* In-order DFS fill of the heap-indexed node tree.  Node u owns the slots
 * [u*b, min(u*b+b, n)); its children are nodes u*(b+1)+1+j, j = 0..b.  The
 * visit order is child 0, key 0, child 1, key 1, ..., key b-1, child b.
 * Recursion depth is the node-tree height (<= 64). */
static void inorder_rec(uint64_t u, uint64_t n, uint64_t b, uint64_t nodes, uint64_t *pos,
                        uint64_t *rank) {
    uint64_t first = u * b;
    uint64_t cnt = (n - first < b) ? n - first : b;
    for (uint64_t j = 0; j <= cnt; j++) {
        uint64_t child = u * (b + 1) + 1 + j;
        if (child < nodes) inorder_rec(child, n, b, nodes, pos, rank);
        if (j < cnt) pos[(*rank)++] = first + j;
    }
}

int or_inorder_btree(uint64_t n, uint64_t b, uint64_t *pos_of_rank) {
    if (b == 0) return -1;
    if (n == 0) return 0;
    uint64_t rank = 0;
    inorder_rec(0, n, b, (n + b - 1) / b, pos_of_rank, &rank);
    return rank == n ? 0 : -2;
}

/* Closed form of the same map.  Non-perfect trees are first reduced to their
 * perfect internal part the way the reference does it
 * (cpu/src/btree-cycles.h:29-36, 71-91): the first l+i ranks alternate b leaf
 * keys and one separator, then come the keys of the partial leaf, then the
 * remaining internal keys.  For a perfect tree of (b+1)^d - 1 keys the 1-based
 * rank r = (u, a, 0^s) in base b+1 (a != 0) lives on level t = d-1-s at slot
 * (b+1)^t - 1 + u*b + a - 1. */
uint64_t or_btree_pos(uint64_t x, uint64_t n, uint64_t b) {
    uint64_t k1 = b + 1;
    uint32_t h = or_ilog(n, k1);
    uint64_t P = or_ipow(k1, h); /* (b+1)^h <= n */
    int perfect = (n / k1 >= P - 1) && (n == P * k1 - 1);
    if (!perfect) {
        uint64_t nint = P - 1, nleaf = n - nint;
        uint64_t rem = nleaf % b, l = nleaf - rem, i = l / b;
        uint64_t span = l + i;
        if (x < span) {
            uint64_t row = x / k1, c = x % k1;
            if (c < b) return nint + row * b + c;
            x = row; /* the row-th separator is internal rank row */
        } else if (x < span + rem) {
            return nint + l + (x - span);
        } else {
            x = i + (x - span - rem);
        }
        n = nint;
        if (n == 0) return 0;
    }
    uint64_t r = x + 1, div = 1;
    while (r % k1 == 0) {
        r /= k1;
        div *= k1;
    }
    uint64_t a = r % k1, u = r / k1;
    uint64_t level_start = (n + 1) / (div * k1) - 1; /* (b+1)^t - 1 */
    return level_start + u * b + a - 1;
}

/* Structural definition of the modified vEB of n keys.  Perfect n = 2^d - 1:
 * top tree (r = 2^root_d - 1 keys) first, then the r+1 bottom trees of
 * l = 2^leaf_d - 1 keys, each laid out recursively.  Non-perfect: the top
 * tree, num_full = (n-r)/l full bottom trees, then the incomplete bottom tree
 * of inc_n keys laid out recursively as its own tree; the top keys whose
 * bottom trees are empty are the largest ranks
 * (gpu/src/mveb-cycles.cuh:299-414). */
static uint64_t veb_pos(uint64_t x, uint64_t n, split_fn split) {
    uint64_t base = 0;
    for (;;) {
        if (n <= 1) return base;
        uint32_t d = floor_log2(n) + 1, rd, ld;
        split(d, &rd, &ld);
        uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
        uint64_t full = (d >= 64) ? ~0ull : (1ull << d) - 1;
        uint64_t nf, inc;
        if (n == full) {
            nf = r + 1; /* the last bottom tree sits after the last top key */
            inc = 0;
        } else {
            nf = (n - r) / l;
            inc = n - r - nf * l;
        }
        uint64_t span = nf * (l + 1);
        if (n == full) span = r * (l + 1) + l;
        if (x < span) {
            uint64_t k = x / (l + 1), c = x % (l + 1);
            if (c == l) { /* top key k */
                n = r;
                x = k;
                continue;
            }
            base += r + k * l;
            n = l;
            x = c;
            continue;
        }
        if (x < span + inc) {
            base += r + nf * l;
            x -= span;
            n = inc;
            continue;
        }
        x = nf + (x - span - inc); /* top keys without a bottom tree */
        n = r;
    }
}

uint64_t or_mveb_pos(uint64_t x, uint64_t n) { return veb_pos(x, n, or_mveb_split); }

/* the same structure with the classic split: the CPU vEB layout
 * (cpu/src/veb-cycles.h:34-160) */
uint64_t or_classic_pos(uint64_t x, uint64_t n) { return veb_pos(x, n, or_classic_split); }

int or_mveb_layout(uint64_t n, uint64_t *pos_of_rank) {
    for (uint64_t x = 0; x < n; x++) pos_of_rank[x] = or_mveb_pos(x, n);
    return 0;
}

int or_classic_layout(uint64_t n, uint64_t *pos_of_rank) {
    for (uint64_t x = 0; x < n; x++) pos_of_rank[x] = or_classic_pos(x, n);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Brodal-Fagerberg-Jacob position table with the mvEB split
 * (gpu/src/query-mveb.cu:25-47).  Entry t describes the bottom trees that
 * start at depth t: their size L, the size R of the top tree above them and
 * the depth D of that top tree's root. */
static void build_table(or_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth,
                        split_fn split) {
    if (n <= 1) return;
    uint32_t rd, ld;
    split(d, &rd, &ld);
    uint64_t root_n = (1ull << rd) - 1;
    uint64_t leaf_n = (1ull << ld) - 1;
    uint32_t i = d - ld;
    table[i].L = leaf_n;
    table[i].R = root_n;
    table[i].D = root_depth;
    build_table(table, root_n, rd, root_depth, split);
    build_table(&table[i], leaf_n, ld, i + root_depth, split);
}

void or_build_table(or_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth) {
    build_table(table, n, d, root_depth, or_mveb_split);
}

/* the classic rule (cpu/src/query-veb.cpp:25-41) */
void or_build_table_classic(or_veb_table *table, uint64_t n, uint32_t d, uint32_t root_depth) {
    build_table(table, n, d, root_depth, or_classic_split);
}

/* Table set for a non-perfect mvEB: one table per nested incomplete bottom
 * tree.  idx[j] is where structure j ends (the start of the incomplete tree
 * it contains), the last entry is n; table_idx[j] is the offset of table j in
 * the packed table array (gpu/test/query-mveb-fixedN.cu:42-165). */
static int tableset(uint64_t n, uint32_t *d_out, uint32_t *num_tables, uint32_t *table_size,
                    or_veb_table *tables, uint64_t *idx, uint32_t *table_idx, split_fn split) {
    if (n == 0) return -1;
    uint32_t d = floor_log2(n) + 1;
    *d_out = d;
    uint64_t full = (1ull << d) - 1;
    if (n == full) {
        memset(tables, 0, sizeof(or_veb_table) * d);
        build_table(tables, n, d, 0, split);
        *num_tables = 1;
        *table_size = d;
        idx[0] = n;
        table_idx[0] = 0;
        return 0;
    }
    /* walk the chain of incomplete bottom trees */
    uint32_t nt = 1, tsize = d;
    uint64_t sizes[64];
    uint32_t depths[64];
    sizes[0] = n;
    depths[0] = d;
    uint64_t tn = n;
    uint32_t td = d;
    while (td >= 2) {
        uint32_t rd, ld;
        split(td, &rd, &ld);
        uint64_t r = (1ull << rd) - 1, l = (1ull << ld) - 1;
        uint64_t nf = (tn - r) / l;
        uint64_t inc = tn - r - nf * l;
        if (inc == 0) break;
        uint32_t incd = floor_log2(inc) + 1;
        sizes[nt] = inc;
        depths[nt] = incd;
        nt++;
        tsize += incd;
        tn = inc;
        td = incd;
        if (inc == (1ull << incd) - 1) break;
    }
    memset(tables, 0, sizeof(or_veb_table) * tsize);
    uint32_t off = 0;
    for (uint32_t j = 0; j < nt; j++) {
        table_idx[j] = off;
        build_table(&tables[off], sizes[j], depths[j], 0, split);
        off += depths[j];
        idx[j] = (j + 1 < nt) ? n - sizes[j + 1] : n;
    }
    *num_tables = nt;
    *table_size = tsize;
    return 0;
}

int or_veb_tableset(uint64_t n, uint32_t *d_out, uint32_t *num_tables, uint32_t *table_size,
                    or_veb_table *tables, uint64_t *idx, uint32_t *table_idx) {
    return tableset(n, d_out, num_tables, table_size, tables, idx, table_idx, or_mveb_split);
}

/* the classic vEB's table set, built the same way with the classic rule
 * (cpu/test/query-veb-fixedN.cpp:37-158, where it is an array of pointers) */
int or_classic_tableset(uint64_t n, uint32_t *d_out, uint32_t *num_tables, uint32_t *table_size,
                        or_veb_table *tables, uint64_t *idx, uint32_t *table_idx) {
    return tableset(n, d_out, num_tables, table_size, tables, idx, table_idx, or_classic_split);
}

/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t seed, uint64_t i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* ---- involution helpers (cpu/src/involutions.cpp) ----------------------- */

/* cpu/src/involutions.cpp:44-63: digits taken from the top (power k down to
 * 1) and laid down from the bottom */
uint64_t or_rev_base_d(uint64_t x, uint64_t base, uint64_t d) {
    if (d == 0) return x;
    uint64_t r = 0, j = 1, k = or_ipow(base, (uint32_t)(d - 1));
    for (uint64_t i = d - 1; i > 0; --i) {
        const uint64_t digit = x / k;
        r += digit * j;
        j *= base;
        x -= digit * k;
        k /= base;
    }
    return r + x * j;
}

/* cpu/src/involutions.cpp:66-80: the top d - b digits kept in place */
uint64_t or_rev_base_b(uint64_t x, uint64_t base, uint64_t d, uint64_t b) {
    uint64_t r = 0, k = or_ipow(base, (uint32_t)(d - 1));
    for (uint64_t i = d - 1; i + 1 > b; --i) {   /* i >= b without wrapping at b = 0 */
        const uint64_t digit = x / k;
        r += digit * k;
        x -= digit * k;
        k /= base;
        if (i == 0) break;
    }
    return r + or_rev_base_d(x, base, b);
}

/* cpu/src/involutions.cpp:83-93 */
static uint64_t or_gcd(uint64_t a, uint64_t b) {
    while (a) {
        const uint64_t c = a;
        a = b % a;
        b = c;
    }
    return b;
}

/* cpu/src/involutions.cpp:96-109, recursive: returns g, with *x a coefficient
 * of a (a * x + b * y = g) */
static uint64_t or_egcd(uint64_t a, uint64_t b, int64_t *x, int64_t *y) {
    if (a == 0) {
        *x = 0;
        *y = 1;
        return b;
    }
    int64_t x1, y1;
    const uint64_t g = or_egcd(b % a, a, &x1, &y1);
    *x = y1 - (int64_t)(b / a) * x1;
    *y = x1;
    return g;
}

/* cpu/src/involutions.cpp:112-121 */
uint64_t or_involution(uint64_t r, uint64_t x, uint64_t m) {
    const uint64_t g = or_gcd(x, m);
    int64_t inv, y;
    or_egcd(x / g, m / g, &inv, &y);
    if (inv < 0) inv += (int64_t)(m / g);
    /* exact where the reference's 64-bit product r * inv would wrap (2^64 and
       more; no test size reaches it) */
    return g * (uint64_t)((unsigned __int128)r * (uint64_t)inv % (m / g));
}
