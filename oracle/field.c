/*
 * Oracle: Goldilocks field arithmetic and divisor helpers.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates ntt/src/field/algo/generic.rs, ntt/src/field/algo/mod.rs:9-20,
 * ntt/src/field/mod.rs:42-50 and ntt/src/divisors.rs.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

typedef unsigned __int128 u128;

static const uint64_t P = OR_P;

/* generic.rs:17-26: add via a - (p - b) with a 2^32-1 borrow fix-up. */
uint64_t or_add(uint64_t a, uint64_t b) {
    uint64_t nb = P - b;
    uint64_t r = a - nb;
    uint32_t adj = (uint32_t)0 - (uint32_t)(a < nb);
    return r - (uint64_t)adj;
}

/* generic.rs:31-39 */
uint64_t or_sub(uint64_t a, uint64_t b) {
    uint64_t r = a - b;
    uint32_t adj = (uint32_t)0 - (uint32_t)(a < b);
    return r - (uint64_t)adj;
}

/* generic.rs:77-108: reduce low + mid*2^64 + high*2^96 using
 * 2^64 = 2^32 - 1 and 2^96 = -1 (mod p). */
static uint64_t reduce_159(uint64_t low, uint32_t mid, uint64_t high) {
    uint64_t low2 = low - high;
    if (low < high) low2 += P;
    uint64_t product = (uint64_t)mid << 32;
    product -= product >> 32;
    uint64_t result = product + low2;
    if (result < product) result -= P;
    if (result >= P) result -= P;
    return result;
}

/* generic.rs:69-71 */
static uint64_t reduce_128(u128 n) {
    return reduce_159((uint64_t)n, (uint32_t)(n >> 64), (uint64_t)(n >> 96));
}

/* generic.rs:44-48 */
uint64_t or_mul(uint64_t a, uint64_t b) { return reduce_128((u128)a * b); }

/* generic.rs:114-124 (Montgomery reduction, R = 2^64) */
static uint64_t mont_reduce_128(u128 x) {
    uint64_t x0 = (uint64_t)x, x1 = (uint64_t)(x >> 64);
    uint64_t a = x0 + (x0 << 32);
    uint64_t e = a < x0;
    uint64_t b = a - (a >> 32) - e;
    uint64_t r = x1 - b;
    uint32_t adj = (uint32_t)0 - (uint32_t)(x1 < b);
    return r - (uint64_t)adj;
}

uint64_t or_mont_mul(uint64_t a, uint64_t b) { return mont_reduce_128((u128)a * b); }

uint64_t or_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = or_mul(r, a);
        a = or_mul(a, a);
        e >>= 1;
    }
    return r;
}

/* generic.rs:171-193: a * 2^s using 2^96 = -1, 2^192 = 1. */
uint64_t or_shift(uint64_t a, uint64_t n) {
    uint64_t q = n / 96, r = n % 96;
    if (q & 1) a = or_sub(0, a);
    if (r == 0) return a;
    if (r <= 32) return reduce_159(a << r, (uint32_t)(a >> (64 - r)), 0);
    if (r < 64) return reduce_159(a << r, (uint32_t)(a >> (64 - r)), a >> (96 - r));
    return reduce_159(0, (uint32_t)(a << (r - 64)), a >> (96 - r));
}

/* generic.rs:198-208: a * w_384^i with w_384 = 2^24 - 2^72 (a square root
 * of 2), so odd powers are a sum of two shifts. */
uint64_t or_root_384(uint64_t a, uint64_t i) {
    if ((i & 1) == 0) return or_shift(a, i >> 1);
    i >>= 1;
    return or_add(or_shift(a, i + 24), or_shift(a, i + 168));
}

static int ctz64(uint64_t x) { return __builtin_ctzll(x); }

/* generic.rs:220-271: binary half-extended gcd; the accumulated power of two
 * is removed with a shift by 191*k (2 has order 192). */
uint64_t or_inv(uint64_t a) {
    if (a == 0) return 0;
    uint64_t u = P, v = a;
    uint64_t t0 = 0, t1 = 1;
    uint64_t twos = (uint64_t)ctz64(v);
    v >>= twos;
    twos += 96;
    for (;;) {
        u -= v;
        t0 += t1;
        int count = ctz64(u);
        u >>= count;
        t1 <<= count;
        twos += (uint64_t)count;
        if (u < v) {
            uint64_t t = u; u = v; v = t;
            t = t0; t0 = t1; t1 = t;
            twos += 96;
        }
        if (u == v) break;
    }
    twos *= 191;
    return or_shift(t0, twos);
}

/* field/mod.rs:42-50 (canonical value of the Montgomery-stored root). */
uint64_t or_root(uint64_t order) {
    if (order == 0 || (P - 1) % order != 0) return 0;
    return or_pow(OR_GENERATOR, (P - 1) / order);
}

/* ---- divisors.rs ---- */
static const uint64_t ORDER = 0xffffffff00000000ULL;
static const uint64_t SMOOTH_ORDER = 64424509440ULL;
static const uint64_t FACTORS[5] = {3, 5, 17, 257, 65537};

int or_is_divisor(uint64_t n) { return n > 0 && ORDER % n == 0; }
int or_is_smooth(uint64_t n) { return n > 0 && SMOOTH_ORDER % n == 0; }

static int cmp_u64(const void *x, const void *y) {
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return (a > b) - (a < b);
}

static uint64_t DIVS[1056];
static size_t NDIVS = 0;

static void init_divisors(void) {
    if (NDIVS) return;
    uint64_t d[1056];
    size_t n = 0;
    for (int i = 0; i <= 32; i++) d[n++] = 1ULL << i;
    for (int f = 0; f < 5; f++) {
        size_t m = n;
        for (size_t i = 0; i < m; i++) d[n++] = d[i] * FACTORS[f];
    }
    qsort(d, n, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < n; i++) DIVS[i] = d[i];
    NDIVS = n;
}

size_t or_divisors(uint64_t *out) {
    init_divisors();
    if (out)
        for (size_t i = 0; i < NDIVS; i++) out[i] = DIVS[i];
    return NDIVS;
}

/* index of the first divisor >= n (binary search insertion point) */
static size_t lower_bound(uint64_t n) {
    size_t lo = 0, hi = NDIVS;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (DIVS[mid] < n) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* divisors.rs:52-59; returns 0 for n == 0 and for n above the group order. */
uint64_t or_next(uint64_t n) {
    if (n == 0) return 0;
    init_divisors();
    size_t i = lower_bound(n);
    return i < NDIVS ? DIVS[i] : 0;
}

/* divisors.rs:67-89: largest divisor d <= ~sqrt(n) with d | n. */
uint64_t or_split(uint64_t n) {
    if (!or_is_divisor(n)) return 0;
    init_divisors();
    uint64_t isqrt = (uint64_t)sqrt((double)n);
    isqrt = (isqrt + n / isqrt) / 2;
    size_t i = lower_bound(isqrt);
    if (i >= NDIVS) i = NDIVS - 1;
    for (size_t j = i + 1; j-- > 0;) {
        uint64_t d = DIVS[j];
        if (n % d == 0 && n / d >= d) return d;
    }
    return 1;
}
