// Goldilocks field arithmetic (p = 2^64 - 2^32 + 1) for gfx950 and the host.
//
// Elements are raw u64 words in [0, p).  The reference stores Montgomery
// representations (ntt/src/field/mod.rs:8-27); every operation used by the
// transforms is linear or a product with a canonical constant, so standard
// arithmetic on the raw words gives the reference's raw output words.
//
// The reduction uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod p), the identities
// behind reduce_159 (ntt/src/field/algo/generic.rs:77-108).  Products are
// built from 32x32->64 partial products so the compiler can use
// v_mad_u64_u32; no MFMA is involved (this is integer work).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GL_HD __host__ __device__ __forceinline__

namespace gl {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u64 P = 0xffffffff00000001ull;
constexpr u64 EPS = 0xffffffffull;  // 2^64 mod p = 2^32 - 1

// a + b for canonical a, b (generic.rs:17-26 form: a - (p - b), fix the
// borrow by adding p, i.e. subtracting 2^32 - 1 modulo 2^64).
GL_HD u64 add(u64 a, u64 b) {
    u64 nb = P - b;
    u64 r = a - nb;
    return (a < nb) ? r - EPS : r;
}

// a - b for canonical a, b (generic.rs:31-39).
GL_HD u64 sub(u64 a, u64 b) {
    u64 r = a - b;
    return (a < b) ? r - EPS : r;
}

GL_HD u64 neg(u64 a) { return a ? P - a : 0; }

// low + mid*2^64 + high*2^96 mod p, with mid < 2^32 and high < 2^63;
// canonical result (generic.rs:77-108).
GL_HD u64 reduce159(u64 low, u32 mid, u64 high) {
    u64 low2 = low - high;
    if (low < high) low2 -= EPS;  // + p (mod 2^64)
    u64 prod = ((u64)mid << 32) - (u64)mid;  // mid * (2^32 - 1), exact
    u64 r = prod + low2;
    if (r < prod) r += EPS;       // carry: + 2^64 = + (2^32 - 1)
    return (r >= P) ? r - P : r;
}

// 64x64 -> 128 product from four 32x32 -> 64 partial products.
GL_HD void mul128(u64 a, u64 b, u64 &lo, u64 &hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 m1 = (u64)a0 * b1 + (p00 >> 32);
    u64 m2 = (u64)a1 * b0 + (u32)m1;
    lo = (m2 << 32) | (u32)p00;
    hi = (u64)a1 * b1 + (m1 >> 32) + (m2 >> 32);
}

// a * b mod p for any u64 a, b; canonical result (generic.rs:44-48,69-71).
GL_HD u64 mul(u64 a, u64 b) {
    u64 lo, hi;
    mul128(a, b, lo, hi);
    return reduce159(lo, (u32)hi, hi >> 32);
}

// Montgomery product a * b * 2^-64 mod p (generic.rs:52-54,114-124):
// canonical whenever a * b < p * 2^64 (e.g. a < p).  Twiddle tables are kept
// in Montgomery form (w * 2^64 mod p), so mont_mul(x, table[i]) = x * w.
GL_HD u64 mont_mul(u64 a, u64 b) {
    u64 x0, x1;
    mul128(a, b, x0, x1);
    u64 s = x0 + (x0 << 32);
    u64 e = s < x0;
    u64 t = s - (s >> 32) - e;
    u64 r = x1 - t;
    return (x1 < t) ? r - EPS : r;
}

// Montgomery form of a canonical value: a * 2^64 mod p.
GL_HD u64 to_mont(u64 a) { return mul(a, EPS); }

// a * 2^S for canonical a and a compile-time S (generic.rs:171-193).
template <int S_>
GL_HD u64 shl(u64 a) {
    constexpr int S = ((S_ % 192) + 192) % 192;
    if constexpr (S == 0) {
        return a;
    } else if constexpr (S >= 96) {
        return neg(shl<S - 96>(a));
    } else if constexpr (S <= 32) {
        return reduce159(a << S, (u32)(a >> (64 - S)), 0);
    } else if constexpr (S < 64) {
        return reduce159(a << S, (u32)(a >> (64 - S)), a >> (96 - S));
    } else if constexpr (S == 64) {
        return reduce159(0, (u32)a, a >> 32);
    } else {
        return reduce159(0, (u32)(a << (S - 64)), a >> (96 - S));
    }
}

// (a - b) * 2^S with the sign of 2^96 = -1 folded into the subtraction.
template <int S_>
GL_HD u64 sub_shl(u64 a, u64 b) {
    constexpr int S = ((S_ % 192) + 192) % 192;
    if constexpr (S >= 96) {
        return shl<S - 96>(sub(b, a));
    } else {
        return shl<S>(sub(a, b));
    }
}

GL_HD u64 pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = mul(a, a);
        e >>= 1;
    }
    return r;
}

// Inverse by Fermat (host-side planning only; 0 -> 0).
GL_HD u64 inv(u64 a) { return pow(a, P - 2); }

// Root of unity of order n (canonical): 2717^((p-1)/n)
// (ntt/src/field/mod.rs:42-50); 0 when n does not divide p - 1.
GL_HD u64 root(u64 n) {
    if (n == 0 || (P - 1) % n != 0) return 0;
    return pow(2717, (P - 1) / n);
}

}  // namespace gl
