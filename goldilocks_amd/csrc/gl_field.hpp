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

GL_HD u32 lo32(u64 x) { return (u32)x; }
GL_HD u32 hi32(u64 x) { return (u32)(x >> 32); }
GL_HD u64 mk64(u32 lo, u32 hi) { return ((u64)hi << 32) | lo; }

// 64-bit a - b and a + b as explicit 32-bit carry chains, so the borrow or
// carry stays in a lane mask instead of being recomputed by a compare.
GL_HD u64 sub_borrow(u64 a, u64 b, u32 &c) {
    u32 c1;
    u32 r0 = __builtin_subc(lo32(a), lo32(b), 0u, &c1);
    u32 r1 = __builtin_subc(hi32(a), hi32(b), c1, &c);
    return mk64(r0, r1);
}
GL_HD u64 add_carry(u64 a, u64 b, u32 &c) {
    u32 c1;
    u32 r0 = __builtin_addc(lo32(a), lo32(b), 0u, &c1);
    u32 r1 = __builtin_addc(hi32(a), hi32(b), c1, &c);
    return mk64(r0, r1);
}

// Keeps a value opaque to the optimizer (no instruction): a product with
// it stays one v_mad instead of being rewritten as selects.
GL_HD int32_t opaque(int32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(x));
#endif
    return x;
}

// Same for a 64-bit value, which also keeps it in one register pair.
GL_HD u64 opaque64(u64 x) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(x));
#endif
    return x;
}

// r + k (2^32 - 1) mod 2^64 for k in {0, 1}: one v_mad_u64_u32.
GL_HD u64 plus_eps(u64 r, bool k) { return opaque64(r) + (u64)(u32)opaque(k ? 1 : 0) * EPS; }

// r - k (2^32 - 1) = r + k p mod 2^64 for k in {0, 1}: one v_mad_i64_i32,
// with 2^32 - 1 = 65535 * 65537.
GL_HD u64 minus_eps(u64 r, bool k) { return opaque64(r) + (u64)((int64_t)opaque(k ? -65535 : 0) * 65537); }

// a + b for canonical a, b: same result as generic.rs:17-26.  a + b >= p
// exactly when the sum carried (then s + EPS < p) or s >= p (then s + EPS
// wraps to s - p).
GL_HD u64 add(u64 a, u64 b) {
    u32 k;
    u64 s = add_carry(a, b, k);
    return plus_eps(s, k | (s >= P));
}

// a - b for canonical a, b (generic.rs:31-39): on borrow add p, which is
// - (2^32 - 1) modulo 2^64.
GL_HD u64 sub(u64 a, u64 b) {
    u32 c;
    u64 r = sub_borrow(a, b, c);
    return minus_eps(r, c);
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
// With x = x0 + x1 * 2^64: s = x0 + (x0 << 32), t = s - (s >> 32) - carry,
// result x1 - t, plus p on borrow.
GL_HD u64 mont_mul(u64 a, u64 b) {
    u64 x0, x1;
    mul128(a, b, x0, x1);
    u32 e, c1, c2, bw;
    u32 s1 = __builtin_addc(hi32(x0), lo32(x0), 0u, &e);
    u32 t0 = __builtin_subc(lo32(x0), s1, e, &c1);
    u32 t1 = __builtin_subc(s1, 0u, c1, &c2);
    u64 r = sub_borrow(x1, mk64(t0, t1), bw);
    return minus_eps(r, bw);
}

// Montgomery form of a canonical value: a * 2^64 mod p.
GL_HD u64 to_mont(u64 a) { return mul(a, EPS); }

// a * 2^S for any u64 a and a compile-time 0 < S < 96; canonical result.
// The product a * 2^S = L + H * 2^64 (L = a << S, H = a >> (64 - S)) is
// reduced with 2^64 = 2^32 - 1 and 2^96 = -1 (generic.rs:144-193); the
// multiply by 2^32 - 1 folds into one v_mad_u64_u32.
template <int S>
GL_HD u64 shl_lt96(u64 a) {
    static_assert(S > 0 && S < 96, "shift range");
    if constexpr (S <= 32) {
        // H < 2^32: value = L + H * EPS, one carry at most.
        u32 k;
        u64 r = add_carry((u64)(u32)(a >> (64 - S)) * EPS, a << S, k);
        return plus_eps(r, k | (r >= P));
    } else if constexpr (S < 64) {
        // H = h0 + h1 2^32 < 2^S, L = a << S has a zero low word:
        // value = (Lh + h0) 2^32 - h0 - h1.  With u0 + c 2^32 = Lh + h0,
        // X = {c ? EPS : 0, u0} equals (Lh + h0) 2^32 mod p and X < p, so
        // X - h0 - h1 (h0 + h1 < 2^33) needs at most one + p.
        const u32 a0 = lo32(a), a1 = hi32(a);
        const u32 h0 = (u32)(a >> (64 - S)), h1 = a1 >> (64 - S);
        u32 c, b1, b2;
        const u32 u0 = __builtin_addc(a0 << (S - 32), h0, 0u, &c);
        u64 r = sub_borrow(mk64(c ? 0xffffffffu : 0u, u0), (u64)h0, b1);
        r = sub_borrow(r, (u64)h1, b2);
        return minus_eps(r, b1 | b2);
    } else if constexpr (S == 64) {
        // value = a0 * EPS - a1
        u32 b;
        u64 d = sub_borrow((u64)lo32(a) * EPS, (u64)hi32(a), b);
        return minus_eps(d, b);
    } else {
        // limbs c2 = bits [64, 96), u = bits [96, 160): value = c2 * EPS - u
        u32 b;
        u64 d = sub_borrow((u64)(u32)(a << (S - 64)) * EPS, a >> (96 - S), b);
        return minus_eps(d, b);
    }
}

// a * 2^S for canonical a (any integer S; 2^192 = 1, 2^96 = -1).
template <int S_>
GL_HD u64 shl(u64 a) {
    constexpr int S = ((S_ % 192) + 192) % 192;
    if constexpr (S == 0) {
        return a;
    } else if constexpr (S == 96) {
        return neg(a);
    } else if constexpr (S > 96) {
        return neg(shl_lt96<S - 96>(a));
    } else {
        return shl_lt96<S>(a);
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
