// Register-resident DFT codelets for gfx950.
//
// dft<N>(x) computes X[k] = sum_j x[j] w_N^(jk) in place on N registers,
// natural order in and out, w_N = 2717^((p-1)/N).  The element set follows
// the sizes the reference's generated codelets cover (ntt/src/ntt/small.rs,
// generated by ntt/src/bin/codegen.rs):
//   - powers of two up to 64: radix-2 decimation in frequency whose twiddles
//     are all shifts, because w_192 = 2 (codegen.rs:10-36);
//   - 3: w_3 = 2^64, again shifts only;
//   - 5 and 17: Rader's algorithm over a shift-only DFT of size p-1 with
//     p-1 constant multiplies (rader.rs:10-94, generators from rader.rs:21-29);
//   - coprime products (15, 3*2^k, 5*2^k, 17*2^k, ...): Good-Thomas index
//     maps with no twiddles (good_thomas.rs:26-126).
// All index maps are compile-time, so permutations are register renames.
#pragma once

#include "gl_field.hpp"

namespace gl {

// ---- compile-time field helpers (for constant tables) ----
constexpr u64 cx_mul(u64 a, u64 b) { return (u64)(((unsigned __int128)a * b) % P); }
constexpr u64 cx_add(u64 a, u64 b) { return (u64)(((unsigned __int128)a + b) % P); }
constexpr u64 cx_pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = cx_mul(r, a);
        a = cx_mul(a, a);
        e >>= 1;
    }
    return r;
}
constexpr u64 cx_root(u64 n) { return cx_pow(2717, (P - 1) / n); }
constexpr u64 cx_inv(u64 a) { return cx_pow(a, P - 2); }
constexpr int cx_modpow(int g, int e, int m) {
    int r = 1;
    for (int i = 0; i < e; i++) r = (r * g) % m;
    return r;
}
constexpr int cx_modinv(int a, int m) {
    for (int x = 1; x < m; x++)
        if ((a * x) % m == 1) return x;
    return 0;
}
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr int log2i(int n) { return n <= 1 ? 0 : 1 + log2i(n / 2); }
constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}
constexpr int odd_part(int n) { return n % 2 ? n : odd_part(n / 2); }

template <int N>
GL_HD void dft(u64 *x);

// ---- powers of two: radix-2 DIF, twiddle w_N^j = 2^(192 j / N) ----
template <int N, int O, int J>
GL_HD void dif_layer(u64 *x) {
    if constexpr (J < N / 2) {
        u64 a = x[O + J], b = x[O + J + N / 2];
        x[O + J] = add(a, b);
        x[O + J + N / 2] = sub_shl<(192 / N) * J>(a, b);
        dif_layer<N, O, J + 1>(x);
    }
}

template <int N, int O>
GL_HD void dif(u64 *x) {
    if constexpr (N > 1) {
        dif_layer<N, O, 0>(x);
        dif<N / 2, O>(x);
        dif<N / 2, O + N / 2>(x);
    }
}

template <int N>
GL_HD void dft_pow2(u64 *x) {
    static_assert(N <= 64 && 192 % N == 0, "shift-only power-of-two codelet");
    dif<N, 0>(x);
    u64 t[N];
#pragma unroll
    for (int k = 0; k < N; k++) t[k] = x[bitrev(k, log2i(N))];
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = t[k];
}

// dft_pow2 for inputs whose upper half x[N/2 .. N) is zero (and is not
// read): the first layer's butterflies (a + 0, (a - 0) w_N^J) reduce to
// x[J + N/2] = x[J] w_N^J; the results are the same canonical words.
template <int N, int J>
GL_HD void dif_layer_upper_zero(u64 *x) {
    if constexpr (J < N / 2) {
        x[J + N / 2] = shl<(192 / N) * J>(x[J]);
        dif_layer_upper_zero<N, J + 1>(x);
    }
}

template <int N>
GL_HD void dft_pow2_upper_zero(u64 *x) {
    static_assert(N <= 64 && 192 % N == 0, "shift-only power-of-two codelet");
    dif_layer_upper_zero<N, 0>(x);
    dif<N / 2, 0>(x);
    dif<N / 2, N / 2>(x);
    u64 t[N];
#pragma unroll
    for (int k = 0; k < N; k++) t[k] = x[bitrev(k, log2i(N))];
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = t[k];
}

// ---- 3: w_3 = 2^64 ----
GL_HD void dft3(u64 *x) {
    u64 a = x[0], b = x[1], c = x[2];
    u64 s = add(b, c);
    u64 u = sub(shl<64>(b), shl<32>(c));  // w b + w^2 c, w^2 = 2^128 = -2^32
    x[0] = add(a, s);
    x[1] = add(a, u);
    x[2] = sub(sub(a, s), u);              // w^2 b + w c = -(b + c) - u
}

// ---- Rader for primes 5 and 17 ----
template <int Pr>
struct RaderGen;
template <> struct RaderGen<5> { static constexpr int gi = 2, gk = 3; };
template <> struct RaderGen<17> { static constexpr int gi = 3, gk = 6; };

// DFT_(Pr-1) of w_Pr^(gi^q) / (Pr-1), entry q (rader.rs:41-48).
template <int Pr>
constexpr u64 rader_tw(int q) {
    constexpr int M = Pr - 1;
    u64 w = cx_root(Pr), wm = cx_root(M), scale = cx_inv(M), acc = 0;
    for (int j = 0; j < M; j++) {
        u64 c = cx_mul(cx_pow(w, (u64)cx_modpow(RaderGen<Pr>::gi, j, Pr)), scale);
        acc = cx_add(acc, cx_mul(c, cx_pow(wm, (u64)((j * q) % M))));
    }
    return acc;
}

template <int Pr, int Q>
GL_HD void rader_scale(u64 *c) {
    if constexpr (Q < Pr - 1) {
        constexpr u64 t = cx_mul(rader_tw<Pr>(Q), 0xffffffffull);  // Montgomery form
        c[Q] = mont_mul(c[Q], t);
        rader_scale<Pr, Q + 1>(c);
    }
}

template <int Pr>
GL_HD void dft_rader(u64 *x) {
    constexpr int M = Pr - 1, gi = RaderGen<Pr>::gi, gk = RaderGen<Pr>::gk;
    u64 c[M];
#pragma unroll
    for (int q = 0; q < M; q++) c[q] = x[cx_modpow(gk, q, Pr)];
    dft<M>(c);
    u64 x0 = x[0];
    x[0] = add(x0, c[0]);
    rader_scale<Pr, 0>(c);
    c[0] = add(c[0], x0);
    u64 d[M];
    d[0] = c[0];
#pragma unroll
    for (int q = 1; q < M; q++) d[q] = c[M - q];
    dft<M>(d);
#pragma unroll
    for (int q = 0; q < M; q++) x[cx_modpow(gi, q, Pr)] = d[q];
}

// ---- Good-Thomas N = N1 * N2, gcd(N1, N2) = 1 (good_thomas.rs:104-126) ----
template <int N1, int N2>
GL_HD void dft_gt(u64 *x) {
    constexpr int N = N1 * N2;
    constexpr int k3 = N2 * cx_modinv(N2 % N1, N1), k4 = N1 * cx_modinv(N1 % N2, N2);
    u64 row2[N1][N2];
#pragma unroll
    for (int i1 = 0; i1 < N1; i1++) {
#pragma unroll
        for (int i2 = 0; i2 < N2; i2++) row2[i1][i2] = x[(i1 * N2 + i2 * N1) % N];
        dft<N2>(row2[i1]);
    }
    u64 row1[N2][N1];
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
#pragma unroll
        for (int i1 = 0; i1 < N1; i1++) row1[k2][i1] = row2[i1][k2];
        dft<N1>(row1[k2]);
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++)
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) x[(k1 * k3 + k2 * k4) % N] = row1[k2][k1];
}

template <int N>
GL_HD void dft(u64 *x) {
    if constexpr (N == 1) {
    } else if constexpr (is_pow2(N)) {
        dft_pow2<N>(x);
    } else if constexpr (N == 3) {
        dft3(x);
    } else if constexpr (N == 5 || N == 17) {
        dft_rader<N>(x);
    } else if constexpr (N == 15) {
        dft_gt<3, 5>(x);
    } else if constexpr (N == 51) {
        dft_gt<3, 17>(x);
    } else if constexpr (N == 85) {
        dft_gt<5, 17>(x);
    } else {
        static_assert(odd_part(N) != N && (odd_part(N) == 3 || odd_part(N) == 5 ||
                                           odd_part(N) == 15 || odd_part(N) == 17),
                      "unsupported codelet size");
        dft_gt<odd_part(N), N / odd_part(N)>(x);
    }
}

}  // namespace gl
