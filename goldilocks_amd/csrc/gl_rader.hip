// Radix-257 Stockham pass by Rader's algorithm (ntt/src/ntt/rader.rs:10-94).
//
// DFT_257 of an instance is a 256-point cyclic convolution: gather
// a_q = x[86^q mod 257], A = DFT_256(a), X0 = x0 + A[0], C = A * TW
// (TW = DFT_256(w_257^(3^q) / 256), precomputed), C[0] += x0, D =
// DFT_256(C reversed on [1..]), X[3^q mod 257] = D[q].  Generators
// (gi, gk) = (3, 86) as rader.rs:21-29.  Both 256-point transforms run as
// 16 x 16 with register codelets; the second is taken as the unscaled
// inverse DFT of C (the same numbers), which starts on the registers the
// first one ends in, so the instance crosses the LDS tile twice, not three
// times.  T = 16 instances per workgroup, one DFT_16 task per thread per
// sub-step.  The pass then applies the usual output twiddle w_M^(m*k) and
// Stockham store (see PassArgs).
//
// a.tw_inner holds w_256^(k1*j) at j*16 + k1, the spectrum TW at 256 + k
// and w_256^(-j*q1) at 512 + j*16 + q1 (Montgomery form).  The Rader permutations 86^q and 3^q mod 257
// are computed in registers: 86^(16*q1 + j) = 86^(16*q1) * 86^j, the first
// factor a compile-time constant, reduced with 2^8 = -1 (mod 257).
#include "gl_dft.hpp"
#include "gl_internal.hpp"

namespace gl {

constexpr int RD_T = 16, RD_NT = 256;
// One LDS tile: instance t owns RD_TS words, row r of it RD_RS words (odd
// strides: row-wise and column-wise accesses both spread over the banks).
constexpr int RD_RS = 17, RD_TS = 16 * RD_RS + 1;
// With GL_RADER_HALF_EXCHANGE (the default) the tile holds 32-bit halves
// (low words, then high words), RD_TS32 entries per instance: conflict-free
// over 32 banks for both access patterns and half the LDS of the 64-bit
// tile, so five workgroups fit per CU instead of four (the VGPR limit at 96
// registers).  On a 257 * 2^20-word transform the pass ran 1590 -> 1515 us;
// cache-resident sizes are unchanged.
#ifndef GL_RADER_HALF_EXCHANGE
#define GL_RADER_HALF_EXCHANGE 1
#endif
constexpr int RD_TS32 = 16 * RD_RS + 2;
constexpr int RD_LDS_WORDS = GL_RADER_HALF_EXCHANGE ? RD_T * RD_TS32 / 2 : RD_T * RD_TS;

// x mod 257 for x <= 256 * 256 (x = 256 a + b = b - a, |b - a| <= 256)
__device__ __forceinline__ u32 mod257(u32 x) {
    const int r = (int)(x & 255u) - (int)(x >> 8);
    return (u32)(r < 0 ? r + 257 : r);
}

constexpr u32 cx_pow257(u32 g, int e) {
    u32 r = 1;
    for (int i = 0; i < e; i++) r = r * g % 257u;
    return r;
}

// G^(16 q) mod 257 for q < 16, evaluated at compile time (namespace-scope
// constexpr tables, indexed by unrolled loop counters).
struct Pow257x16 {
    u32 v[16];
};
template <u32 G>
constexpr Pow257x16 pow257_x16() {
    Pow257x16 t{};
    for (int q = 0; q < 16; q++) t.v[q] = cx_pow257(G, 16 * q);
    return t;
}
constexpr Pow257x16 kGk16 = pow257_x16<86>(), kGi16 = pow257_x16<3>();

// G^j mod 257 for a runtime j < 16 (square-and-multiply over j's 4 bits)
template <u32 G>
__device__ __forceinline__ u32 pow257_small(u32 j) {
    constexpr u32 g2 = cx_pow257(G, 2), g4 = cx_pow257(G, 4), g8 = cx_pow257(G, 8);
    u32 r = 1;
    r = (j & 1u) ? mod257(r * G) : r;
    r = (j & 2u) ? mod257(r * g2) : r;
    r = (j & 4u) ? mod257(r * g4) : r;
    r = (j & 8u) ? mod257(r * g8) : r;
    return r;
}

__device__ __forceinline__ void rader257_body(const PassArgs &a, u32 blk, u64 *lds) {
    const int tid = threadIdx.x;
    const int t = tid % RD_T, j = tid / RD_T;  // j in [0, 16)
    const u32 I = blk * (u32)RD_T + (u32)t;
    const bool valid = I < a.total_inst;
    const u64 *w256 = a.tw_inner + j * 16, *spec = a.tw_inner + 256 + j;
    u32 b = 0, i = 0, s = 0, m = 0;
    if (valid) locate(a, I, b, i, s, m);
    const u64 *src = a.in + (u64)b * a.bstride;
    auto load = [&](u32 r) -> u64 {
        if (a.flags & FLAG_REVERSE) {
            u64 pos = (u64)m + (u64)a.Mp * r;
            pos = pos ? a.n_rev - pos : 0;
            return src[s + a.S * pos];
        }
        return src[(u64)i + a.SMp * r];
    };
    // v[k] (word k of this thread's column j) -> out[k] (word k of row j)
    auto exchange = [&](const u64 *v, u64 *out) {
        if constexpr (GL_RADER_HALF_EXCHANGE) {
            u32 *r32 = (u32 *)lds + t * RD_TS32;
#pragma unroll
            for (int k = 0; k < 16; k++) r32[k * RD_RS + j] = lo32(v[k]);
            __syncthreads();
            u32 lo[16];
#pragma unroll
            for (int k = 0; k < 16; k++) lo[k] = r32[j * RD_RS + k];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 16; k++) r32[k * RD_RS + j] = hi32(v[k]);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 16; k++) out[k] = mk64(lo[k], r32[j * RD_RS + k]);
        } else {
            u64 *row = lds + t * RD_TS;
#pragma unroll
            for (int k = 0; k < 16; k++) row[k * RD_RS + j] = v[k];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 16; k++) out[k] = row[j * RD_RS + k];
        }
    };
    u64 w[16], x[16];
    // sub-step a: DFT_16 over q1 of a[q1*16 + j] = x[86^(q1*16 + j)]
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) w[k1] = w256[k1];
    const u32 gk = pow257_small<86>((u32)j);
#pragma unroll
    for (int q1 = 0; q1 < 16; q1++) x[q1] = valid ? load(mod257(kGk16.v[q1] * gk)) : 0;
    const u64 x0 = (j == 0 && valid) ? load(0) : 0;
    dft<16>(x);
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) x[k1] = mont_mul(x[k1], w[k1]);
    // sub-step b: thread j holds A[j + 16 k2]; pointwise with the spectrum:
    // X0 = x0 + A[0], C[0] += x0
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) w[k2] = spec[16 * k2];
    exchange(x, x);
    dft<16>(x);
    u64 X0 = 0;
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
        u64 c = mont_mul(x[k2], w[k2]);
        if (k2 == 0 && j == 0) {
            X0 = add(x0, x[0]);
            c = add(c, x0);
        }
        x[k2] = c;
    }
    // sub-step c: D = DFT_256(C reversed on [1..]) is the unscaled inverse
    // DFT of C: D[q1 + 16 q2] = sum_j w_16^(-j q2) w_256^(-j q1) F_j[q1],
    // F_j[q1] = sum_k2 C[j + 16 k2] w_16^(-k2 q1).  Thread j holds
    // C[j + 16 k2], so F_j is its own DFT_16 read at -q1: no exchange.
    const u64 *winv = a.tw_inner + 512 + j * 16;
#pragma unroll
    for (int q1 = 1; q1 < 16; q1++) w[q1] = winv[q1];
    dft<16>(x);
    u64 g[16];
    g[0] = x[0];
#pragma unroll
    for (int q1 = 1; q1 < 16; q1++) g[q1] = mont_mul(x[16 - q1], w[q1]);
    __syncthreads();  // every sub-step b read is done before the tile is reused
    // sub-step d: thread j (as q1) transforms column j of the tile over the
    // rows, read at -q2: D[j + 16 q2] = x[-q2] -> X[3^(j + 16 q2) mod 257]
    exchange(g, x);
    dft<16>(x);
    if (!valid) return;
    u64 *dst = a.out + (u64)b * a.bstride + s + a.S * 257ull * (u64)m;
    auto store = [&](u64 kk, u64 v) {
        if (a.flags & FLAG_TWIDDLE_FULL) {
            v = mont_mul(v, a.tw_lo[kk + 257ull * (u64)m]);
        } else if (a.flags & FLAG_TWIDDLE) {
            u64 e = (u64)m * kk;
            v = mont_mul(v, mont_mul(a.tw_lo[e & ((1ull << a.lo_bits) - 1)], a.tw_hi[e >> a.lo_bits]));
        }
        if (a.flags & FLAG_SCALE) v = mont_mul(v, a.scale);
        dst[a.S * kk] = v;
    };
    const u32 hi = pow257_small<3>((u32)j);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) store(mod257(kGi16.v[k2] * hi), x[(16 - k2) & 15]);
    if (j == 0) store(0, X0);
}

__global__ __launch_bounds__(RD_NT) void rader257_kernel(PassArgs a) {
    __shared__ u64 lds[RD_LDS_WORDS];
    rader257_body(a, blockIdx.x, lds);
}

__global__ __launch_bounds__(RD_NT) void rader257_grouped_kernel(const PassArgs *__restrict__ descs,
                                                                 const u32 *__restrict__ starts, u32 nd) {
    __shared__ u64 lds[RD_LDS_WORDS];
    const u32 bb = blockIdx.x;
    u32 lo = 0, hi = nd;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) / 2;
        if (starts[mid] <= bb) lo = mid; else hi = mid;
    }
    const PassArgs a = descs[lo];
    rader257_body(a, bb - starts[lo], lds);
}

static void launch_rader257(const PassArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL(rader257_kernel, dim3(grid), dim3(RD_NT), 0, s, a);
}

static void launch_rader257_grouped(const PassArgs *descs, const u32 *starts, u32 nd, unsigned grid,
                                    hipStream_t s) {
    hipLaunchKernelGGL(rader257_grouped_kernel, dim3(grid), dim3(RD_NT), 0, s, descs, starts, nd);
}

static const PassVariant kRader257 = {257, 1, RD_T, RD_NT, RD_LDS_WORDS * 8, &launch_rader257,
                                      &launch_rader257_grouped, VK_RADER257, 1};

const PassVariant *rader257_variant() { return &kRader257; }

}  // namespace gl
