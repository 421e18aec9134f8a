// Radix-257 Stockham pass by Rader's algorithm (ntt/src/ntt/rader.rs:10-94).
//
// DFT_257 of an instance is a 256-point cyclic convolution: gather
// a_q = x[86^q mod 257], A = DFT_256(a), X0 = x0 + A[0], C = A * TW
// (TW = DFT_256(w_257^(3^q) / 256), precomputed), C[0] += x0, D =
// DFT_256(C reversed on [1..]), X[3^q mod 257] = D[q].  Generators
// (gi, gk) = (3, 86) as rader.rs:21-29.  Both DFT_256 run as 16 x 16 with
// register codelets and two LDS buffers; T = 16 instances per workgroup, one
// DFT_16 task per thread per sub-step.  The pass then applies the usual
// output twiddle w_M^(m*k) and Stockham store (see PassArgs).
//
// a.tw_inner holds four 256-entry tables: w_256^e, TW, 86^q mod 257 and
// 3^q mod 257.
#include "gl_dft.hpp"
#include "gl_internal.hpp"

namespace gl {

constexpr int RD_T = 16, RD_NT = 256, RD_STRIDE = 257;

__device__ __forceinline__ int rd_col(int hi, int lo) { return lo * 16 + (hi ^ lo); }

__device__ __forceinline__ void rader257_body(const PassArgs &a, u32 blk, u64 *bufA, u64 *bufB) {
    const int tid = threadIdx.x;
    const int t = tid % RD_T, j = tid / RD_T;  // j in [0, 16)
    const u32 I = blk * (u32)RD_T + (u32)t;
    const bool valid = I < a.total_inst;
    const u64 *w256 = a.tw_inner, *spec = a.tw_inner + 256, *permk = a.tw_inner + 512,
              *permi = a.tw_inner + 768;
    u32 b = 0, i = 0, s = 0, m = 0;
    if (valid) locate(a, I, b, i, s, m);
    const u64 *src = a.in + (u64)b * a.bstride;
    auto load = [&](u64 r) -> u64 {
        u64 pos = (u64)m + (u64)a.Mp * r;
        if (a.flags & FLAG_REVERSE) pos = pos ? a.n_rev - pos : 0;
        return src[s + a.S * pos];
    };
    u64 *rowA = bufA + t * RD_STRIDE, *rowB = bufB + t * RD_STRIDE;
    // sub-step a: DFT_16 over q1 of a[q1*16 + j]
    u64 x[16];
    u64 x0 = 0;
#pragma unroll
    for (int q1 = 0; q1 < 16; q1++) x[q1] = valid ? load(permk[q1 * 16 + j]) : 0;
    if (j == 0 && valid) x0 = load(0);
    dft<16>(x);
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) x[k1] = mont_mul(x[k1], w256[(k1 * j) & 255]);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) rowA[k1 * 16 + (j ^ k1)] = x[k1];
    __syncthreads();
    // sub-step b: thread j = k1 holds A[k1 + 16 k2]; pointwise with the
    // spectrum, X0 = x0 + A[0], C[0] += x0, then store C reversed on [1..].
    u64 X0 = 0;
#pragma unroll
    for (int q2 = 0; q2 < 16; q2++) x[q2] = rowA[j * 16 + (q2 ^ j)];
    dft<16>(x);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) {
        const int k = j + 16 * k2;
        u64 c = mont_mul(x[k2], spec[k]);
        if (k == 0) {
            X0 = add(x0, x[0]);
            c = add(c, x0);
        }
        const int pos = (256 - k) & 255;
        rowB[rd_col(pos >> 4, pos & 15)] = c;
    }
    __syncthreads();
    // sub-step c: DFT_16 over q1 of D_in[q1*16 + j]
#pragma unroll
    for (int q1 = 0; q1 < 16; q1++) x[q1] = rowB[rd_col(q1, j)];
    dft<16>(x);
#pragma unroll
    for (int k1 = 1; k1 < 16; k1++) x[k1] = mont_mul(x[k1], w256[(k1 * j) & 255]);
    __syncthreads();  // rowA is reused: every sub-step b read has completed
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) rowA[k1 * 16 + (j ^ k1)] = x[k1];
    __syncthreads();
    // sub-step d: D[j + 16 k2] -> X[3^k mod 257]
#pragma unroll
    for (int q2 = 0; q2 < 16; q2++) x[q2] = rowA[j * 16 + (q2 ^ j)];
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
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) store(permi[j + 16 * k2], x[k2]);
    if (j == 0) store(0, X0);
}

__global__ __launch_bounds__(RD_NT) void rader257_kernel(PassArgs a) {
    __shared__ u64 bufA[RD_T * RD_STRIDE], bufB[RD_T * RD_STRIDE];
    rader257_body(a, blockIdx.x, bufA, bufB);
}

__global__ __launch_bounds__(RD_NT) void rader257_grouped_kernel(const PassArgs *__restrict__ descs,
                                                                 const u32 *__restrict__ starts, u32 nd) {
    __shared__ u64 bufA[RD_T * RD_STRIDE], bufB[RD_T * RD_STRIDE];
    const u32 bb = blockIdx.x;
    u32 lo = 0, hi = nd;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) / 2;
        if (starts[mid] <= bb) lo = mid; else hi = mid;
    }
    const PassArgs a = descs[lo];
    rader257_body(a, bb - starts[lo], bufA, bufB);
}

static void launch_rader257(const PassArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL(rader257_kernel, dim3(grid), dim3(RD_NT), 0, s, a);
}

static void launch_rader257_grouped(const PassArgs *descs, const u32 *starts, u32 nd, unsigned grid,
                                    hipStream_t s) {
    hipLaunchKernelGGL(rader257_grouped_kernel, dim3(grid), dim3(RD_NT), 0, s, descs, starts, nd);
}

static const PassVariant kRader257 = {257, 1, RD_T, RD_NT, 2 * RD_T * RD_STRIDE * 8, &launch_rader257,
                                      &launch_rader257_grouped, VK_RADER257, 1};

const PassVariant *rader257_variant() { return &kRader257; }

}  // namespace gl
