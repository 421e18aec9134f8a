// Elementwise kernels of the NTT runtime for gfx950 (random fill, twiddle
// tables, the radix-65537 Rader pass pieces, pointwise products, the
// commit's coset extension) and the assembled table of pass kernel
// variants, whose instantiations live in gl_variants_*.hip (templates in
// gl_pass.hpp).
#include <vector>

#include "gl_dft.hpp"
#include "gl_internal.hpp"

namespace gl {

const PassVariant *pass_variants_part0(int *count);
const PassVariant *pass_variants_part1(int *count);
const PassVariant *pass_variants_part2(int *count);
const PassVariant *pass_variants_part3(int *count);
const PassVariant *pass_variants_part4(int *count);
const PassVariant *pass_variants_part5(int *count);
const PassVariant *pass_variants_part6(int *count);
const PassVariant *pass_variants_part7(int *count);

// The variants in one fixed order (their index is the variant id of the C
// ABI): the parts concatenated.
const PassVariant *pass_variants(int *count) {
    static const std::vector<PassVariant> all = [] {
        const PassVariant *(*parts[])(int *) = {pass_variants_part0, pass_variants_part1, pass_variants_part2,
                                                pass_variants_part3, pass_variants_part4, pass_variants_part5,
                                                pass_variants_part6, pass_variants_part7};
        std::vector<PassVariant> v;
        for (auto part : parts) {
            int c = 0;
            const PassVariant *p = part(&c);
            v.insert(v.end(), p, p + c);
        }
        return v;
    }();
    *count = (int)all.size();
    return all.data();
}

// ---- elementwise kernels ----

__device__ __forceinline__ u64 splitmix64(u64 &state) {
    u64 z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Word i: first splitmix64 output < p of the stream seeded at
// seed + i * 0x9e3779b97f4a7c15 (the same map the tests' generator uses).
__global__ void fill_random_kernel(u64 *v, u64 n, u64 seed) {
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        u64 st = seed + i * 0x9e3779b97f4a7c15ull, x;
        do x = splitmix64(st); while (x >= P);
        v[i] = x;
    }
}

void launch_fill_random(u64 *v, u64 n, u64 seed, hipStream_t s) {
    unsigned grid = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    if (grid == 0) return;
    hipLaunchKernelGGL(fill_random_kernel, dim3(grid), dim3(256), 0, s, v, n, seed);
}

// Twiddle matrix in output order, Montgomery form: out[k + R*m] = w^(m*k)
// for m < count / R, k < R, where w^e = lo[e & mask] * hi[e >> lo_bits].
__global__ void twiddle_matrix_kernel(u64 *out, u64 count, u64 R, const u64 *lo, const u64 *hi, u32 lo_bits) {
    for (u64 idx = blockIdx.x * (u64)blockDim.x + threadIdx.x; idx < count; idx += (u64)gridDim.x * blockDim.x) {
        const u64 m = idx / R, k = idx - m * R, e = m * k;
        out[idx] = mont_mul(lo[e & ((1ull << lo_bits) - 1)], hi[e >> lo_bits]);
    }
}

void launch_twiddle_matrix(u64 *out, u64 count, u64 R, const u64 *lo, const u64 *hi, u32 lo_bits, hipStream_t s) {
    unsigned grid = (unsigned)((count + 255) / 256 < 65536 ? (count + 255) / 256 : 65536);
    if (grid == 0) return;
    hipLaunchKernelGGL(twiddle_matrix_kernel, dim3(grid), dim3(256), 0, s, out, count, R, lo, hi, lo_bits);
}

// out[j] = lo[j & mask] * hi[j >> lo_bits] (Montgomery forms): the powers of
// a table's base for j < count.
__global__ void power_series_kernel(u64 *out, u64 count, const u64 *lo, const u64 *hi, u32 lo_bits) {
    for (u64 j = blockIdx.x * (u64)blockDim.x + threadIdx.x; j < count; j += (u64)gridDim.x * blockDim.x)
        out[j] = mont_mul(lo[j & ((1ull << lo_bits) - 1)], hi[j >> lo_bits]);
}

__global__ void scale_table_kernel(u64 *out, const u64 *in, u64 count, u64 scale) {
    for (u64 j = blockIdx.x * (u64)blockDim.x + threadIdx.x; j < count; j += (u64)gridDim.x * blockDim.x)
        out[j] = mont_mul(in[j], scale);
}

void launch_scale_table(u64 *out, const u64 *in, u64 count, u64 scale, hipStream_t s) {
    if (count == 0) return;
    const u64 blocks = (count + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(scale_table_kernel, dim3(grid), dim3(256), 0, s, out, in, count, scale);
}

void launch_power_series(u64 *out, u64 count, const u64 *lo, const u64 *hi, u32 lo_bits, hipStream_t s) {
    unsigned grid = (unsigned)((count + 255) / 256 < 65536 ? (count + 255) / 256 : 65536);
    if (grid == 0) return;
    hipLaunchKernelGGL(power_series_kernel, dim3(grid), dim3(256), 0, s, out, count, lo, hi, lo_bits);
}

// ---- radix-65537 Rader pass pieces (rader.rs:65-93 with (gi, gk) = (3, 21846)).
// Instance I (of the pass) owns C[I*65536 .. +65536) and x0[I].

// C[I][q] = x[gk^q mod 65537], x0[I] = x[0] for the pass's input view.
__global__ void rader_big_gather_kernel(PassArgs a, const u32 *__restrict__ permk, u64 *__restrict__ C,
                                        u64 *__restrict__ x0) {
    const u64 total = (u64)a.total_inst * 65536ull;
    for (u64 g = blockIdx.x * (u64)blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        const u32 I = (u32)(g >> 16), q = (u32)(g & 65535);
        u32 b, i, s, m;
        locate(a, I, b, i, s, m);
        const u64 *src = a.in + (u64)b * a.bstride;
        auto at = [&](u64 r) {
            u64 pos = (u64)m + (u64)a.Mp * r;
            if (a.flags & FLAG_REVERSE) pos = pos ? a.n_rev - pos : 0;
            return src[s + a.S * pos];
        };
        C[g] = at(permk[q]);
        if (q == 0) x0[I] = at(0);
    }
}

// X0 = x0 + A[0]; C = A * TW (Montgomery spectrum); C[0] += x0.  X0 is kept
// in x0[] for the scatter.
__global__ void rader_big_pointwise_kernel(u64 *__restrict__ C, u64 *__restrict__ x0, const u64 *__restrict__ spec,
                                           u64 ninst) {
    const u64 total = ninst * 65536ull;
    for (u64 g = blockIdx.x * (u64)blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        const u32 q = (u32)(g & 65535);
        u64 c = mont_mul(C[g], spec[q]);
        if (q == 0) {
            const u64 I = g >> 16, v = x0[I];
            x0[I] = add(v, C[g]);
            c = add(c, v);
        }
        C[g] = c;
    }
}

// out[s + S*(k + 65537*m)] = D[q] (k = gi^q mod 65537), out at k = 0 = X0,
// with the pass's output twiddle and scale.
__global__ void rader_big_scatter_kernel(PassArgs a, const u32 *__restrict__ permi, const u64 *__restrict__ D,
                                         const u64 *__restrict__ X0) {
    const u64 total = (u64)a.total_inst * 65537ull;
    for (u64 g = blockIdx.x * (u64)blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        const u32 I = (u32)(g / 65537ull), j = (u32)(g - (u64)I * 65537ull);
        u32 b, i, s, m;
        locate(a, I, b, i, s, m);
        u64 v, k;
        if (j == 65536) {
            v = X0[I];
            k = 0;
        } else {
            v = D[(u64)I * 65536ull + j];
            k = permi[j];
        }
        if (a.flags & FLAG_TWIDDLE_FULL) {
            v = mont_mul(v, a.tw_lo[k + 65537ull * (u64)m]);
        } else if (a.flags & FLAG_TWIDDLE) {
            const u64 e = (u64)m * k;
            v = mont_mul(v, mont_mul(a.tw_lo[e & ((1ull << a.lo_bits) - 1)], a.tw_hi[e >> a.lo_bits]));
        }
        if (a.flags & FLAG_SCALE) v = mont_mul(v, a.scale);
        a.out[(u64)b * a.bstride + s + a.S * (k + 65537ull * (u64)m)] = v;
    }
}

static unsigned grid_for(u64 work) {
    u64 g = (work + 255) / 256;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

void launch_rader_big_gather(const PassArgs &a, const u32 *permk, u64 *C, u64 *x0, hipStream_t s) {
    hipLaunchKernelGGL(rader_big_gather_kernel, dim3(grid_for((u64)a.total_inst * 65536)), dim3(256), 0, s, a, permk,
                       C, x0);
}
void launch_rader_big_pointwise(u64 *C, u64 *x0, const u64 *spec, u64 ninst, hipStream_t s) {
    hipLaunchKernelGGL(rader_big_pointwise_kernel, dim3(grid_for(ninst * 65536)), dim3(256), 0, s, C, x0, spec,
                       ninst);
}
void launch_rader_big_scatter(const PassArgs &a, const u32 *permi, const u64 *D, const u64 *X0, hipStream_t s) {
    hipLaunchKernelGGL(rader_big_scatter_kernel, dim3(grid_for((u64)a.total_inst * 65537)), dim3(256), 0, s, a,
                       permi, D, X0);
}

// a[i] = a[i] * b[i] as the reference's Field product (Montgomery).
__global__ void pointwise_mont_kernel(u64 *a, const u64 *b, u64 n) {
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
        a[i] = mont_mul(a[i], b[i]);
}

void launch_pointwise_mont(u64 *a, const u64 *b, u64 n, hipStream_t s) {
    unsigned grid = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    if (grid == 0) return;
    hipLaunchKernelGGL(pointwise_mont_kernel, dim3(grid), dim3(256), 0, s, a, b, n);
}

// Coset extension for the commit: ext[r][c] = coef[r][c] * g^r for r < rows,
// zero for rows <= r < blowup*rows (g^r = lo[r & mask] * hi[r >> lo_bits]).
__global__ void coset_extend_kernel(const u64 *coef, u64 *ext, u64 rows, u64 cols, u64 blowup,
                                    const u64 *lo, const u64 *hi, u32 lo_bits) {
    u64 total = rows * blowup * cols;
    for (u64 g = blockIdx.x * (u64)blockDim.x + threadIdx.x; g < total; g += (u64)gridDim.x * blockDim.x) {
        u64 r = g / cols;
        u64 v = 0;
        if (r < rows) v = mont_mul(coef[g], mont_mul(lo[r & ((1ull << lo_bits) - 1)], hi[r >> lo_bits]));
        ext[g] = v;
    }
}

void launch_coset_extend(const u64 *coef, u64 *ext, u64 rows, u64 cols, u64 blowup,
                         const u64 *pw_lo, const u64 *pw_hi, u32 lo_bits, hipStream_t s) {
    u64 total = rows * blowup * cols;
    unsigned grid = (unsigned)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
    if (grid == 0) return;
    hipLaunchKernelGGL(coset_extend_kernel, dim3(grid), dim3(256), 0, s, coef, ext, rows, cols,
                       blowup, pw_lo, pw_hi, lo_bits);
}

}  // namespace gl
