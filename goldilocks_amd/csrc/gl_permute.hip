// The reference's public permutations on the GPU (ntt/src/permute/mod.rs):
// matrix transposes and the bit-reversal permutation of u64 words.  All are
// pure data movement, so the kernels are bound by HBM: every word is read
// once and written once, through LDS tiles that keep both the reads and the
// writes in runs of 32 consecutive words (256 bytes).
//
// - transpose_tile_kernel: out-of-place, tile (by, bx) of the rows x cols
//   input lands transposed in tile (bx, by) of the cols x rows output.
// - transpose_square_kernel: in place for square matrices; a workgroup owns
//   the tile pair (i, j), (j, i) (one tile on the diagonal) and swaps them
//   through LDS, so no other workgroup reads what it writes.
// - bit_reverse_tile_kernel: in place for n = 2^k, k >= 10.  With
//   i = a*2^(k-5) + b*32 + c (a, c < 32), rev(i) = rev5(c)*2^(k-5) +
//   rev(b)*32 + rev5(a): the 32 x 32 tile of middle index b trades places
//   with the tile of rev(b), transposed and with both tile indices
//   bit-reversed.  Shorter lengths swap pairs directly.
#include "gl_internal.hpp"

namespace gl {

constexpr int TP_T = 32;                   // tile edge (words)
constexpr int TP_NT = 256;                 // threads per workgroup
constexpr int TP_ROWS = TP_NT / TP_T;      // tile rows per sweep
constexpr int TP_LD = TP_T + 1;            // LDS row stride (odd: no bank conflicts on the column reads)

__global__ __launch_bounds__(TP_NT) void transpose_tile_kernel(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                              u64 rows, u64 cols, u64 tiles_x) {
    __shared__ u64 tile[TP_T * TP_LD];
    const u64 t = blockIdx.x;
    const u64 by = t / tiles_x, bx = t - by * tiles_x;
    const int tx = threadIdx.x % TP_T, ty = threadIdx.x / TP_T;
    const u64 r0 = by * TP_T, c0 = bx * TP_T;
#pragma unroll
    for (int k = 0; k < TP_T; k += TP_ROWS) {
        const u64 r = r0 + ty + k, c = c0 + tx;
        if (r < rows && c < cols) tile[(ty + k) * TP_LD + tx] = in[r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TP_T; k += TP_ROWS) {
        const u64 oc = c0 + ty + k, orow = r0 + tx;  // output row oc (an input column), column orow
        if (oc < cols && orow < rows) out[oc * rows + orow] = tile[tx * TP_LD + ty + k];
    }
}

// Tile pair id -> (i, j), i <= j, in the order (0,0), (0,1), (1,1), (0,2), ...
__device__ __forceinline__ void tri_pair(u64 id, u64 &i, u64 &j) {
    u64 jj = (u64)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while (jj * (jj + 1) / 2 > id) jj--;
    while ((jj + 1) * (jj + 2) / 2 <= id) jj++;
    j = jj;
    i = id - jj * (jj + 1) / 2;
}

__global__ __launch_bounds__(TP_NT) void transpose_square_kernel(u64 *__restrict__ m, u64 n) {
    __shared__ u64 ta[TP_T * TP_LD], tb[TP_T * TP_LD];
    u64 bi, bj;
    tri_pair(blockIdx.x, bi, bj);
    const int tx = threadIdx.x % TP_T, ty = threadIdx.x / TP_T;
    const u64 ra = bi * TP_T, ca = bj * TP_T;  // tile A at rows ra.., cols ca..; tile B is its mirror
#pragma unroll
    for (int k = 0; k < TP_T; k += TP_ROWS) {
        const u64 r = ra + ty + k, c = ca + tx;
        if (r < n && c < n) ta[(ty + k) * TP_LD + tx] = m[r * n + c];
        const u64 r2 = ca + ty + k, c2 = ra + tx;
        if (bi != bj && r2 < n && c2 < n) tb[(ty + k) * TP_LD + tx] = m[r2 * n + c2];
    }
    __syncthreads();
    const u64 *src_a = ta, *src_b = bi != bj ? tb : ta;
#pragma unroll
    for (int k = 0; k < TP_T; k += TP_ROWS) {
        // tile (bi, bj) <- transpose of tile (bj, bi), and the mirror
        const u64 r = ra + ty + k, c = ca + tx;
        if (r < n && c < n) m[r * n + c] = src_b[tx * TP_LD + ty + k];
        const u64 r2 = ca + ty + k, c2 = ra + tx;
        if (bi != bj && r2 < n && c2 < n) m[r2 * n + c2] = src_a[tx * TP_LD + ty + k];
    }
}

__device__ __forceinline__ u32 rev_bits(u32 x, int bits) { return bits ? __brev(x) >> (32 - bits) : 0; }

__global__ __launch_bounds__(TP_NT) void bit_reverse_tile_kernel(u64 *__restrict__ v, int k) {
    __shared__ u64 ta[TP_T * TP_LD], tb[TP_T * TP_LD];
    const int mid = k - 10;
    const u32 b = blockIdx.x, b2 = rev_bits(b, mid);
    if (b > b2) return;  // the pair is handled by block b2
    const u64 row = 1ull << (k - 5);  // stride of the high index a
    const int tx = threadIdx.x % TP_T, ty = threadIdx.x / TP_T;
    u64 *pa = v + (u64)b * TP_T, *pb = v + (u64)b2 * TP_T;
#pragma unroll
    for (int q = 0; q < TP_T; q += TP_ROWS) {
        const int a = ty + q;
        ta[a * TP_LD + tx] = pa[a * row + tx];
        if (b != b2) tb[a * TP_LD + tx] = pb[a * row + tx];
    }
    __syncthreads();
    const u64 *sa = ta, *sb = b != b2 ? tb : ta;
    const int rc = (int)rev_bits((u32)tx, 5);
#pragma unroll
    for (int q = 0; q < TP_T; q += TP_ROWS) {
        const int a = ty + q, ra = (int)rev_bits((u32)a, 5);
        // v[a, b, c] = v[rev5(c), rev(b), rev5(a)]
        pa[a * row + tx] = sb[rc * TP_LD + ra];
        if (b != b2) pb[a * row + tx] = sa[rc * TP_LD + ra];
    }
}

__global__ void bit_reverse_pairs_kernel(u64 *__restrict__ v, u64 n, int k) {
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        const u64 j = rev_bits((u32)i, k);
        if (i < j) {
            const u64 x = v[i];
            v[i] = v[j];
            v[j] = x;
        }
    }
}

static unsigned grid_for(u64 blocks) { return (unsigned)blocks; }

void launch_transpose(const u64 *in, u64 *out, u64 rows, u64 cols, hipStream_t s) {
    const u64 tx = (cols + TP_T - 1) / TP_T, ty = (rows + TP_T - 1) / TP_T;
    hipLaunchKernelGGL(transpose_tile_kernel, dim3(grid_for(tx * ty)), dim3(TP_NT), 0, s, in, out, rows, cols, tx);
}

void launch_transpose_square(u64 *m, u64 n, hipStream_t s) {
    const u64 nb = (n + TP_T - 1) / TP_T;
    hipLaunchKernelGGL(transpose_square_kernel, dim3(grid_for(nb * (nb + 1) / 2)), dim3(TP_NT), 0, s, m, n);
}

void launch_bit_reverse(u64 *v, u64 n, int k, hipStream_t s) {
    if (k >= 10) {
        hipLaunchKernelGGL(bit_reverse_tile_kernel, dim3(grid_for(1ull << (k - 10))), dim3(TP_NT), 0, s, v, k);
    } else {
        hipLaunchKernelGGL(bit_reverse_pairs_kernel, dim3(1), dim3(TP_NT), 0, s, v, n, k);
    }
}

}  // namespace gl
