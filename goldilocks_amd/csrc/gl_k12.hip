// Keccak-p[1600,12] / KangarooTwelve layer hashing and the row tree for gfx950.
//
// One 25-lane state per thread (50 VGPRs): a layer of the reference's Merkle
// tree (pcs/src/k12/mod.rs:41-60) hashes independent 160-byte blocks, so each
// thread absorbs its block, pads it as pcs/src/k12/generic.rs:62-84 does
// (customization length byte 0x00, suffix 0x07, final bit 0x80 at byte 167),
// runs the 12 rounds of generic.rs:25-60 and writes the first 32 bytes.
// The round constants are the last 12 of FIPS 202's; the rotation offsets are
// the standard Keccak rho offsets, applied with the combined rho-pi lane map.
#include "gl_internal.hpp"

namespace gl {

constexpr u64 kRC[12] = {
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull,
    0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,
    0x000000000000800aull, 0x800000008000000aull, 0x8000000080008081ull,
    0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull,
};

// A 64-bit lane as two 32-bit halves.
struct Lane {
    u32 lo, hi;
};

__device__ __forceinline__ u32 xor3(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }

// rotl64 by a compile-time R in (0, 64): two funnel shifts (v_alignbit_b32).
template <int R>
__device__ __forceinline__ Lane rotl(Lane x) {
    if constexpr (R == 0) {
        return x;
    } else if constexpr (R == 32) {
        return {x.hi, x.lo};
    } else if constexpr (R > 32) {
        return rotl<R - 32>(Lane{x.hi, x.lo});
    } else {
        return {__builtin_amdgcn_alignbit(x.lo, x.hi, 32 - R), __builtin_amdgcn_alignbit(x.hi, x.lo, 32 - R)};
    }
}

// chi: b0 ^ (~b1 & b2) on both halves
__device__ __forceinline__ Lane chi(Lane b0, Lane b1, Lane b2) {
    return {b0.lo ^ (~b1.lo & b2.lo), b0.hi ^ (~b1.hi & b2.hi)};
}

// theta's column sums and rho: b = rotl(a ^ C[x-1] ^ rotl1(C[x+1]), r)
template <int R>
__device__ __forceinline__ Lane theta_rho(Lane a, Lane cm, Lane cp1) {
    return rotl<R>(Lane{xor3(a.lo, cm.lo, cp1.lo), xor3(a.hi, cm.hi, cp1.hi)});
}

// One round on lanes a[x + 5y] (theta, rho + pi, chi, iota).
template <int ROUND>
__device__ __forceinline__ void keccak_round(Lane *a) {
    Lane c[5], r[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
        c[x].lo = xor3(xor3(a[x].lo, a[x + 5].lo, a[x + 10].lo), a[x + 15].lo, a[x + 20].lo);
        c[x].hi = xor3(xor3(a[x].hi, a[x + 5].hi, a[x + 10].hi), a[x + 15].hi, a[x + 20].hi);
    }
#pragma unroll
    for (int x = 0; x < 5; x++) r[x] = rotl<1>(c[x]);
    // b[y, 2x+3y] = rot(a[x, y] ^ d[x], rho[x, y]), d[x] = c[x-1] ^ rotl1(c[x+1])
#define TR(X, Y, RHO) theta_rho<RHO>(a[(X) + 5 * (Y)], c[((X) + 4) % 5], r[((X) + 1) % 5])
    const Lane b00 = TR(0, 0, 0), b10 = TR(1, 1, 44), b20 = TR(2, 2, 43), b30 = TR(3, 3, 21), b40 = TR(4, 4, 14);
    const Lane b01 = TR(3, 0, 28), b11 = TR(4, 1, 20), b21 = TR(0, 2, 3), b31 = TR(1, 3, 45), b41 = TR(2, 4, 61);
    const Lane b02 = TR(1, 0, 1), b12 = TR(2, 1, 6), b22 = TR(3, 2, 25), b32 = TR(4, 3, 8), b42 = TR(0, 4, 18);
    const Lane b03 = TR(4, 0, 27), b13 = TR(0, 1, 36), b23 = TR(1, 2, 10), b33 = TR(2, 3, 15), b43 = TR(3, 4, 56);
    const Lane b04 = TR(2, 0, 62), b14 = TR(3, 1, 55), b24 = TR(4, 2, 39), b34 = TR(0, 3, 41), b44 = TR(1, 4, 2);
#undef TR
    // chi (row y holds b0y..b4y as x = 0..4) + iota
    constexpr u32 rc_lo = (u32)kRC[ROUND], rc_hi = (u32)(kRC[ROUND] >> 32);
    a[0] = chi(b00, b10, b20);
    a[0].lo ^= rc_lo;
    a[0].hi ^= rc_hi;
    a[1] = chi(b10, b20, b30);
    a[2] = chi(b20, b30, b40);
    a[3] = chi(b30, b40, b00);
    a[4] = chi(b40, b00, b10);
    a[5] = chi(b01, b11, b21);
    a[6] = chi(b11, b21, b31);
    a[7] = chi(b21, b31, b41);
    a[8] = chi(b31, b41, b01);
    a[9] = chi(b41, b01, b11);
    a[10] = chi(b02, b12, b22);
    a[11] = chi(b12, b22, b32);
    a[12] = chi(b22, b32, b42);
    a[13] = chi(b32, b42, b02);
    a[14] = chi(b42, b02, b12);
    a[15] = chi(b03, b13, b23);
    a[16] = chi(b13, b23, b33);
    a[17] = chi(b23, b33, b43);
    a[18] = chi(b33, b43, b03);
    a[19] = chi(b43, b03, b13);
    a[20] = chi(b04, b14, b24);
    a[21] = chi(b14, b24, b34);
    a[22] = chi(b24, b34, b44);
    a[23] = chi(b34, b44, b04);
    a[24] = chi(b44, b04, b14);
}

template <int ROUND>
__device__ __forceinline__ void keccak_rounds(Lane *a) {
    if constexpr (ROUND < 12) {
        keccak_round<ROUND>(a);
        keccak_rounds<ROUND + 1>(a);
    }
}

// Keccak-p[1600, 12] (rounds 12..23 of Keccak-f; generic.rs:25-60) on
// 64-bit lanes a[x + 5y], fully unrolled so the pi lane map is register
// renaming.
__device__ __forceinline__ void keccak_p12(u64 *a) {
    Lane l[25];
#pragma unroll
    for (int i = 0; i < 25; i++) l[i] = Lane{(u32)a[i], (u32)(a[i] >> 32)};
    keccak_rounds<0>(l);
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] = ((u64)l[i].hi << 32) | l[i].lo;
}

// Byte-wise absorb of a block of l <= 166 bytes with K12 padding.  Lane
// indices are compile-time so the state stays in registers.
__device__ __forceinline__ void absorb_padded(u64 *a, const uint8_t *in, u64 l) {
#pragma unroll
    for (int w = 0; w < 25; w++) {
        u64 v = 0;
        if (w < 21) {
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const u64 j = 8 * w + b;
                if (j < l) v |= (u64)in[j] << (8 * b);
                else if (j == l + 1) v |= (u64)0x07 << (8 * b);
            }
        }
        a[w] = v;
    }
    a[20] ^= 0x8000000000000000ull;
}

// out[32i .. 32i+32] = K12(in[160i .. 160i+160]) for the nfull full blocks
// of an 8-byte-aligned layer: 20 word loads, the padding words are constants
// (0x00, 0x07 at bytes 160, 161; 0x80 at byte 167).  Block nfull, when
// nblocks = nfull + 1, is the partial last block, read from the 21 padded
// words k12_stage_tail_kernel wrote.  Kept free of byte-wise paths so the
// state and the round temporaries fit in registers.
__global__ __launch_bounds__(256) void k12_layer_kernel(const u64 *__restrict__ in, u64 nfull, u64 nblocks,
                                                        const u64 *__restrict__ stage, u64 *__restrict__ out,
                                                        u64 *__restrict__ stage_next, u64 next_first) {
    for (u64 blk = blockIdx.x * (u64)blockDim.x + threadIdx.x; blk < nblocks;
         blk += (u64)gridDim.x * blockDim.x) {
        u64 a[25];
        if (blk < nfull) {
            const u64 *w = in + blk * 20;
#pragma unroll
            for (int i = 0; i < 20; i++) a[i] = w[i];
            a[20] = 0x8000000000000700ull;
        } else {
#pragma unroll
            for (int i = 0; i < 21; i++) a[i] = stage[i];
        }
#pragma unroll
        for (int i = 21; i < 25; i++) a[i] = 0;
        keccak_p12(a);
        u64 *o = out + blk * 4;
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
        o[3] = a[3];
        // Outputs of blocks >= next_first form the next layer's partial
        // block (32 q bytes, q = nblocks - next_first in 1..4): copy them
        // into its padded form; the last block adds the padding words
        // (0x00, 0x07 at bytes 32q, 32q + 1; 0x80 at byte 167).
        if (stage_next && blk >= next_first) {
            u64 *t = stage_next + 4 * (blk - next_first);
            t[0] = a[0];
            t[1] = a[1];
            t[2] = a[2];
            t[3] = a[3];
            if (blk + 1 == nblocks) {
                const u64 q4 = 4 * (nblocks - next_first);
                for (u64 w = q4; w < 21; w++)
                    stage_next[w] = (w == q4 ? 0x0700ull : 0) | (w == 20 ? 0x8000000000000000ull : 0);
            }
        }
    }
}

// The partial block [160 first, len) padded into 21 words (lane w builds
// word w): data bytes, 0x00 at byte l, 0x07 at byte l + 1, 0x80 at byte 167.
__global__ __launch_bounds__(64) void k12_stage_tail_kernel(const uint8_t *__restrict__ in, u64 len, u64 first,
                                                            u64 *__restrict__ stage) {
    const int w = threadIdx.x;
    if (w >= 21) return;
    const uint8_t *blk = in + first * 160;
    const u64 l = len - first * 160;
    u64 v = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const u64 j = 8 * (u64)w + b;
        if (j < l) v |= (u64)blk[j] << (8 * b);
        else if (j == l + 1) v |= (u64)0x07 << (8 * b);
    }
    if (w == 20) v ^= 0x8000000000000000ull;
    stage[w] = v;
}

// Blocks [first, nblocks) byte by byte: every block of a layer whose input
// or output is not 8-byte aligned.
__global__ __launch_bounds__(64) void k12_layer_bytes_kernel(const uint8_t *__restrict__ in, u64 len,
                                                             uint8_t *__restrict__ out, u64 first, u64 nblocks) {
    for (u64 blk = first + blockIdx.x * (u64)blockDim.x + threadIdx.x; blk < nblocks;
         blk += (u64)gridDim.x * blockDim.x) {
        u64 a[25];
        const u64 off = blk * 160;
        const u64 l = (len - off) < 160 ? (len - off) : 160;
        absorb_padded(a, in + off, l);
        keccak_p12(a);
#pragma unroll
        for (int w = 0; w < 4; w++)
#pragma unroll
            for (int b = 0; b < 8; b++) out[blk * 32 + 8 * w + b] = (uint8_t)(a[w] >> (8 * b));
    }
}

bool launch_k12_layer(const uint8_t *in, u64 len, uint8_t *out, u64 *stage_in, bool stage_ready, u64 *stage_out,
                      hipStream_t s) {
    const u64 nblocks = (len + 159) / 160;
    if (nblocks == 0) return false;
    const bool aligned = ((uintptr_t)in % 8) == 0 && ((uintptr_t)out % 8) == 0;
    if (!aligned) {  // every block byte by byte
        u64 g = (nblocks + 63) / 64;
        unsigned grid = (unsigned)(g < 65536 ? g : 65536);
        hipLaunchKernelGGL(k12_layer_bytes_kernel, dim3(grid), dim3(64), 0, s, in, len, out, (u64)0, nblocks);
        return false;
    }
    const u64 nfull = len / 160;
    if (nfull < nblocks && !stage_ready)
        hipLaunchKernelGGL(k12_stage_tail_kernel, dim3(1), dim3(64), 0, s, in, len, nfull, stage_in);
    // the next layer (32 bytes per block) has a partial last block to stage
    const u64 next_len = nblocks * 32;
    const bool produce = stage_out && next_len > 32 && next_len % 160 != 0;
    const u64 next_first = produce ? (next_len / 160) * 5 : nblocks;
    u64 g = (nblocks + 255) / 256;
    unsigned grid = (unsigned)(g < 65536 ? g : 65536);
    hipLaunchKernelGGL(k12_layer_kernel, dim3(grid), dim3(256), 0, s, (const u64 *)in, nfull, nblocks,
                       (const u64 *)stage_in, (u64 *)out, produce ? stage_out : nullptr, next_first);
    return produce;
}

// ---- Row tree: the binary tree of row hashes (no reference function).
// Leaf r = K12(row r): the row's 8 * cols bytes and the empty customization's
// length byte 0x00, absorbed 21 words (168 bytes) at a time by Keccak-p[1600,
// 12] with suffix 0x07 and 0x80 at byte 167 of the last block (TurboSHAKE128,
// D = 0x07: K12's single node, rows up to kSingleNodeCols words).  The word
// after the data holds the bytes 0x00, 0x07: the constant 0x0700.  Wider
// rows use K12's tree mode (k12_rows_cv_kernel, k12_rows_node_kernel).  A
// parent is K12(left || right), 64 bytes in one permutation; the last node
// of an odd layer moves up unchanged.

// Absorb and permute row w (cols words) into a[0..3] = its 32-byte hash.
__device__ __forceinline__ void row_leaf(const u64 *w, u32 cols, u64 *a) {
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] = 0;
    const u32 nb = cols / 21 + 1;
    for (u32 b = 0; b + 1 < nb; b++) {
        const u64 *blk = w + 21 * b;
#pragma unroll
        for (int i = 0; i < 21; i++) a[i] ^= blk[i];
        keccak_p12(a);
    }
    const u32 base = 21 * (nb - 1);
#pragma unroll
    for (int i = 0; i < 21; i++) {
        const u32 j = base + i;
        a[i] ^= j < cols ? w[j] : (j == cols ? 0x0700ull : 0);
    }
    a[20] ^= 0x8000000000000000ull;
    keccak_p12(a);
}

// h = K12(l || r) for 32-byte l, r.
__device__ __forceinline__ void pair_hash(const u64 *l, const u64 *r, u64 *h) {
    u64 a[25];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        a[i] = l[i];
        a[4 + i] = r[i];
    }
    a[8] = 0x0700ull;
#pragma unroll
    for (int i = 9; i < 25; i++) a[i] = 0;
    a[20] = 0x8000000000000000ull;
    keccak_p12(a);
#pragma unroll
    for (int i = 0; i < 4; i++) h[i] = a[i];
}

__global__ __launch_bounds__(256) void k12_rows_kernel(const u64 *__restrict__ rows, u64 nrows, u32 cols,
                                                       u64 *__restrict__ out) {
    for (u64 r = blockIdx.x * (u64)blockDim.x + threadIdx.x; r < nrows; r += (u64)gridDim.x * blockDim.x) {
        u64 a[25];
        row_leaf(rows + r * cols, cols, a);
#pragma unroll
        for (int i = 0; i < 4; i++) out[4 * r + i] = a[i];
    }
}

// Tree-mode leaves (RFC 9861 section 3.2), rows of kSingleNodeCols + 1 words
// or more: S = row bytes || 0x00 is cut into kChunkWords-word (8192-byte)
// chunks S_0 .. S_{n-1} (n = cols / kChunkWords + 1); chunk i >= 1 gives the
// chaining value CV_i = TurboSHAKE128(S_i, D = 0x0B, 32 bytes), and the leaf
// is TurboSHAKE128(S_0 || 0x03 0x00^7 || CV_1 .. CV_{n-1} ||
// length_encode(n - 1) || 0xFF 0xFF, D = 0x06).  Every piece is whole words
// (the rows are words), so each message is produced word by word.
constexpr u32 kChunkWords = 1024;

// TurboSHAKE128 of the nw words word(0 .. nw-1), the last of which already
// holds the domain byte; 0x80 goes to byte 167 of the last block.
template <class Word>
__device__ __forceinline__ void absorb_words(Word word, u32 nw, u64 *a) {
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] = 0;
    const u32 nb = (nw + 20) / 21;
    for (u32 b = 0; b < nb; b++) {
        const u32 base = 21 * b;
#pragma unroll
        for (int i = 0; i < 21; i++)
            if (base + i < nw) a[i] ^= word(base + i);
        if (b + 1 == nb) a[20] ^= 0x8000000000000000ull;
        keccak_p12(a);
    }
}

// One thread per (row, chunk c >= 1): cv[(row * (nch - 1) + c - 1) * 4 ..].
// Chunk c holds row words [c * kChunkWords, min((c + 1) * kChunkWords, cols))
// and, if it is the last, the byte 0x00 of S: its domain word is 0x0B00
// then, else 0x0B (the chunk ends on a word boundary).
__global__ __launch_bounds__(256) void k12_rows_cv_kernel(const u64 *__restrict__ rows, u64 nrows, u32 cols,
                                                          u32 nch, u64 *__restrict__ cv) {
    const u64 per = nch - 1, total = nrows * per;
    for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < total; t += (u64)gridDim.x * blockDim.x) {
        const u64 r = t / per;
        const u32 c = (u32)(t - r * per) + 1;
        const u64 *src = rows + r * cols + (u64)kChunkWords * c;
        const bool last = c + 1 == nch;
        const u32 nd = last ? cols - kChunkWords * c : kChunkWords;
        const u64 dom = last ? 0x0B00ull : 0x0Bull;
        u64 a[25];
        absorb_words([&](u32 j) { return j < nd ? src[j] : dom; }, nd + 1, a);
#pragma unroll
        for (int i = 0; i < 4; i++) cv[4 * t + i] = a[i];
    }
}

// One thread per row: the final node S_0 || 0x03 0x00^7 || CVs || tail,
// tail = length_encode(n - 1) (n - 1 < 256: two bytes), 0xFF 0xFF and the
// domain byte 0x06 in one word.
__global__ __launch_bounds__(256) void k12_rows_node_kernel(const u64 *__restrict__ rows, u64 nrows, u32 cols,
                                                            u32 nch, const u64 *__restrict__ cv,
                                                            u64 *__restrict__ out) {
    const u32 ncv = 4 * (nch - 1);
    const u64 tail = (u64)(nch - 1) | (0x01ull << 8) | (0xFFull << 16) | (0xFFull << 24) | (0x06ull << 32);
    for (u64 r = blockIdx.x * (u64)blockDim.x + threadIdx.x; r < nrows; r += (u64)gridDim.x * blockDim.x) {
        const u64 *w = rows + r * cols, *c = cv + r * ncv;
        u64 a[25];
        absorb_words(
            [&](u32 j) {
                return j < kChunkWords ? w[j] : j == kChunkWords ? 0x03ull : j <= kChunkWords + ncv ? c[j - kChunkWords - 1] : tail;
            },
            kChunkWords + 2 + ncv, a);
#pragma unroll
        for (int i = 0; i < 4; i++) out[4 * r + i] = a[i];
    }
}

// Layer of n nodes (4 words each) -> its ceil(n / 2) parents.
__global__ __launch_bounds__(256) void k12_pairs_kernel(const u64 *__restrict__ in, u64 n, u64 *__restrict__ out) {
    const u64 m = (n + 1) / 2;
    for (u64 j = blockIdx.x * (u64)blockDim.x + threadIdx.x; j < m; j += (u64)gridDim.x * blockDim.x) {
        u64 h[4];
        if (2 * j + 1 < n) {
            pair_hash(in + 8 * j, in + 8 * j + 4, h);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) h[i] = in[8 * j + i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) out[4 * j + i] = h[i];
    }
}

// The last layers in one workgroup: a layer of n <= kRowTail nodes is read
// into LDS once and every layer above it is built there (and written out)
// up to the root.  Every thread runs the same layer loop (n is uniform), so
// each barrier is reached by the whole workgroup.
constexpr u32 kRowTail = 512;
__global__ __launch_bounds__(256) void k12_pairs_tail_kernel(const u64 *__restrict__ in, u64 n,
                                                             u64 *__restrict__ out) {
    __shared__ u64 buf[2][kRowTail * 4];
    for (u32 i = threadIdx.x; i < 4 * n; i += blockDim.x) buf[0][i] = in[i];
    __syncthreads();
    int cur = 0;
    while (n > 1) {
        const u64 m = (n + 1) / 2;
        for (u32 j = threadIdx.x; j < m; j += blockDim.x) {
            const u64 *src = buf[cur];
            u64 h[4];
            if (2 * j + 1 < n) {
                pair_hash(src + 8 * j, src + 8 * j + 4, h);
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) h[i] = src[8 * j + i];
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                buf[1 - cur][4 * j + i] = h[i];
                out[4 * j + i] = h[i];
            }
        }
        __syncthreads();
        out += 4 * m;
        n = m;
        cur = 1 - cur;
    }
}

static unsigned grid_for(u64 threads) {
    u64 g = (threads + 255) / 256;
    return (unsigned)(g < 65536 ? g : 65536);
}

u64 row_tree_cv_words(u64 nrows, u32 cols) {
    return cols > kSingleNodeCols ? nrows * (u64)(cols / kChunkWords) * 4 : 0;
}

void launch_row_leaves(const u64 *rows, u64 nrows, u32 cols, u64 *out, u64 *cv, hipStream_t s) {
    if (nrows == 0) return;
    if (cols <= kSingleNodeCols) {
        hipLaunchKernelGGL(k12_rows_kernel, dim3(grid_for(nrows)), dim3(256), 0, s, rows, nrows, cols, out);
        return;
    }
    const u32 nch = cols / kChunkWords + 1;
    hipLaunchKernelGGL(k12_rows_cv_kernel, dim3(grid_for(nrows * (nch - 1))), dim3(256), 0, s, rows, nrows, cols, nch,
                       cv);
    hipLaunchKernelGGL(k12_rows_node_kernel, dim3(grid_for(nrows)), dim3(256), 0, s, rows, nrows, cols, nch,
                       (const u64 *)cv, out);
}

void launch_row_tree(const u64 *rows, u64 nrows, u32 cols, u64 *out, u64 *cv, hipStream_t s) {
    if (nrows == 0) return;
    launch_row_leaves(rows, nrows, cols, out, cv, s);
    u64 n = nrows;
    u64 *layer = out;
    while (n > kRowTail) {
        hipLaunchKernelGGL(k12_pairs_kernel, dim3(grid_for((n + 1) / 2)), dim3(256), 0, s, (const u64 *)layer, n,
                           layer + 4 * n);
        layer += 4 * n;
        n = (n + 1) / 2;
    }
    if (n > 1) hipLaunchKernelGGL(k12_pairs_tail_kernel, dim3(1), dim3(256), 0, s, (const u64 *)layer, n, layer + 4 * n);
}

// Opening i of leaf idx[i]: for layer l < depth, the sibling of node
// idx[i] >> l (4 words), or zeros where that node has none (it moves up
// unchanged) or idx[i] is not a leaf.  One thread per (opening, layer).
__global__ __launch_bounds__(256) void row_open_kernel(RowGeometry g, const u64 *__restrict__ layers, const u64 *__restrict__ idx, u64 k,
                                u64 *__restrict__ out) {
    const u64 total = k * (u64)g.depth;
    for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < total; t += (u64)gridDim.x * blockDim.x) {
        const u64 i = t / g.depth, l = t % g.depth;
        const u64 leaf = idx[i];
        const u64 sib = (leaf >> l) ^ 1;
        const bool has = leaf < g.n[0] && sib < g.n[l];
#pragma unroll
        for (int w = 0; w < 4; w++) out[4 * t + w] = has ? layers[4 * (g.off[l] + sib) + w] : 0;
    }
}

// ok[i] = 1 when leaf hash leaves[i] (of the opened row) at leaf idx[i] with
// opening paths[i] hashes up to root.
__global__ __launch_bounds__(64) void row_verify_kernel(RowGeometry g, const u64 *__restrict__ leaves,
                                                        const u64 *__restrict__ idx, u64 k,
                                                        const u64 *__restrict__ paths,
                                                        const u64 *__restrict__ root, int *__restrict__ ok) {
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < k; i += (u64)gridDim.x * blockDim.x) {
        u64 node = idx[i];
        if (node >= g.n[0]) {
            ok[i] = 0;
            continue;
        }
        u64 h[4] = {leaves[4 * i], leaves[4 * i + 1], leaves[4 * i + 2], leaves[4 * i + 3]};
        for (int l = 0; l < g.depth; l++) {
            const u64 *sib = paths + 4 * (i * (u64)g.depth + l);
            if ((node ^ 1) < g.n[l]) {
                if (node & 1) pair_hash(sib, h, h);
                else pair_hash(h, sib, h);
            }
            node >>= 1;
        }
        ok[i] = (h[0] == root[0] && h[1] == root[1] && h[2] == root[2] && h[3] == root[3]) ? 1 : 0;
    }
}

void launch_row_open(const RowGeometry &g, const u64 *layers, const u64 *idx, u64 k, u64 *out, hipStream_t s) {
    const u64 total = k * (u64)g.depth;
    if (total == 0) return;
    u64 grid = (total + 255) / 256;
    hipLaunchKernelGGL(row_open_kernel, dim3((unsigned)(grid < 65536 ? grid : 65536)), dim3(256), 0, s, g, layers, idx,
                       k, out);
}

void launch_row_verify(const RowGeometry &g, const u64 *leaves, const u64 *idx, u64 k, const u64 *paths,
                       const u64 *root, int *ok, hipStream_t s) {
    if (k == 0) return;
    u64 grid = (k + 63) / 64;
    hipLaunchKernelGGL(row_verify_kernel, dim3((unsigned)(grid < 65536 ? grid : 65536)), dim3(64), 0, s, g, leaves,
                       idx, k, paths, root, ok);
}

// Proof-of-work scan (pcs/src/pow.rs:32-60): nonces [start, start+count) are
// tested; the smallest nonce whose K12(challenge || nonce_le)[0..8] read as a
// little-endian u64 is >= threshold is folded into *best with atomicMin.
__global__ __launch_bounds__(256) void pow_kernel(const u64 *__restrict__ ch, u64 threshold, u64 start,
                                                  u64 count, unsigned long long *best) {
    u64 c0 = ch[0], c1 = ch[1], c2 = ch[2], c3 = ch[3];
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < count; i += (u64)gridDim.x * blockDim.x) {
        u64 nonce = start + i;
        if (nonce >= *(volatile unsigned long long *)best) return;
        u64 a[25];
        a[0] = c0; a[1] = c1; a[2] = c2; a[3] = c3;
        a[4] = nonce;
        a[5] = 0x0700;  // bytes 40 (0x00) and 41 (0x07)
#pragma unroll
        for (int j = 6; j < 25; j++) a[j] = 0;
        a[20] = 0x8000000000000000ull;
        keccak_p12(a);
        if (a[0] >= threshold) atomicMin(best, (unsigned long long)nonce);
    }
}

// Single K12 hash with up to 168 output bytes (k12::<H>, generic.rs:62-84).
__global__ __launch_bounds__(64) void k12_single_kernel(const uint8_t *in, u64 len, uint8_t *out, u64 outlen) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    u64 a[25];
    absorb_padded(a, in, len);
    keccak_p12(a);
#pragma unroll
    for (int w = 0; w < 21; w++)
#pragma unroll
        for (int b = 0; b < 8; b++)
            if ((u64)(8 * w + b) < outlen) out[8 * w + b] = (uint8_t)(a[w] >> (8 * b));
}

void launch_k12_single(const uint8_t *in, u64 len, uint8_t *out, u64 outlen, hipStream_t s) {
    hipLaunchKernelGGL(k12_single_kernel, dim3(1), dim3(64), 0, s, in, len, out, outlen);
}

// ---- Merkle openings.  Layer l (l = 0: the leaf bytes) has len[l] bytes and
// starts at base[l]; an opening of leaf block b holds, for every hashed layer
// l < depth, the block b_l = b / 5^l of layer l (160 bytes, zero padded past
// the end of the layer): hashing it gives entry b_l of layer l + 1, which
// sits in block b_{l+1} at byte 32 * (b_l % 5).

// out[(i * depth + l) * 160 + j], one thread per byte.
__global__ __launch_bounds__(256) void merkle_open_kernel(MerkleGeometry g, const u64 *__restrict__ blocks, u64 k, uint8_t *__restrict__ out) {
    const u64 total = k * (u64)g.depth * 160;
    for (u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x; t < total; t += (u64)gridDim.x * blockDim.x) {
        const u64 j = t % 160, il = t / 160, l = il % g.depth, i = il / g.depth;
        u64 b = blocks[i];
        for (u64 q = 0; q < l; q++) b /= 5;
        const u64 pos = b * 160 + j;
        out[t] = pos < g.len[l] ? g.base[l][pos] : 0;
    }
}

// ok[i] = 1 when opening i of leaf block blocks[i] hashes up to root32.
__global__ __launch_bounds__(64) void merkle_verify_kernel(MerkleGeometry g, const uint8_t *__restrict__ paths, const u64 *__restrict__ blocks,
                                     u64 k, const uint8_t *__restrict__ root32, int *__restrict__ ok) {
    for (u64 i = blockIdx.x * (u64)blockDim.x + threadIdx.x; i < k; i += (u64)gridDim.x * blockDim.x) {
        u64 b = blocks[i];
        bool good = b * 160 < g.len[0];
        for (int l = 0; l < g.depth && good; l++) {
            const uint8_t *blk = paths + (i * (u64)g.depth + l) * 160;
            const u64 rem = g.len[l] - b * 160, bl = rem < 160 ? rem : 160;
            u64 a[25];
            absorb_padded(a, blk, bl);
            keccak_p12(a);
            const uint8_t *want = l + 1 < g.depth ? paths + (i * (u64)g.depth + l + 1) * 160 + 32 * (b % 5) : root32;
            for (int w = 0; w < 4 && good; w++)
                for (int q = 0; q < 8; q++)
                    if ((uint8_t)(a[w] >> (8 * q)) != want[8 * w + q]) good = false;
            b /= 5;
        }
        ok[i] = good ? 1 : 0;
    }
}

void launch_merkle_open(const MerkleGeometry &g, const u64 *blocks, u64 k, uint8_t *out, hipStream_t s) {
    const u64 total = k * (u64)g.depth * 160;
    if (total == 0) return;
    u64 grid = (total + 255) / 256;
    hipLaunchKernelGGL(merkle_open_kernel, dim3((unsigned)(grid < 65536 ? grid : 65536)), dim3(256), 0, s, g, blocks,
                       k, out);
}

void launch_merkle_verify(const MerkleGeometry &g, const uint8_t *paths, const u64 *blocks, u64 k,
                          const uint8_t *root32, int *ok, hipStream_t s) {
    if (k == 0) return;
    u64 grid = (k + 63) / 64;
    hipLaunchKernelGGL(merkle_verify_kernel, dim3((unsigned)(grid < 65536 ? grid : 65536)), dim3(64), 0, s, g, paths,
                       blocks, k, root32, ok);
}

void launch_pow_search(const uint8_t *challenge32, u64 threshold, u64 start, u64 count,
                       unsigned long long *best, hipStream_t s) {
    u64 g = (count + 255) / 256;
    unsigned grid = (unsigned)(g < 16384 ? g : 16384);
    if (grid == 0) return;
    hipLaunchKernelGGL(pow_kernel, dim3(grid), dim3(256), 0, s, (const u64 *)challenge32, threshold,
                       start, count, best);
}

}  // namespace gl
