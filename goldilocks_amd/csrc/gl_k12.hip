// Keccak-p[1600,12] / single-block KangarooTwelve layer hashing for gfx950.
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

__device__ __constant__ u64 kRC[12] = {
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull,
    0x8000000000008003ull, 0x8000000000008002ull, 0x8000000000000080ull,
    0x000000000000800aull, 0x800000008000000aull, 0x8000000080008081ull,
    0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull,
};

__device__ __forceinline__ u64 rotl64(u64 x, int r) { return (x << r) | (x >> (64 - r)); }

// 12 rounds of theta, rho+pi, chi, iota on lanes a[x + 5y].
__device__ __forceinline__ void keccak_p12(u64 *a) {
#pragma unroll 1
    for (int round = 0; round < 12; round++) {
        u64 c0 = a[0] ^ a[5] ^ a[10] ^ a[15] ^ a[20];
        u64 c1 = a[1] ^ a[6] ^ a[11] ^ a[16] ^ a[21];
        u64 c2 = a[2] ^ a[7] ^ a[12] ^ a[17] ^ a[22];
        u64 c3 = a[3] ^ a[8] ^ a[13] ^ a[18] ^ a[23];
        u64 c4 = a[4] ^ a[9] ^ a[14] ^ a[19] ^ a[24];
        u64 d0 = c4 ^ rotl64(c1, 1), d1 = c0 ^ rotl64(c2, 1), d2 = c1 ^ rotl64(c3, 1);
        u64 d3 = c2 ^ rotl64(c4, 1), d4 = c3 ^ rotl64(c0, 1);
        // theta + rho + pi: b[y, 2x+3y] = rot(a[x, y] ^ d[x], r[x, y])
        u64 b00 = a[0] ^ d0;
        u64 b10 = rotl64(a[6] ^ d1, 44);
        u64 b20 = rotl64(a[12] ^ d2, 43);
        u64 b30 = rotl64(a[18] ^ d3, 21);
        u64 b40 = rotl64(a[24] ^ d4, 14);
        u64 b01 = rotl64(a[3] ^ d3, 28);
        u64 b11 = rotl64(a[9] ^ d4, 20);
        u64 b21 = rotl64(a[10] ^ d0, 3);
        u64 b31 = rotl64(a[16] ^ d1, 45);
        u64 b41 = rotl64(a[22] ^ d2, 61);
        u64 b02 = rotl64(a[1] ^ d1, 1);
        u64 b12 = rotl64(a[7] ^ d2, 6);
        u64 b22 = rotl64(a[13] ^ d3, 25);
        u64 b32 = rotl64(a[19] ^ d4, 8);
        u64 b42 = rotl64(a[20] ^ d0, 18);
        u64 b03 = rotl64(a[4] ^ d4, 27);
        u64 b13 = rotl64(a[5] ^ d0, 36);
        u64 b23 = rotl64(a[11] ^ d1, 10);
        u64 b33 = rotl64(a[17] ^ d2, 15);
        u64 b43 = rotl64(a[23] ^ d3, 56);
        u64 b04 = rotl64(a[2] ^ d2, 62);
        u64 b14 = rotl64(a[8] ^ d3, 55);
        u64 b24 = rotl64(a[14] ^ d4, 39);
        u64 b34 = rotl64(a[15] ^ d0, 41);
        u64 b44 = rotl64(a[21] ^ d1, 2);
        // chi (row y holds b0y..b4y as x = 0..4) + iota
        a[0] = b00 ^ (~b10 & b20) ^ kRC[round];
        a[1] = b10 ^ (~b20 & b30);
        a[2] = b20 ^ (~b30 & b40);
        a[3] = b30 ^ (~b40 & b00);
        a[4] = b40 ^ (~b00 & b10);
        a[5] = b01 ^ (~b11 & b21);
        a[6] = b11 ^ (~b21 & b31);
        a[7] = b21 ^ (~b31 & b41);
        a[8] = b31 ^ (~b41 & b01);
        a[9] = b41 ^ (~b01 & b11);
        a[10] = b02 ^ (~b12 & b22);
        a[11] = b12 ^ (~b22 & b32);
        a[12] = b22 ^ (~b32 & b42);
        a[13] = b32 ^ (~b42 & b02);
        a[14] = b42 ^ (~b02 & b12);
        a[15] = b03 ^ (~b13 & b23);
        a[16] = b13 ^ (~b23 & b33);
        a[17] = b23 ^ (~b33 & b43);
        a[18] = b33 ^ (~b43 & b03);
        a[19] = b43 ^ (~b03 & b13);
        a[20] = b04 ^ (~b14 & b24);
        a[21] = b14 ^ (~b24 & b34);
        a[22] = b24 ^ (~b34 & b44);
        a[23] = b34 ^ (~b44 & b04);
        a[24] = b44 ^ (~b04 & b14);
    }
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

// out[32i .. 32i+32] = K12(in[160i .. min(160i+160, len)]).
__global__ __launch_bounds__(256) void k12_layer_kernel(const uint8_t *__restrict__ in, u64 len,
                                                        uint8_t *__restrict__ out, u64 nblocks,
                                                        int aligned) {
    for (u64 blk = blockIdx.x * (u64)blockDim.x + threadIdx.x; blk < nblocks;
         blk += (u64)gridDim.x * blockDim.x) {
        u64 a[25];
        const u64 off = blk * 160;
        const u64 l = (len - off) < 160 ? (len - off) : 160;
        if (l == 160 && aligned) {
            const u64 *w = (const u64 *)(in + off);
#pragma unroll
            for (int i = 0; i < 20; i++) a[i] = w[i];
            a[20] = 0x8000000000000700ull;  // 0x00, 0x07 at bytes 160,161; 0x80 at 167
#pragma unroll
            for (int i = 21; i < 25; i++) a[i] = 0;
        } else {
            absorb_padded(a, in + off, l);
        }
        keccak_p12(a);
        u64 *o = (u64 *)(out + blk * 32);
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2];
        o[3] = a[3];
    }
}

void launch_k12_layer(const uint8_t *in, u64 len, uint8_t *out, hipStream_t s) {
    u64 nblocks = (len + 159) / 160;
    if (nblocks == 0) return;
    u64 g = (nblocks + 255) / 256;
    unsigned grid = (unsigned)(g < 65536 ? g : 65536);
    int aligned = ((uintptr_t)in % 8) == 0;
    hipLaunchKernelGGL(k12_layer_kernel, dim3(grid), dim3(256), 0, s, in, len, out, nblocks, aligned);
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
__global__ void k12_single_kernel(const uint8_t *in, u64 len, uint8_t *out, u64 outlen) {
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

void launch_pow_search(const uint8_t *challenge32, u64 threshold, u64 start, u64 count,
                       unsigned long long *best, hipStream_t s) {
    u64 g = (count + 255) / 256;
    unsigned grid = (unsigned)(g < 16384 ? g : 16384);
    if (grid == 0) return;
    hipLaunchKernelGGL(pow_kernel, dim3(grid), dim3(256), 0, s, (const u64 *)challenge32, threshold,
                       start, count, best);
}

}  // namespace gl
