// Internal declarations shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gl {

typedef uint64_t u64;
typedef uint32_t u32;

// One Stockham decimation-in-frequency pass of radix R = R1 * R2.
//
// The array holds `count` independent items of `bstride` words.  Inside an
// item the pass sees S interleaved sub-problems of length M = R * Mp; word
// (s, n) of sub-problem s sits at s + S*n.  Instance i = s + S*m (m < Mp)
// gathers x_r = in[i + S*Mp*r] (r < R), computes y = DFT_R(x), multiplies
// y_k by w_M^(m*k) and scatters to out[s + S*k + S*R*m].  After the last
// pass (Mp = 1) the item is in natural order (a standard Stockham autosort
// ordering, so no separate transpose pass is needed).
struct PassArgs {
    const u64 *in;
    u64 *out;
    u64 bstride;       // words per batch item
    u64 S;             // interleaved sub-problems
    u64 SMp;           // S * Mp: input stride between r and r+1
    u64 n_rev;         // transform length when reversing the input (intt)
    u32 ninst;         // S * Mp instances per item
    u32 total_inst;    // count * ninst
    u32 Mp;
    u32 lo_bits;       // output twiddle w_M^e = lo[e & mask] * hi[e >> lo_bits]
    const u64 *tw_inner;  // w_R^e, e < R
    const u64 *tw_lo;
    const u64 *tw_hi;
    u64 scale;         // multiplier applied to every output when FLAG_SCALE or FLAG_SCALE_TW
    const u64 *coset;  // FLAG_COSET: input row j < coset_rows is multiplied by coset[j]
    u64 coset_rows;    //   (Montgomery g^j); rows >= coset_rows read as zero (also FLAG_ZEXT)
    const u64 *row_scale;  // FLAG_ROW_SCALE: output row r is multiplied by row_scale[r]
    u64 SS;            //   S / S0: rows between consecutive k (row of position q is q / S0)
    u32 flags;
    // x / ninst, x / S and x / S0 for x < 2^31 as (x * mul) >> sh (fastdiv_params)
    u32 ninst_mul, ninst_sh, S_mul, S_sh, S0_mul, S0_sh;
};

// Multiply-shift constants for dividing any x < 2^31 by d >= 1: with
// l = ceil(log2 d), mul = ceil(2^(31+l) / d) < 2^32 and x * mul / 2^(31+l)
// exceeds x / d by less than 1/d, so the floor is exact.  d >= 2^32 gives
// mul = 0 (quotient 0 for every such x).
inline void fastdiv_params(u64 d, u32 &mul, u32 &sh) {
    if (d >= (1ull << 32)) {
        mul = 0;
        sh = 0;
        return;
    }
    const u32 l = d <= 1 ? 0 : 64 - (u32)__builtin_clzll(d - 1);  // ceil(log2 d) <= 32
    const u64 num = 1ull << (31 + l);                               // <= 2^63: no 128-bit division
    mul = (u32)((num + d - 1) / d);
    sh = 31 + l;
}

__host__ __device__ __forceinline__ u32 fastdiv(u32 x, u32 mul, u32 sh) { return (u32)(((u64)x * mul) >> sh); }

// Instance I of a pass -> batch item b, instance i within the item, and its
// (s, m) with i = s + S*m.
__device__ __forceinline__ void locate(const PassArgs &a, u32 I, u32 &b, u32 &i, u32 &s, u32 &m) {
    b = fastdiv(I, a.ninst_mul, a.ninst_sh);
    i = I - b * a.ninst;
    m = fastdiv(i, a.S_mul, a.S_sh);
    s = i - m * (u32)a.S;
}

enum : u32 {
    FLAG_TWIDDLE = 1u,  // apply w_M^(m*k) to the outputs (Mp > 1)
    FLAG_REVERSE = 2u,  // read x[(n - j) mod n] (first pass of intt)
    FLAG_SCALE = 4u,    // multiply outputs by `scale` (1/n for intt)
    FLAG_TWIDDLE_FULL = 8u,  // tw_lo is a full table: output k of instance m uses tw_lo[k + R*m]
    FLAG_COSET = 16u,  // first pass of the commit's extension: coset shift + zero rows
    FLAG_ZEXT = 32u,   // first pass: rows >= coset_rows read as zero (no shift)
    FLAG_ROW_SCALE = 64u,  // last pass: output row r times row_scale[r] (replaces FLAG_SCALE)
    FLAG_ZEXT_HALF = 128u,  // with FLAG_ZEXT: coset_rows = M/2, so inputs r >= R/2 are the zero ones
    FLAG_SCALE_TW = 256u,   // with FLAG_TWIDDLE: the recurrence starts at scale * w^(m*k0)
};
// All twiddle tables and `scale` hold Montgomery forms (w * 2^64 mod p) and
// are applied with mont_mul.

// Which inner twiddle table a variant reads from PassArgs::tw_inner.
enum : int {
    VK_TWO_STAGE = 0,    // R1 x R2 (or one codelet): entry rb*R1 + ka = w_R^(ka*rb)
    VK_THREE_STAGE = 1,  // R1 x R2 x R3: entry e = w_R^e
    VK_RADER257 = 2,     // radix-257 Rader tables
    VK_COMPOSITE = 3,    // radix-65537 Rader: several kernels, no single launch
};

struct PassVariant {
    int R1, R2, T, NT;
    size_t lds_bytes;
    void (*launch)(const PassArgs &, unsigned grid, hipStream_t);
    // descs/starts in device memory; descriptor j owns blocks [starts[j], starts[j+1])
    void (*launch_grouped)(const PassArgs *descs, const u32 *starts, u32 nd, unsigned grid, hipStream_t);
    int kind;
    int R3;  // three-stage variants: last radix (R2 then holds R2 * R3); else 1
};

// All compiled pass variants (gl_ntt_kernels.hip).
const PassVariant *pass_variants(int *count);

// Radix-257 Rader pass (gl_rader.hip).
const PassVariant *rader257_variant();

// Elementwise kernels (gl_ntt_kernels.hip).
void launch_fill_random(u64 *v, u64 n, u64 seed, hipStream_t s);
// Permutations (gl_permute.hip): out (cols x rows) = transpose of in (rows x
// cols), in != out; an n x n matrix transposed in place; v[i] <- v[rev(i)]
// for n = 2^k in place.
void launch_transpose(const u64 *in, u64 *out, u64 rows, u64 cols, hipStream_t s);
void launch_transpose_square(u64 *m, u64 n, hipStream_t s);
void launch_bit_reverse(u64 *v, u64 n, int k, hipStream_t s);
void launch_twiddle_matrix(u64 *out, u64 count, u64 R, const u64 *lo, const u64 *hi, u32 lo_bits,
                           hipStream_t s);
void launch_rader_big_gather(const PassArgs &a, const u32 *permk, u64 *C, u64 *x0, hipStream_t s);
void launch_rader_big_pointwise(u64 *C, u64 *x0, const u64 *spec, u64 ninst, hipStream_t s);
void launch_rader_big_scatter(const PassArgs &a, const u32 *permi, const u64 *D, const u64 *X0, hipStream_t s);
void launch_pointwise_mont(u64 *a, const u64 *b, u64 n, hipStream_t s);
void launch_power_series(u64 *out, u64 count, const u64 *lo, const u64 *hi, u32 lo_bits, hipStream_t s);
// out[j] = in[j] * scale (Montgomery product; scale in Montgomery form), j < count.
void launch_scale_table(u64 *out, const u64 *in, u64 count, u64 scale, hipStream_t s);
void launch_coset_extend(const u64 *coef, u64 *ext, u64 rows, u64 cols, u64 blowup,
                         const u64 *pw_lo, const u64 *pw_hi, u32 lo_bits, hipStream_t s);

// K12 kernels (gl_k12.hip).
// Hash one layer (len bytes at in) into out.  stage_in: 21 words of device
// scratch holding this layer's padded partial last block; when stage_ready
// is false it is built first.  stage_out (21 words, distinct from
// stage_in, or null): the kernel also writes the next layer's padded
// partial last block there.  Returns true when it did (the next call can
// pass stage_out as its stage_in with stage_ready = true).
bool launch_k12_layer(const uint8_t *in, u64 len, uint8_t *out, u64 *stage_in, bool stage_ready, u64 *stage_out,
                      hipStream_t s);
void launch_k12_single(const uint8_t *in, u64 len, uint8_t *out, u64 outlen, hipStream_t s);

// Merkle tree geometry for openings: layer l has len[l] bytes at base[l]
// (l = 0: the leaves); depth = number of hashed layers (the root layer,
// 32 bytes, is layer depth).
constexpr int kMaxMerkleDepth = 48;
struct MerkleGeometry {
    const uint8_t *base[kMaxMerkleDepth + 1];
    u64 len[kMaxMerkleDepth + 1];
    int depth;
};
void launch_merkle_open(const MerkleGeometry &g, const u64 *blocks, u64 k, uint8_t *out, hipStream_t s);
void launch_merkle_verify(const MerkleGeometry &g, const uint8_t *paths, const u64 *blocks, u64 k,
                          const uint8_t *root32, int *ok, hipStream_t s);
// Row tree (the binary tree of row hashes): leaves = K12 of each row (K12's
// single node up to kSingleNodeCols words, its tree mode up to
// kRowTreeMaxCols, whose chaining values need row_tree_cv_words scratch
// words at cv), then the layers of parents up to the root, all written to
// out (4 words a node).  Layer l has n[l] nodes starting at node off[l] of
// that output; depth = hashed layers above the leaves.
constexpr unsigned kSingleNodeCols = 1023;
constexpr unsigned kRowTreeMaxCols = 256 * 1024 - 1;  // at most 255 chunk chaining values
constexpr int kMaxRowDepth = 64;
struct RowGeometry {
    u64 off[kMaxRowDepth + 1];
    u64 n[kMaxRowDepth + 1];
    int depth;
};
u64 row_tree_cv_words(u64 nrows, u32 cols);
void launch_row_leaves(const u64 *rows, u64 nrows, u32 cols, u64 *out, u64 *cv, hipStream_t s);
void launch_row_tree(const u64 *rows, u64 nrows, u32 cols, u64 *out, u64 *cv, hipStream_t s);
void launch_row_open(const RowGeometry &g, const u64 *layers, const u64 *idx, u64 k, u64 *out, hipStream_t s);
void launch_row_verify(const RowGeometry &g, const u64 *leaves, const u64 *idx, u64 k, const u64 *paths,
                       const u64 *root, int *ok, hipStream_t s);
void launch_pow_search(const uint8_t *challenge32, u64 threshold, u64 start, u64 count,
                       unsigned long long *best, hipStream_t s);

}  // namespace gl
