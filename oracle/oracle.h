/*
 * CPU oracle for the Goldilocks NTT and K12 Merkle path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * (recmo/goldilocks, Rust) used as the correctness checker by tests/, by
 * __graft_entry__.smoke() and as the CPU baseline leg of bench.py.  Nothing in
 * the product library (goldilocks_amd/) links, loads or calls it.
 *
 * Conventions
 *   - A "Field" word is the raw u64 the reference stores (Montgomery form,
 *     canonical in [0,p)).  Every transform here is linear, so transforming
 *     raw words with standard arithmetic and canonical roots reproduces the
 *     reference's raw output words exactly (ntt/src/field/mod.rs:8-50).
 *   - Roots: w_n = 2717^((p-1)/n) (ntt/src/field/mod.rs:42-50).
 *   - ntt computes X[k] = sum_j x[j] w_n^(jk) in natural order
 *     (ntt/src/ntt/naive.rs:3-22); intt scales by 1/n, reverses [1..] and runs
 *     the forward transform (ntt/src/ntt/mod.rs:104-120).
 *
 * Pinning status (what checks each function against something other than
 * itself) is listed next to each declaration; the checks live in
 * tests/test_oracle_*.py.
 */
#ifndef GOLDILOCKS_ORACLE_H
#define GOLDILOCKS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 0xffffffff00000001ULL
#define OR_GENERATOR 2717ULL

/* ---- field (ntt/src/field/algo/generic.rs) ---------------------------------
 * Pinned: random + edge operands vs Python big-int closed forms. */
uint64_t or_add(uint64_t a, uint64_t b);            /* generic.rs:17-26 */
uint64_t or_sub(uint64_t a, uint64_t b);            /* generic.rs:31-39 */
uint64_t or_mul(uint64_t a, uint64_t b);            /* generic.rs:44-48,67-108 */
uint64_t or_mont_mul(uint64_t a, uint64_t b);       /* generic.rs:52-54,114-124 */
uint64_t or_pow(uint64_t a, uint64_t e);            /* standard-form a^e */
uint64_t or_inv(uint64_t a);                        /* generic.rs:220-271 (0 -> 0) */
uint64_t or_shift(uint64_t a, uint64_t s);          /* generic.rs:171-193, a*2^s */
uint64_t or_root_384(uint64_t a, uint64_t i);       /* generic.rs:198-208 */
uint64_t or_root(uint64_t order);                   /* field/mod.rs:42-50, canonical; 0 if none */

/* ---- divisors (ntt/src/divisors.rs) ----------------------------------------
 * Pinned: reference KATs divisors().len()==1056, next(1000)=1020,
 * next(1234)=1280 (divisors.rs:101-115) and brute-force split. */
int or_is_divisor(uint64_t n);                      /* divisors.rs:21-23 */
int or_is_smooth(uint64_t n);                       /* divisors.rs:26-28 */
size_t or_divisors(uint64_t *out);                  /* divisors.rs:32-47, out may be NULL */
uint64_t or_next(uint64_t n);                       /* divisors.rs:52-59 */
uint64_t or_split(uint64_t n);                      /* divisors.rs:67-89 */

/* ---- NTT (ntt/src/ntt/ mod, naive, cooley_tukey, good_thomas, rader) -------------------------------------------------
 * Return 0 on success, -1 on an unsupported size (where the reference
 * panics, ntt/mod.rs:55-58).
 * Pinned: naive vs Python big-int DFT; fast vs naive for every legal
 * n <= 4096; delta/constant closed forms; sampled X[k] at large n. */
int or_naive_ntt(uint64_t *v, size_t n);            /* naive.rs:3-22 */
int or_naive_intt(uint64_t *v, size_t n);           /* naive.rs:24-49 */
int or_ntt(uint64_t *v, size_t n);                  /* mod.rs:98-102 */
int or_intt(uint64_t *v, size_t n);                 /* mod.rs:104-120 */
int or_ntt_batch(uint64_t *v, size_t n, size_t count, int inverse); /* strategy(n).ntt */
/* convolve.rs:13-20.  Pinned: the KAT [1,2,3]*[4,5,6] = [31,31,28]
 * (convolve.rs:44-60) and the naive cyclic sum for small n. */
int or_convolve(uint64_t *a, const uint64_t *b_ntt, size_t n);
/* X[k] of the forward transform, evaluated directly in O(n). */
uint64_t or_eval_dft(const uint64_t *x, size_t n, uint64_t k);
/* Host thread count used by or_ntt/or_ntt_batch outer loops, the commit's
 * column loop and each Merkle layer (default 1). */
void or_set_threads(int threads);
/* fn(ctx, i) for i < count on up to or_set_threads() threads (serial below
 * 2^15 units of work, or when already inside a parallel loop). */
void or_par_for(size_t count, size_t work_per_item, void (*fn)(void *, size_t), void *ctx);

/* ---- K12 / Merkle (pcs/src/k12/{generic,mod}.rs) ----------------------------
 * Pinned: 24-round permutation vs hashlib.sha3_256; XKCP KATs
 * (generic.rs:113-160); PoW KATs (pow.rs:67-106, aarch64.rs:234-240). */
void or_keccak_p1600(uint64_t st[25], int rounds);  /* generic.rs:25-60 (rounds=12) */
int or_k12(const uint8_t *in, size_t len, uint8_t *out, size_t outlen); /* generic.rs:62-84 */
void or_sha3_256(const uint8_t *in, size_t len, uint8_t out[32]); /* pinning aid */
/* Keccak sponge, byte-granular padding (suffix byte, 0x80 at rate - 1);
 * outlen <= rate.  Pinned: SHAKE128 / SHA3-256 parameters vs hashlib. */
int or_keccak_sponge(const uint8_t *in, size_t len, size_t rate, uint8_t suffix, int rounds,
                     uint8_t *out, size_t outlen);
/* KangarooTwelve, empty customization, any length (single node up to 8191
 * bytes, tree mode above); equals or_k12 up to 166 bytes.  Pinned: or_k12
 * for len <= 166 and the XKCP / RFC 9861 vectors for 17^2, 17^3, 8191,
 * 8192, 17^4 and 17^5 pattern bytes. */
int or_k12_long(const uint8_t *in, size_t len, uint8_t *out, size_t outlen);
size_t or_next_layer_size(size_t len);              /* mod.rs:33-35 */
void or_process_layer(const uint8_t *in, size_t len, uint8_t *out); /* generic.rs:86-94 */
/* All layers concatenated in order (mod.rs:41-60); *out_len = total bytes.
 * Returns -1 if cap is too small (then *out_len holds the required size). */
int or_merkle_tree(const uint8_t *in, size_t len, uint8_t *out, size_t cap, size_t *out_len);
size_t or_merkle_tree_size(size_t len);
/* Row tree (no reference function; the optional binary tree of row hashes):
 * leaf r = or_k12_long(row r bytes), node = or_k12(left || right), the last
 * node of an odd layer moves up unchanged; every layer, leaves first.
 * Returns -1 for empty shapes or cap too small.  Pinned by a
 * Python assembly of the layers over or_k12_long / or_k12. */
int or_row_tree(const uint64_t *rows, size_t nrows, size_t cols, uint8_t *out, size_t cap, size_t *out_len);
size_t or_row_tree_size(size_t leaves);
int or_pow_verify(const uint8_t challenge[32], uint64_t threshold, uint64_t nonce); /* pow.rs:32-37 */
uint64_t or_pow_solve(const uint8_t challenge[32], uint64_t threshold);            /* pow.rs:39-60 */

/* ---- commit (no reference function; composed from reference primitives) ----
 * Column-wise coset LDE (intt, zero-extend to blowup*rows, multiply
 * coefficient i by 2717^i, ntt) of a row-major rows x cols matrix, then
 * merkle_tree over the row-major LDE bytes.  lde_out: (blowup*rows)*cols
 * words; layers_out: or_merkle_tree layout.  Pinned through its parts plus a
 * direct polynomial-evaluation check at tiny sizes. */
int or_commit(const uint64_t *m, size_t rows, size_t cols, unsigned blowup,
              uint64_t *lde_out, uint8_t *layers_out, size_t cap, size_t *out_len);
/* The same extension with the row tree over the LDE rows (or_row_tree). */
int or_commit_rows(const uint64_t *m, size_t rows, size_t cols, unsigned blowup,
                   uint64_t *lde_out, uint8_t *layers_out, size_t cap, size_t *out_len);

/* ---- permutations (ntt/src/permute/mod.rs, permutation.rs) ------------------
 * Pinned: the reference's KATs (mod.rs test_copy_2_3 / test_copy_3_3,
 * permutation.rs test_transpose) and numpy's transpose / an independent
 * bit-reversal on random shapes (tests/test_oracle_permute.py). */
size_t or_transpose_index(size_t width, size_t height, size_t i);  /* permutation.rs:1-4 */
int or_transpose(uint64_t *m, size_t rows, size_t cols);         /* mod.rs:53-67 */
size_t or_bit_reverse_index(size_t size, size_t i);              /* permutation.rs:9-16 */
int or_bit_reverse(uint64_t *v, size_t n);                        /* mod.rs:71-74; -1 unless a power of two */

/* ---- seeded inputs ----------------------------------------------------------
 * Word i = the first splitmix64 output of the stream seeded with
 * seed + i*0x9e3779b97f4a7c15 that is < p (re-mixing on rejection), so any
 * element can be generated independently (the GPU library uses the same map). */
void or_fill_random(uint64_t *v, size_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
