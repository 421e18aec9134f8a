/*
 * Oracle: Keccak-p[1600,12], single-block KangarooTwelve, layer hashing and
 * the 5-ary byte Merkle tree, plus the PoW check built on them; the Keccak
 * sponge, single-node KangarooTwelve of longer messages and the binary row
 * tree (no reference function) that the row-tree commit uses.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates pcs/src/k12/generic.rs:4-94, pcs/src/k12/mod.rs:28-60 and
 * pcs/src/pow.rs:32-60.  The round constants and the rho/pi lane walk are the
 * Keccak specification's tables (FIPS 202), which the reference also uses.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* FIPS 202 round constants; Keccak-p[1600,12] uses the last 12
 * (generic.rs:4-17). */
static const uint64_t RC24[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL,
    0x8000000080008000ULL, 0x000000000000808bULL, 0x0000000080000001ULL,
    0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008aULL,
    0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL,
    0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
    0x000000000000800aULL, 0x800000008000000aULL, 0x8000000080008081ULL,
    0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL,
};
/* generic.rs:18-23 */
static const int PI[24] = {10, 7, 11, 17, 18, 3, 5, 16, 8, 21, 24, 4,
                           15, 23, 19, 13, 12, 2, 20, 14, 22, 9, 6, 1};
static const int RHO[24] = {1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 2, 14,
                            27, 41, 56, 8, 25, 43, 62, 18, 39, 61, 20, 44};

static uint64_t rotl(uint64_t x, int r) { return r ? (x << r) | (x >> (64 - r)) : x; }

/* generic.rs:25-60 for rounds = 12; rounds = 24 gives Keccak-f[1600]. */
void or_keccak_p1600(uint64_t st[25], int rounds) {
    uint64_t bc[5];
    for (int round = 24 - rounds; round < 24; round++) {
        for (int x = 0; x < 5; x++) bc[x] = st[x] ^ st[x + 5] ^ st[x + 10] ^ st[x + 15] ^ st[x + 20];
        for (int x = 0; x < 5; x++) {
            uint64_t t = bc[(x + 4) % 5] ^ rotl(bc[(x + 1) % 5], 1);
            for (int y = 0; y < 25; y += 5) st[y + x] ^= t;
        }
        uint64_t last = st[1];
        for (int x = 0; x < 24; x++) {
            uint64_t t = st[PI[x]];
            st[PI[x]] = rotl(last, RHO[x]);
            last = t;
        }
        for (int y = 0; y < 25; y += 5) {
            for (int x = 0; x < 5; x++) bc[x] = st[y + x];
            for (int x = 0; x < 5; x++) st[y + x] = bc[x] ^ ((~bc[(x + 1) % 5]) & bc[(x + 2) % 5]);
        }
        st[0] ^= RC24[round];
    }
}

static void load_state(uint64_t st[25], const uint8_t bytes[200]) {
    for (int i = 0; i < 25; i++) {
        uint64_t w = 0;
        for (int b = 7; b >= 0; b--) w = (w << 8) | bytes[8 * i + b];
        st[i] = w;
    }
}

static void store_state(const uint64_t st[25], uint8_t bytes[200]) {
    for (int i = 0; i < 25; i++)
        for (int b = 0; b < 8; b++) bytes[8 * i + b] = (uint8_t)(st[i] >> (8 * b));
}

/* generic.rs:62-84: single-block K12 with the empty customization string. */
int or_k12(const uint8_t *in, size_t len, uint8_t *out, size_t outlen) {
    if (len > 166 || outlen > 168) return -1;
    uint8_t s[200];
    memset(s, 0, sizeof s);
    memcpy(s, in, len);
    s[len] = 0x00;      /* length_encode(0) of the customization string */
    s[len + 1] = 0x07;  /* single-node suffix */
    s[167] ^= 0x80;     /* final padding bit, rate 168 */
    uint64_t st[25];
    load_state(st, s);
    or_keccak_p1600(st, 12);
    store_state(st, s);
    memcpy(out, s, outlen);
    return 0;
}

/* The Keccak sponge over Keccak-p[1600, rounds] with byte-granular input:
 * absorb `rate`-byte blocks, then the last partial block followed by the
 * byte `suffix` (domain bits plus the first padding bit), zeros, and 0x80
 * XORed into byte rate - 1 (pad10*1); squeeze outlen <= rate bytes.
 * SHA3-256 is (rate 136, suffix 0x06, 24 rounds) and SHAKE128 (168, 0x1f,
 * 24) in FIPS 202, which pins this function against hashlib; single-node
 * KangarooTwelve is (168, 0x07, 12) over M || 0x00 (the KangarooTwelve
 * specification, RFC 9861 section 3). */
int or_keccak_sponge(const uint8_t *in, size_t len, size_t rate, uint8_t suffix, int rounds,
                     uint8_t *out, size_t outlen) {
    if (rate == 0 || rate > 200 || rate % 8 != 0 || outlen > rate || rounds < 1 || rounds > 24) return -1;
    uint64_t st[25] = {0}, t[25];
    uint8_t block[200];
    for (;;) {
        const size_t take = len < rate ? len : rate;
        memset(block, 0, sizeof block);
        memcpy(block, in, take);
        const int last = len < rate;
        if (last) {
            block[take] ^= suffix;
            block[rate - 1] ^= 0x80;
        }
        load_state(t, block);
        for (size_t i = 0; i < rate / 8; i++) st[i] ^= t[i];
        or_keccak_p1600(st, rounds);
        if (last) break;
        in += rate;
        len -= rate;
    }
    store_state(st, block);
    memcpy(out, block, outlen);
    return 0;
}

/* SHA3-256: used only to pin the permutation against a library. */
void or_sha3_256(const uint8_t *in, size_t len, uint8_t out[32]) {
    or_keccak_sponge(in, len, 136, 0x06, 24, out, 32);
}

/* length_encode (RFC 9861 section 3.3): x as big-endian bytes without
 * leading zeros, then the count of those bytes (0 -> the single byte 0x00). */
static size_t length_encode(size_t x, uint8_t out[9]) {
    uint8_t be[8];
    size_t n = 0;
    for (; x; x >>= 8) be[n++] = (uint8_t)x;
    for (size_t i = 0; i < n; i++) out[i] = be[n - 1 - i];
    out[n] = (uint8_t)n;
    return n + 1;
}

/* KangarooTwelve with the empty customization string (RFC 9861 section 3):
 * S = M || length_encode(0) = M || 0x00.  Up to 8192 bytes of S (len <=
 * 8191) it is TurboSHAKE128(S, D = 0x07), for len <= 166 the single
 * permutation of or_k12 (generic.rs:62-84); longer S is cut into 8192-byte
 * chunks, chunk i >= 1 gives CV_i = TurboSHAKE128(S_i, D = 0x0B, 32 bytes)
 * and the result is TurboSHAKE128(S_0 || 0x03 0x00^7 || CV_1 .. CV_{n-1} ||
 * length_encode(n-1) || 0xFF 0xFF, D = 0x06). */
int or_k12_long(const uint8_t *in, size_t len, uint8_t *out, size_t outlen) {
    if (outlen > 168) return -1;
    const size_t slen = len + 1, chunk = 8192;
    uint8_t *S = (uint8_t *)malloc(slen);
    if (!S) return -1;
    if (len) memcpy(S, in, len);
    S[len] = 0x00; /* length_encode(0) of the customization string */
    int rc;
    if (slen <= chunk) {
        rc = or_keccak_sponge(S, slen, 168, 0x07, 12, out, outlen);
    } else {
        const size_t nch = (slen + chunk - 1) / chunk;
        uint8_t enc[9];
        const size_t el = length_encode(nch - 1, enc);
        const size_t nlen = chunk + 8 + 32 * (nch - 1) + el + 2;
        uint8_t *node = (uint8_t *)malloc(nlen);
        if (!node) {
            free(S);
            return -1;
        }
        memcpy(node, S, chunk);
        node[chunk] = 0x03;
        memset(node + chunk + 1, 0, 7);
        rc = 0;
        for (size_t i = 1; i < nch && !rc; i++) {
            const size_t off = i * chunk, l = slen - off < chunk ? slen - off : chunk;
            rc = or_keccak_sponge(S + off, l, 168, 0x0B, 12, node + chunk + 8 + 32 * (i - 1), 32);
        }
        memcpy(node + chunk + 8 + 32 * (nch - 1), enc, el);
        node[nlen - 2] = 0xFF;
        node[nlen - 1] = 0xFF;
        if (!rc) rc = or_keccak_sponge(node, nlen, 168, 0x06, 12, out, outlen);
        free(node);
    }
    free(S);
    return rc;
}

#define BLOCK_SIZE 160
#define HASH_SIZE 32

/* mod.rs:33-35 */
size_t or_next_layer_size(size_t len) { return (len + BLOCK_SIZE - 1) / BLOCK_SIZE * HASH_SIZE; }

typedef struct {
    const uint8_t *in;
    size_t len;
    uint8_t *out;
} LayerCtx;

/* Blocks [64 i, 64 i + 64) of a layer. */
static void layer_range(void *arg, size_t i) {
    const LayerCtx *cx = (const LayerCtx *)arg;
    for (size_t b = 64 * i; b < 64 * i + 64 && b * BLOCK_SIZE < cx->len; b++) {
        size_t l = cx->len - b * BLOCK_SIZE;
        if (l > BLOCK_SIZE) l = BLOCK_SIZE;
        or_k12(cx->in + b * BLOCK_SIZE, l, cx->out + b * HASH_SIZE, HASH_SIZE);
    }
}

/* generic.rs:86-94: out[32i..32i+32] = K12(in[160i..min(160i+160,len)]).
 * Blocks are independent, so ranges of 64 run in parallel (or_set_threads). */
void or_process_layer(const uint8_t *in, size_t len, uint8_t *out) {
    size_t blocks = (len + BLOCK_SIZE - 1) / BLOCK_SIZE;
    LayerCtx cx = {in, len, out};
    or_par_for((blocks + 63) / 64, 64 * 1024, layer_range, &cx);
}

size_t or_merkle_tree_size(size_t len) {
    size_t total = 0;
    while (len > HASH_SIZE) {
        len = or_next_layer_size(len);
        total += len;
    }
    return total;
}

/* mod.rs:41-60: hash layers until at most 32 bytes remain. */
int or_merkle_tree(const uint8_t *in, size_t len, uint8_t *out, size_t cap, size_t *out_len) {
    size_t need = or_merkle_tree_size(len);
    if (out_len) *out_len = need;
    if (cap < need) return -1;
    const uint8_t *layer = in;
    uint8_t *dst = out;
    while (len > HASH_SIZE) {
        or_process_layer(layer, len, dst);
        layer = dst;
        len = or_next_layer_size(len);
        dst += len;
    }
    return 0;
}

/* ---- Row tree: the binary tree of row hashes (no reference function).
 * Leaf r = K12(row r), the row's 8 * cols little-endian bytes (K12's single
 * node up to 1023 words, its tree mode above); a layer of n nodes gives
 * ceil(n / 2) nodes, node j =
 * K12(node 2j || node 2j + 1) (64 bytes, one permutation), the last node of
 * an odd layer moving up unchanged; layers until one node, the root.  The
 * output is every layer in order, leaves first, 32 bytes per node. */
size_t or_row_tree_size(size_t leaves) {
    if (leaves == 0) return 0;
    size_t total = HASH_SIZE * leaves;
    while (leaves > 1) {
        leaves = (leaves + 1) / 2;
        total += HASH_SIZE * leaves;
    }
    return total;
}

typedef struct {
    const uint8_t *in;
    size_t n, row_bytes;
    uint8_t *out;
} RowCtx;

/* Leaves [64 i, 64 i + 64). */
static void leaf_range(void *arg, size_t i) {
    const RowCtx *cx = (const RowCtx *)arg;
    for (size_t r = 64 * i; r < 64 * i + 64 && r < cx->n; r++)
        or_k12_long(cx->in + r * cx->row_bytes, cx->row_bytes, cx->out + r * HASH_SIZE, HASH_SIZE);
}

/* Parents [64 i, 64 i + 64) of a layer of n nodes. */
static void pair_range(void *arg, size_t i) {
    const RowCtx *cx = (const RowCtx *)arg;
    const size_t m = (cx->n + 1) / 2;
    for (size_t j = 64 * i; j < 64 * i + 64 && j < m; j++) {
        if (2 * j + 1 < cx->n)
            or_k12(cx->in + 2 * j * HASH_SIZE, 2 * HASH_SIZE, cx->out + j * HASH_SIZE, HASH_SIZE);
        else
            memcpy(cx->out + j * HASH_SIZE, cx->in + 2 * j * HASH_SIZE, HASH_SIZE);
    }
}

int or_row_tree(const uint64_t *rows, size_t nrows, size_t cols, uint8_t *out, size_t cap, size_t *out_len) {
    const size_t need = or_row_tree_size(nrows);
    if (out_len) *out_len = need;
    if (nrows == 0 || cols == 0 || cap < need) return -1;
    /* Field words are little-endian u64 on this host: the byte view is the memory */
    RowCtx cx = {(const uint8_t *)rows, nrows, 8 * cols, out};
    or_par_for((nrows + 63) / 64, 64 * (cols / 21 + 1) * 1024, leaf_range, &cx);
    size_t n = nrows;
    uint8_t *layer = out;
    while (n > 1) {
        RowCtx px = {layer, n, 0, layer + n * HASH_SIZE};
        or_par_for(((n + 1) / 2 + 63) / 64, 64 * 1024, pair_range, &px);
        layer += n * HASH_SIZE;
        n = (n + 1) / 2;
    }
    return 0;
}

/* pow.rs:32-37 */
int or_pow_verify(const uint8_t challenge[32], uint64_t threshold, uint64_t nonce) {
    uint8_t msg[40], h[8];
    memcpy(msg, challenge, 32);
    for (int b = 0; b < 8; b++) msg[32 + b] = (uint8_t)(nonce >> (8 * b));
    or_k12(msg, 40, h, 8);
    uint64_t v = 0;
    for (int b = 7; b >= 0; b--) v = (v << 8) | h[b];
    return v >= threshold;
}

/* pow.rs:39-60: the smallest satisfying nonce (sequential scan). */
uint64_t or_pow_solve(const uint8_t challenge[32], uint64_t threshold) {
    for (uint64_t nonce = 0;; nonce++)
        if (or_pow_verify(challenge, threshold, nonce)) return nonce;
}
