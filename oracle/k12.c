/*
 * Oracle: Keccak-p[1600,12], single-block KangarooTwelve, layer hashing and
 * the 5-ary byte Merkle tree, plus the PoW check built on them.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates pcs/src/k12/generic.rs:4-94, pcs/src/k12/mod.rs:28-60 and
 * pcs/src/pow.rs:32-60.  The round constants and the rho/pi lane walk are the
 * Keccak specification's tables (FIPS 202), which the reference also uses.
 */
#include "oracle.h"

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

/* SHA3-256 (24 rounds, rate 136, suffix 0x06): used only to pin the
 * permutation against an independent library implementation. */
void or_sha3_256(const uint8_t *in, size_t len, uint8_t out[32]) {
    uint64_t st[25] = {0};
    uint8_t block[200];
    const size_t rate = 136;
    while (len >= rate) {
        memset(block, 0, sizeof block);
        memcpy(block, in, rate);
        uint64_t t[25];
        load_state(t, block);
        for (int i = 0; i < 17; i++) st[i] ^= t[i];
        or_keccak_p1600(st, 24);
        in += rate;
        len -= rate;
    }
    memset(block, 0, sizeof block);
    memcpy(block, in, len);
    block[len] ^= 0x06;
    block[rate - 1] ^= 0x80;
    uint64_t t[25];
    load_state(t, block);
    for (int i = 0; i < 17; i++) st[i] ^= t[i];
    or_keccak_p1600(st, 24);
    store_state(st, block);
    memcpy(out, block, 32);
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
