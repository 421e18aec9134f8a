/*
 * Oracle: column-wise coset low-degree extension + K12 Merkle commit, and the
 * seeded input generator.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no commit function (pcs exposes only merkle_tree,
 * pcs/src/k12/mod.rs:41-60).  This composes reference primitives:
 * intt (ntt/src/ntt/mod.rs:104-120), coefficient scaling by 2717^i
 * (GENERATOR, ntt/src/field/algo/mod.rs:12), ntt (mod.rs:98-102), the raw
 * little-endian byte view of a Field slice (bench/src/main.rs:179-183) and
 * merkle_tree.  or_commit_rows replaces the last step by the binary tree of
 * row hashes (or_row_tree).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    const uint64_t *m;
    uint64_t *lde;
    size_t rows, cols, ext;
} CommitCtx;

/* One column: intt, zero-extend, coefficient r times 2717^r, ntt. */
static void commit_column(void *arg, size_t c) {
    const CommitCtx *cx = (const CommitCtx *)arg;
    uint64_t *col = (uint64_t *)malloc(cx->ext * sizeof(uint64_t));
    for (size_t r = 0; r < cx->rows; r++) col[r] = cx->m[r * cx->cols + c];
    or_intt(col, cx->rows);
    memset(col + cx->rows, 0, (cx->ext - cx->rows) * sizeof(uint64_t));
    uint64_t g = 1;
    for (size_t r = 0; r < cx->rows; r++) {
        col[r] = or_mul(col[r], g);
        g = or_mul(g, OR_GENERATOR);
    }
    or_ntt(col, cx->ext);
    for (size_t r = 0; r < cx->ext; r++) cx->lde[r * cx->cols + c] = col[r];
    free(col);
}

/* The column-wise coset extension into lde_out; -1 for unsupported shapes. */
static int lde(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out) {
    size_t ext = rows * (size_t)blowup;
    if (rows == 0 || cols == 0 || blowup == 0 || !or_is_divisor(rows) || !or_is_divisor(ext))
        return -1;
    CommitCtx cx = {m, lde_out, rows, cols, ext};
    /* columns are independent: one per thread (each column's transforms
     * then run serially inside the parallel loop) */
    or_par_for(cols, ext * 64, commit_column, &cx);
    return 0;
}

int or_commit(const uint64_t *m, size_t rows, size_t cols, unsigned blowup,
              uint64_t *lde_out, uint8_t *layers_out, size_t cap, size_t *out_len) {
    if (lde(m, rows, cols, blowup, lde_out)) return -1;
    /* Field words are little-endian u64, so the byte view is the memory. */
    return or_merkle_tree((const uint8_t *)lde_out, rows * (size_t)blowup * cols * 8, layers_out, cap, out_len);
}

int or_commit_rows(const uint64_t *m, size_t rows, size_t cols, unsigned blowup,
                   uint64_t *lde_out, uint8_t *layers_out, size_t cap, size_t *out_len) {
    if (lde(m, rows, cols, blowup, lde_out)) return -1;
    return or_row_tree(lde_out, rows * (size_t)blowup, cols, layers_out, cap, out_len);
}

static uint64_t splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void or_fill_random(uint64_t *v, size_t n, uint64_t seed) {
    for (size_t i = 0; i < n; i++) {
        uint64_t s = seed + (uint64_t)i * 0x9e3779b97f4a7c15ULL;
        uint64_t x;
        do x = splitmix64(&s); while (x >= OR_P);
        v[i] = x;
    }
}
