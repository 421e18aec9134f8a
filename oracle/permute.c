/*
 * Oracle: the reference's public permutations (ntt/src/permute/), restated
 * for the GPU transpose and bit-reversal entry points.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* permutation.rs:1-4: the source position of destination i when a matrix of
 * `width` columns and `height` rows is transposed (q = width*height - 1
 * stays fixed, every other i maps to i*width mod q). */
size_t or_transpose_index(size_t width, size_t height, size_t i) {
    const size_t n = width * height;
    if (n == 0) return 0;
    const size_t q = n - 1;
    return i == q ? q : (size_t)(((unsigned __int128)i * width) % q);
}

/* mod.rs:53-67: a rows x cols row-major matrix becomes its cols x rows
 * transpose in place, written in destination order as transpose_copy does:
 * new[i*rows + j] = old[j*cols + i]. */
int or_transpose(uint64_t *m, size_t rows, size_t cols) {
    const size_t n = rows * cols;
    if (n == 0) return 0;
    uint64_t *copy = malloc(n * sizeof *copy);
    if (!copy) return -1;
    memcpy(copy, m, n * sizeof *copy);
    for (size_t i = 0; i < cols; i++)
        for (size_t j = 0; j < rows; j++) m[i * rows + j] = copy[j * cols + i];
    free(copy);
    return 0;
}

/* permutation.rs:9-16: i -> i with its log2(size) low bits reversed. */
size_t or_bit_reverse_index(size_t size, size_t i) {
    size_t bits = 0, r = 0;
    while (((size_t)1 << bits) < size) bits++;
    for (size_t b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

/* mod.rs:71-74: v[i] <- v[rev(i)] for a power-of-two length (the reference
 * asserts it, permutation.rs:13); -1 otherwise. */
int or_bit_reverse(uint64_t *v, size_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    for (size_t i = 0; i < n; i++) {
        const size_t j = or_bit_reverse_index(n, i);
        if (i < j) {
            const uint64_t t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
    return 0;
}
