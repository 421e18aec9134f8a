/*
 * goldilocks_gpu.h - C ABI of the MI355X (gfx950) Goldilocks NTT and K12
 * Merkle library (libgoldilocks_amd.so).
 *
 * Every pointer is a plain host or device address; every "Field" word is the
 * raw u64 the reference keeps in a `&mut [Field]` (Montgomery form, canonical
 * in [0, p), p = 2^64 - 2^32 + 1; ntt/src/field/mod.rs:8-12).  Results are
 * bit-exact with the reference's CPU path on every output word and hash.
 *
 * Errors: functions return 0 (GL_OK) or a GL_E_* code and never abort the
 * host; gl_last_error() gives a message for the calling thread.  Where the
 * reference panics (ntt/src/ntt/mod.rs:55-58, pcs/src/k12/generic.rs:65-66),
 * these return GL_E_SIZE / GL_E_ARG.
 *
 * Threading: all entry points may be called from several host threads,
 * also on one stream; host-buffer calls use a stream per calling thread.
 * Plans and twiddle tables are shared and reference counted: a call, or a
 * captured batch plan, keeps what it reads alive, and gl_clear_cache frees
 * nothing still in use.  Each call takes its own device scratch, stream
 * ordered (reused by later work on the same stream without waiting).
 * Device-buffer calls ("_dev") enqueue on the caller's stream (NULL = the
 * default stream) and return without synchronizing unless stated otherwise.
 */
#ifndef GOLDILOCKS_GPU_H
#define GOLDILOCKS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GL_OK 0
#define GL_E_SIZE 1        /* length does not divide p - 1 (or is 0 for a forward transform) */
#define GL_E_NOMEM 2       /* device allocation failed or an HBM budget was exceeded */
#define GL_E_DEVICE 3      /* HIP runtime error */
#define GL_E_ARG 4         /* invalid argument (null pointer, short buffer, K12 input > 166 B) */
#define GL_E_UNSUPPORTED 5 /* legal size the device planner cannot schedule */
#define GL_E_BUSY 6        /* server: no slot / budget right now (retry) */

/* Message for the last error on the calling thread ("" if none). */
const char *gl_last_error(void);

/* Select and initialise a device (optional; the first call initialises the
 * current device).  Returns GL_E_DEVICE when no gfx950 device is usable. */
int gl_init(int device);

/* ntt/src/divisors.rs:21-23: n > 0 and n | p - 1. */
int gl_is_divisor(uint64_t n);

/* ntt/src/divisors.rs:52-59 (divisors::next): the smallest supported NTT
 * length >= n; 0 for n = 0, and 0 when n > p - 1 (where the reference's
 * lookup runs off its table).  Host-only. */
uint64_t gl_next_size(uint64_t n);

/* ---- NTT, host buffers (copied to HBM and back; synchronous) ------------- */

/* Forward transform in place.  Replaces goldilocks_ntt::ntt::ntt
 * (ntt/src/ntt/mod.rs:98-102). */
int gl_ntt(uint64_t *v, size_t n);

/* Inverse transform in place: scale by 1/n, reverse v[1..], forward.
 * n <= 1 is a no-op.  Replaces ntt::intt (ntt/src/ntt/mod.rs:104-120). */
int gl_intt(uint64_t *v, size_t n);

/* `count` transforms of length n stored back to back.  Replaces
 * strategy(n).ntt(values) with values.len() == n * count
 * (ntt/src/ntt/mod.rs:48-78, trait Ntt :25-31); inverse != 0 runs intt on
 * each item. */
int gl_ntt_batch(uint64_t *v, size_t n, size_t count, int inverse);

/* k transforms of individual lengths n[i] at v[i] (mixed sizes and radices),
 * scheduled as one batched launch sequence.  No reference counterpart beyond
 * looping ntt::ntt over the slices; results equal that loop. */
int gl_ntt_batch_mixed(uint64_t *const *v, const size_t *n, size_t k, int inverse);

/* ---- NTT, device buffers (words already in HBM) ------------------------- */

/* As gl_ntt_batch on a device buffer of n*count words, enqueued on `stream`. */
int gl_ntt_dev(uint64_t *d, size_t n, size_t count, int inverse, void *stream);

/* gl_ntt_dev that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It has no counterpart in the reference.  The ping-pong
 * scratch is one fixed array of 2^24 words (128 MiB) per device, held until
 * the process ends: n * count above 2^24 is GL_E_ARG, lengths with a factor
 * 65537 are GL_E_UNSUPPORTED.  It builds no twiddle table: run an allocating
 * entry point (gl_ntt_dev, gl_ntt, ...) on the same length once first, else
 * GL_E_UNSUPPORTED.  Every call on a device shares its array: calls on one
 * stream run in order, calls that overlap on one device (other streams or
 * threads) race. */
int gl_ntt_dev_noalloc(uint64_t *d, size_t n, size_t count, int inverse, void *stream);

/* Column transforms of a row-major (n x cols) matrix: column c is words
 * c, c + cols, c + 2*cols, ...; `count` such matrices back to back. */
int gl_ntt_cols_dev(uint64_t *d, size_t n, size_t cols, size_t count, int inverse, void *stream);

/* Mixed-size batch on device buffers d[i] of n[i] words.  Transforms are
 * packed in order into chunks of at most gl_set_batch_chunk_bytes bytes
 * (default 1 GiB, or GOLDILOCKS_BATCH_CHUNK_BYTES); each chunk runs as
 * grouped launches, one per (pass level, kernel variant).  Larger transforms
 * and radix-65537 ones run on their own.  Enqueues on `stream` and returns,
 * like the other device entry points: each (chunk, pass level) is launched
 * as soon as its descriptors are built (uploaded through pinned memory), so
 * the device runs one level while the host plans the next.  Not for graph
 * capture (GL_E_UNSUPPORTED on a capturing stream): gl_batch_plan_create
 * captures a batch. */
int gl_ntt_mixed_dev(uint64_t *const *d, const size_t *n, size_t k, int inverse, void *stream);

/* Chunk size of the grouped batch path; returns the previous value.  Results
 * do not depend on it. */
uint64_t gl_set_batch_chunk_bytes(uint64_t bytes);

/* The same batch, one launch sequence per transform (the unbatched
 * baseline the batched path is measured against). */
int gl_ntt_mixed_dev_unbatched(uint64_t *const *d, const size_t *n, size_t k, int inverse,
                               void *stream);

/* Seeded uniform words in [0, p): word i is the first splitmix64 output
 * below p of the stream seeded at seed + i * 0x9e3779b97f4a7c15. */
int gl_fill_random_dev(uint64_t *d, size_t n, uint64_t seed, void *stream);

/* Describe the device schedule for length n: pass i runs a radix
 * r1[i]*r2[i] kernel variant with t[i] instances per workgroup.  Fills up to
 * cap entries, *npasses = pass count.  (Planner introspection; the
 * reference's counterpart is the plan tree strategy(n) builds.) */
int gl_plan_passes(uint64_t n, uint32_t *r1, uint32_t *r2, uint32_t *t, size_t cap, size_t *npasses);

/* Compiled pass kernel variants: entry i runs DFT_(r1*r2) per instance with
 * t instances per workgroup.  Returns the count; fills up to cap entries. */
size_t gl_variants(uint32_t *r1, uint32_t *r2, uint32_t *t, size_t cap);

/* Replace the schedule for length n by the given variant indices (product
 * of their radices must be n); used by the autotuner.  nvariants = 0 restores
 * the planner's default. */
int gl_plan_set(uint64_t n, const uint32_t *variant_ids, size_t nvariants);

/* The same for column transforms of length n (gl_ntt_cols_dev with cols > 1,
 * and the commit's column transforms), which keep a schedule of their own.
 * Tuning aid, no reference counterpart. */
int gl_plan_set_cols(uint64_t n, const uint32_t *variant_ids, size_t nvariants);

/* Kernel name of variant `id` (ids as for gl_plan_set; count = radix-257
 * Rader kernel, count + 1 = radix-65537 Rader composite); NULL if unknown.
 * Tuning/introspection aid, no reference counterpart. */
const char *gl_variant_name(uint32_t id);

/* Variant ids of the passes of the schedule for length n (the same plan
 * gl_ntt would run). */
int gl_plan_variants(uint64_t n, uint32_t *ids, size_t cap, size_t *npasses);
/* The same for the column-transform schedule of length n. */
int gl_plan_variants_cols(uint64_t n, uint32_t *ids, size_t cap, size_t *npasses);

/* Output twiddles of a pass on sub-problems of length M use one M-word table
 * (one load and one product per output) when M <= words, or M <= 4 * words
 * for passes whose instance stride S is >= 32 (S adjacent instances share
 * each table word), else a two-level power table (w^(m*k0) * (w^(m*kstep))^j).
 * Results are identical either way; returns the previous threshold (default
 * 2^20, or GOLDILOCKS_TWIDDLE_FULL_MAX; 0 disables the full tables). */
uint64_t gl_set_twiddle_full_max(uint64_t words);

/* Pass timing (measurement aid, no reference counterpart).  While enabled,
 * every NTT pass launch (gl_ntt*, gl_ntt_cols_dev, the commit's transforms
 * and grouped batch launches) is bracketed by HIP events on its stream.
 * gl_timing_pending = passes recorded since the last collect.
 * gl_timing_collect waits for them and returns, in launch order, the variant
 * id (as for gl_variant_name) and duration in ms of up to cap of them;
 * *count = number recorded (all are released).  Launches made while a
 * stream is being captured into a graph, or once 2^20 records await
 * collection, are not timed. */
void gl_timing_enable(int on);
size_t gl_timing_pending(void);
int gl_timing_collect(uint32_t *variant_ids, float *ms, size_t cap, size_t *count);

/* Testing aid with no reference counterpart (the reference aborts on
 * allocation failure): the next `count` device scratch allocations of the
 * library fail as if hipMalloc had, so callers can exercise the GL_E_NOMEM
 * paths of every entry point and of the server; 0 clears it. */
void gl_debug_fail_allocs(int count);
/* The same after the next `skip` allocations have succeeded (a failure in
 * the middle of a multi-allocation call, e.g. a later chunk of a pipelined
 * host batch). */
void gl_debug_fail_allocs_after(int skip, int count);

/* Drop cached plans and device twiddle tables (ntt/src/ntt/mod.rs:43-46).
 * Tables still referenced by work being enqueued or by a batch plan stay
 * valid and are freed by a later call; explicit (gl_plan_set) and tuned
 * schedules stay in force.  Synchronizes the device. */
void gl_clear_cache(void);

/* ---- K12 / Merkle (pcs/src/k12) ----------------------------------------- */

/* k12::<H>(input) for len <= 166, outlen <= 168 (pcs/src/k12/generic.rs:62-84). */
int gl_k12(const uint8_t *in, size_t len, uint8_t *out, size_t outlen);

/* next_layer_size (pcs/src/k12/mod.rs:33-35): ceil(len / 160) * 32. */
size_t gl_next_layer_size(size_t len);

/* Total bytes of all layers merkle_tree(len) returns. */
size_t gl_merkle_tree_size(size_t len);

/* process_layer (pcs/src/k12/generic.rs:86-94, xkcp.rs:5-14):
 * out[32i..32i+32] = K12(in[160i .. min(160i+160, len)]). */
int gl_k12_layer(const uint8_t *in, size_t len, uint8_t *out);
int gl_k12_layer_dev(const uint8_t *d_in, size_t len, uint8_t *d_out, void *stream);

/* merkle_tree (pcs/src/k12/mod.rs:41-60): every layer, concatenated in
 * order (sizes follow from gl_next_layer_size), root last.  *out_len gets
 * the total; GL_E_ARG if cap is smaller (then *out_len = required size). */
int gl_merkle_tree(const uint8_t *in, size_t len, uint8_t *out, size_t cap, size_t *out_len);
int gl_merkle_tree_dev(const uint8_t *d_in, size_t len, uint8_t *d_out, size_t cap,
                       size_t *out_len, void *stream);

/* Proof of work (pcs/src/pow.rs:32-60): the smallest nonce with
 * u64_le(K12(challenge || nonce_le)[0..8]) >= threshold, searched on the GPU
 * in increasing order up to max_nonce (exclusive).  GL_E_ARG if none. */
int gl_pow_solve(const uint8_t *challenge32, uint64_t threshold, uint64_t max_nonce, uint64_t *nonce);

/* ---- Convolution (ntt/src/convolve.rs:13-20) -----------------------------
 * circular(a, b): a <- intt(ntt(a) * b) with b already in NTT form; the
 * pointwise product is the reference's Field product (Montgomery, R = 2^64)
 * on the raw words.  Length n must divide p - 1. */
int gl_convolve(uint64_t *a, const uint64_t *b_ntt, size_t n);
int gl_convolve_dev(uint64_t *d_a, const uint64_t *d_b_ntt, size_t n, void *stream);

/* ---- Permutations (ntt/src/permute/mod.rs) --------------------------------
 * Of u64 words (the reference's are generic over Copy types).
 * gl_transpose(_dev): permute::transpose (mod.rs:53-58, the semantics of
 * transpose_copy, mod.rs:60-67): the rows x cols row-major matrix becomes its
 * cols x rows transpose in place.  Square matrices swap tile pairs in place;
 * other shapes go through device scratch (pool) and a copy back.
 * gl_transpose_oop_dev: the same move from src into dst (dst != src, no
 * scratch; the reference has no out-of-place form).  rows * cols == 0 is a
 * no-op; a shape whose byte size overflows is GL_E_ARG.
 * gl_bit_reverse(_dev): permute::bit_reverse (mod.rs:71-74), v[i] <- v[rev(i)]
 * with rev reversing the log2(n) low bits (permutation.rs:9-16); n must be a
 * power of two (the reference asserts it, permutation.rs:13): GL_E_SIZE
 * otherwise, n = 0 included. */
int gl_transpose(uint64_t *m, size_t rows, size_t cols);
int gl_transpose_dev(uint64_t *d, size_t rows, size_t cols, void *stream);
int gl_transpose_oop_dev(const uint64_t *d_src, uint64_t *d_dst, size_t rows, size_t cols, void *stream);
int gl_bit_reverse(uint64_t *v, size_t n);
int gl_bit_reverse_dev(uint64_t *d, size_t n, void *stream);

/* ---- Merkle openings (the step after the commit) --------------------------
 * No reference function: the reference returns the layers (pcs/src/k12/
 * mod.rs:41-60) and stops.  An opening of leaf block b (the 160-byte block
 * b of the leaf bytes) is depth blocks of 160 bytes: for layer l = 0 (the
 * leaves) .. depth-1, block b / 5^l of layer l, zero padded past the end of
 * the layer.  Hashing it (k12::<32> of its unpadded bytes) gives the entry
 * that sits at byte 32 * ((b / 5^l) % 5) of the next block; the last one
 * hashes to the 32-byte root. */

/* Number of hashed layers of a tree over leaf_len bytes (= the number of
 * layers merkle_tree returns). */
size_t gl_merkle_depth(size_t leaf_len);

/* k openings of leaf blocks d_blocks[i] from a device-resident tree (leaf
 * bytes + the layers gl_merkle_tree_dev wrote) into d_out, k * depth * 160
 * bytes.  Block indices past the leaves give all-zero openings. */
int gl_merkle_open_dev(const uint8_t *d_leaves, size_t leaf_len, const uint8_t *d_layers, const uint64_t *d_blocks,
                       size_t k, uint8_t *d_out, void *stream);

/* d_ok[i] = 1 when opening i (layout of gl_merkle_open_dev) of leaf block
 * d_blocks[i] hashes up to d_root32, else 0. */
int gl_merkle_verify_dev(const uint8_t *d_paths, size_t leaf_len, const uint64_t *d_blocks, size_t k,
                         const uint8_t *d_root32, int32_t *d_ok, void *stream);

/* ---- Commit (no single reference function; composed from ntt::intt,
 * coefficient scaling by GENERATOR^i, ntt::ntt and merkle_tree over the raw
 * bytes as bench/src/main.rs:179-183 hashes them) ---------------------------
 * m: row-major rows x cols words.  lde_out: (blowup*rows) x cols row-major
 * coset low-degree extension; layers_out: merkle_tree of its bytes. */
int gl_commit(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out,
              uint8_t *layers_out, size_t cap, size_t *out_len);
int gl_commit_dev(const uint64_t *d_m, size_t rows, size_t cols, unsigned blowup,
                  uint64_t *d_lde, uint8_t *d_layers, size_t cap, size_t *out_len, void *stream);

/* ---- Row tree and row-tree commit (no reference function: the binary tree
 * of row hashes, an alternative to merkle_tree over the commit's bytes) -----
 * Leaf r = KangarooTwelve of row r's 8*cols little-endian bytes (empty
 * customization, 32-byte output, RFC 9861): up to 1023 words K12's single
 * node, TurboSHAKE128 of the row and 0x00 with D = 0x07 over several
 * Keccak-p[1600,12] blocks (the same permutation, padding and constants as
 * k12::<32>, pcs/src/k12/generic.rs:62-84, which is this for inputs up to
 * 166 bytes); from 1024 words K12's tree mode over 8192-byte chunks; rows
 * of at most 262143 words (GL_E_SIZE beyond).  A layer of n nodes gives ceil(n/2) parents,
 * parent j = k12::<32>(node 2j || node 2j+1); the last node of an odd layer
 * moves up unchanged; layers until one node, the root.  Layers output: every
 * layer in order, leaves first, 32 bytes per node, gl_row_tree_size(leaves)
 * bytes; device outputs must be 8-byte aligned. */
size_t gl_row_tree_size(size_t leaves);
/* Layers above the leaves (ceil(log2(leaves))). */
size_t gl_row_tree_depth(size_t leaves);
int gl_row_tree(const uint64_t *rows, size_t nrows, size_t cols, uint8_t *layers_out, size_t cap, size_t *out_len);
int gl_row_tree_dev(const uint64_t *d_rows, size_t nrows, size_t cols, uint8_t *d_layers, size_t cap,
                    size_t *out_len, void *stream);
/* The commit's column coset extension (as gl_commit) with the row tree over
 * its rows; layers_out holds gl_row_tree_size(blowup*rows) bytes. */
int gl_commit_rows(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out,
                   uint8_t *layers_out, size_t cap, size_t *out_len);
int gl_commit_rows_dev(const uint64_t *d_m, size_t rows, size_t cols, unsigned blowup, uint64_t *d_lde,
                       uint8_t *d_layers, size_t cap, size_t *out_len, void *stream);
/* k openings of leaves d_idx[i] into d_out, k * depth * 32 bytes: for layer
 * l < depth the sibling of node d_idx[i] >> l, or 32 zero bytes where that
 * node has none (it moves up unchanged) or d_idx[i] >= leaves. */
int gl_row_open_dev(const uint8_t *d_layers, size_t leaves, const uint64_t *d_idx, size_t k, uint8_t *d_out,
                    void *stream);
/* d_ok[i] = 1 when row i of d_rows (k x cols words: the opened rows' values)
 * at leaf d_idx[i] with opening i (layout of gl_row_open_dev) hashes up to
 * d_root32, else 0. */
int gl_row_verify_dev(const uint64_t *d_rows, size_t cols, size_t leaves, const uint64_t *d_idx, size_t k,
                      const uint8_t *d_paths, const uint8_t *d_root32, int32_t *d_ok, void *stream);

/* ---- Captured batch plans (hipGraph) -------------------------------------
 * A mixed batch over fixed device buffers, planned once: descriptors are
 * uploaded at creation and the launch sequence is captured as a hipGraph, so
 * each gl_batch_plan_launch is one graph launch with no host work.  Results
 * equal gl_ntt_mixed_dev on the same buffers. */
typedef struct gl_batch_plan gl_batch_plan;
int gl_batch_plan_create(uint64_t *const *d, const size_t *n, size_t k, int inverse,
                         gl_batch_plan **out);
int gl_batch_plan_launch(gl_batch_plan *plan, void *stream);
void gl_batch_plan_destroy(gl_batch_plan *plan);

/* ---- Serving: continuous batching of NTT / commit jobs ------------------
 * A server owns `slots` HIP streams and an HBM budget.  Submitted jobs wait
 * in FIFO order and are admitted as soon as a slot is free and their device
 * footprint fits in the unused budget, so work that arrives while other jobs
 * run overlaps with them on the device.  Twiddle tables and plans stay
 * resident across jobs.  Job buffers are host memory unless `device` is set.
 * Results are the same as running each job alone. */
#define GL_JOB_NTT 0
#define GL_JOB_INTT 1
#define GL_JOB_COMMIT 2
#define GL_JOB_COMMIT_ROWS 3  /* gl_commit_rows: layers_out holds gl_row_tree_size(blowup*rows) bytes */

typedef struct {
    int kind;             /* GL_JOB_NTT / GL_JOB_INTT / GL_JOB_COMMIT / GL_JOB_COMMIT_ROWS */
    int device;           /* nonzero: all buffers below are device pointers */
    uint64_t *data;       /* NTT: count*n words, transformed in place; COMMIT: rows x cols input */
    size_t n;             /* NTT length, or commit rows */
    size_t count;         /* NTT: transforms back to back; COMMIT: columns */
    unsigned blowup;      /* COMMIT only */
    uint64_t *lde_out;    /* COMMIT only: (blowup*rows) x cols words */
    uint8_t *layers_out;  /* COMMIT only: gl_merkle_tree_size(blowup*rows*cols*8) bytes */
} gl_job;

typedef struct {
    uint64_t submitted, completed, failed, rejected;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion latency */
    double elements;                        /* words transformed (NTT) / LDE words (commit) */
    double busy_seconds;                    /* wall time with at least one job in flight */
    size_t peak_bytes;                      /* peak admitted device footprint */
} gl_server_stats;

typedef struct gl_server gl_server;
/* hbm_budget_bytes = 0: no budget limit. slots in [1, 16]. */
gl_server *gl_server_create(size_t hbm_budget_bytes, int slots);
/* Queue a job; GL_E_NOMEM if its footprint alone exceeds the budget, GL_E_SIZE /
 * GL_E_ARG for invalid jobs.  The job's buffers must stay valid until done. */
int gl_server_submit(gl_server *srv, const gl_job *job, uint64_t *ticket);
/* Nonblocking: *done = 1 when the job finished; returns the job's status.
 * Once a poll has reported the job done, its ticket is released. */
int gl_server_poll(gl_server *srv, uint64_t ticket, int *done);
/* Block until the job finished; latency_ms (may be NULL) = submit -> done.
 * Returns the job's status and releases the ticket. */
int gl_server_wait(gl_server *srv, uint64_t ticket, double *latency_ms);
/* Block until every submitted job finished. */
int gl_server_drain(gl_server *srv);
int gl_server_stats_get(gl_server *srv, gl_server_stats *out);
void gl_server_destroy(gl_server *srv);

#ifdef __cplusplus
}
#endif
#endif
