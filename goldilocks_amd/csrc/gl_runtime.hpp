// Host runtime: error reporting, device memory pool, twiddle tables, the
// NTT planner and plan execution.  Shared by the C ABI entry points.
//
// Lifetime rules (every entry point may run on several host threads):
// - Plans are reference counted (PlanRef); replacing or clearing a cached
//   plan never frees one a caller still holds.
// - Device tables are reference counted (TableRef).  A table whose last
//   reference drops is retired, not freed; retired tables are freed by
//   clear_caches() after a device synchronization that follows their
//   retirement, so work enqueued while a reference was held has finished.
//   Callers hold the references of every table their enqueued work reads
//   (PassCursor::keep) until the enqueue returns; captured graphs hold them
//   for their lifetime.
// - Device scratch comes from a stream-ordered pool: a block freed on
//   stream s may be handed out again at once for work on s (stream order
//   runs the earlier user first), and to another stream only after the
//   event recorded at the free has completed.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gl_internal.hpp"
#include "../../include/goldilocks_gpu.h"

namespace gl {

// ---- errors ----
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);
#define GL_HIP(call)                                              \
    do {                                                          \
        hipError_t _e = (call);                                   \
        if (_e != hipSuccess) return ::gl::hip_error(_e, #call); \
    } while (0)

int ensure_init();

// ---- device memory pool (stream ordered, reused by size class) ----
void *pool_alloc(size_t bytes, hipStream_t s);
// Return a block; work enqueued on s before this call may still use it.
void pool_free(void *p, hipStream_t s);
// Return a block no queued work uses any more: any stream may take it at once.
void pool_free_idle(void *p);
// Free every idle block whose last use has completed (no device sync).
void pool_release_idle();
// The next `count` pool_alloc calls return nullptr (error-path tests).
void set_fail_allocs(int skip, int count);

// A pool block owned for the duration of one enqueue on a stream.
struct DevBuf {
    void *p = nullptr;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(size_t bytes, hipStream_t stream) : p(bytes ? pool_alloc(bytes, stream) : nullptr), s(stream) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) pool_free(p, s); }
    // Allocate into an empty buffer.
    void alloc(size_t bytes, hipStream_t stream) {
        s = stream;
        p = bytes ? pool_alloc(bytes, stream) : nullptr;
    }
    // Give the block back as idle: the caller has seen every use of it
    // complete (e.g. a synchronous copy ordered after them on its stream).
    void release_idle() {
        if (p) pool_free_idle(p);
        p = nullptr;
    }
    template <class T> T *as() const { return (T *)p; }
};

hipStream_t thread_stream();
// Two more per-thread streams (i = 0, 1) for copies that overlap the
// thread's compute stream (the pipelined host batch).
hipStream_t thread_aux_stream(int i);

// ---- host <-> device copies of the host-buffer entry points ----
// Large copies are staged through pinned chunks (gl_hostio.cpp).
// copy_h2d: returns once h_src may be reused; the device side completes in
// stream order on s.  copy_d2h: ordered after the work on s; returns once
// h_dst holds the bytes (sync = false: a copy too small to be staged is only
// enqueued, and the caller synchronizes s before reading h_dst).
int copy_h2d(void *d_dst, const void *h_src, size_t bytes, hipStream_t s);
int copy_d2h(void *h_dst, const void *d_src, size_t bytes, hipStream_t s, bool sync = true);
// Pinned host memory for small asynchronous uploads (at least `bytes`,
// from a pool of eight slabs, more only while more callers hold one at
// once).  pinned_release records an event
// on s behind the uploads enqueued from the slab; the slab is handed out
// again only after that event has completed.
struct PinnedSlab {
    void *p = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
};
int pinned_acquire(size_t bytes, PinnedSlab &out);
void pinned_release(PinnedSlab &slab, hipStream_t s);

// ---- twiddles (device tables, Montgomery form, cached) ----
using TableRef = std::shared_ptr<const u64>;
using Keep = std::vector<TableRef>;

struct PowTable {  // w^e = lo[e & (2^lo_bits - 1)] * hi[e >> lo_bits], e < order
    TableRef lo, hi;
    u32 lo_bits = 0;
};
// create = false: return the cached table or nullptr, never build one.
// Powers of w_M (root of order M).
PowTable root_table(u64 M, bool create = true);
// Powers of an arbitrary base g for exponents < count (e.g. coset shifts).
PowTable power_table(u64 base, u64 count, bool create = true);
// w_R^e for e < R (inner twiddles of a three-stage pass).
TableRef inner_table(u64 R, bool create = true);
// Inner twiddles of a two-sub-step pass R1 x R2: entry rb*R1 + ka = w_R^(ka*rb).
TableRef inner_matrix(u64 R1, u64 R2, bool create = true);
// Output-order twiddles of a pass of radix R on sub-problems of length M:
// entry k + R*m = w_M^(m*k), M words.
TableRef twiddle_matrix(u64 M, u64 R, bool create = true);
// Largest sub-problem length M whose output twiddles use a full M-word
// table; returns the previous value.
u64 set_full_table_max(u64 words);
const PassVariant *rader65537_variant();
// Tables of the radix-257 Rader pass (see gl_rader.hip).
TableRef rader257_table(bool create = true);
// Montgomery 2717^j for j < rows (the commit's coset shift).
TableRef coset_table(u64 rows);
// Montgomery 2717^j / rows for j < rows: the intt scale and the coset shift
// of the commit's extension in one factor per row.
TableRef scaled_coset_table(u64 rows);

// ---- planning ----
struct PassPlan {
    const PassVariant *v;
    u64 R;
};
struct Plan {
    u64 n = 0;
    std::vector<PassPlan> passes;
};
using PlanRef = std::shared_ptr<const Plan>;
// Returns GL_OK and a pass list for length n, GL_E_SIZE or GL_E_UNSUPPORTED.
// The schedule is the explicit one installed by set_plan, else the tune
// file's, else the planner's default.
int get_plan(u64 n, PlanRef *out);
// Schedule for column transforms (S0 > 1 interleaved transforms).
int get_plan_cols(u64 n, PlanRef *out);
// Install an explicit pass list (variant indices as listed by gl_variants;
// index == variant count means the radix-257 Rader pass); an empty list
// removes it, so the tune file's schedule (else the default) applies again.
// cols: the schedule of column transforms (S0 > 1) instead of contiguous ones
int set_plan(u64 n, const u32 *ids, size_t nids, bool cols = false);

// Pass p of a plan as launch arguments.  Ping-pong: pass 0 reads data (or
// `input`); non-final passes alternate data <-> tmp; the final pass writes
// data (in place when it also reads data).
struct PassCursor {
    const Plan *plan;
    u64 *data, *tmp;
    u64 S0, count, bstride;
    bool inverse;
    bool scale_out = true;  // inverse: multiply the last pass by 1/n
    bool create_tables = true;  // false: a missing table is GL_E_UNSUPPORTED (nothing is allocated)
    const u64 *input = nullptr;  // first pass reads here instead of data (out of place)
    // first pass only (two-sub-step variants): input row j < coset_rows is
    // multiplied by coset[j], rows above read as zero (FLAG_COSET)
    // (coset == nullptr with coset_rows > 0: zero rows only, FLAG_ZEXT)
    const u64 *coset = nullptr;
    u64 coset_rows = 0;
    // inverse, last pass (two-sub-step variants): output row r is multiplied
    // by row_scale[r] instead of 1/n (FLAG_ROW_SCALE)
    const u64 *row_scale = nullptr;
    // evolving state
    size_t p = 0;
    u64 S = 0, M = 0;
    const u64 *src = nullptr;
    bool scale_folded = false;  // 1/n went into the first pass's twiddle recurrence
    Keep keep;                  // tables the filled-in passes read
    void begin() { p = 0; S = S0; M = plan->n; src = input ? input : data; scale_folded = false; }
    bool done() const { return p >= plan->passes.size(); }
    // Fills args/grid for the current pass and advances; returns GL_OK or an error.
    int next(PassArgs &a, unsigned &grid);
};

// Run `count` items of bstride words, each holding S0 interleaved transforms
// of length n (S0 = 1: contiguous transforms; S0 = cols: matrix columns).
// tmp must hold count*bstride words when the plan has more than one pass.
// input (optional): the first pass reads the transform from there and the
// result lands in data (same layout), leaving input untouched.  row_scale
// (optional; last pass a single two- or three-stage kernel): output row r
// is multiplied by row_scale[r] (Montgomery product), replacing the
// inverse's 1/n.
int run_plan(const Plan &plan, u64 *data, u64 *tmp, u64 S0, u64 count, u64 bstride, bool inverse,
             hipStream_t stream, bool scale_out = true, const u64 *input = nullptr, const u64 *coset = nullptr,
             u64 coset_rows = 0, const u64 *row_scale = nullptr);
// Build every table a plan's passes read (forward and inverse), without
// launching anything; the references go to keep.
int warm_plan(const Plan &plan, u64 S0, Keep *keep);
// True when the plan contains a pass that is not a single kernel (the
// radix-65537 Rader pass), which grouped batch launches cannot include.
bool plan_is_composite(const Plan &plan);
size_t plan_tmp_words(const Plan &plan, u64 count, u64 bstride);
// True when a * b * c words (8 bytes each) do not fit in size_t bytes.
inline bool words_overflow(u64 a, u64 b, u64 c = 1) {
    u64 ab, abc, bytes;
    return __builtin_mul_overflow(a, b, &ab) || __builtin_mul_overflow(ab, c, &abc) ||
           __builtin_mul_overflow(abc, (u64)8, &bytes);
}
// Device words of scratch a plan takes beyond tmp (the radix-65537 pass).
size_t plan_extra_words(const Plan &plan, u64 count, u64 bstride);

// Drop cached plans and tables (explicit schedules stay installed), free
// retired tables after a device synchronization, release idle pool blocks.
void clear_caches();

// Per-pass timing (gl_timing_enable / gl_timing_collect).
u32 variant_id(const PassVariant *v);
void timing_enable(bool on);
hipEvent_t timing_begin(hipStream_t s);
void timing_end(const PassVariant *v, hipEvent_t e0, hipStream_t s);
size_t timing_pending();
int timing_collect(uint32_t *ids, float *ms, size_t cap, size_t *count);

// Commit pieces shared by the C ABI and the server (gl_api.cpp).  The tree
// over the extension is the 5-ary byte Merkle tree (TREE_BYTES) or the
// binary tree of row hashes (TREE_ROWS, at most kRowTreeMaxCols columns;
// its layers output must be 8-byte aligned).
enum CommitTree : int { TREE_BYTES = 0, TREE_ROWS = 1 };
int commit_check(size_t rows, size_t cols, unsigned blowup, CommitTree tree = TREE_BYTES);
size_t commit_tree_size(size_t rows, size_t cols, unsigned blowup, CommitTree tree = TREE_BYTES);
// Plans and tables of a commit shape (built once; references in keep).
int commit_prepare(size_t rows, size_t cols, unsigned blowup, Keep *keep);
int commit_enqueue(const u64 *d_m, size_t rows, size_t cols, unsigned blowup, u64 *d_lde, uint8_t *d_layers,
                   u64 *coef, u64 *tmp, hipStream_t s, CommitTree tree = TREE_BYTES);

}  // namespace gl
