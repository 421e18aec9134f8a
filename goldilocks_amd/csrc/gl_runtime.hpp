// Host runtime: error reporting, device memory pool, twiddle tables, the
// NTT planner and plan execution.  Shared by the C ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

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

// ---- device memory pool (buffers are reused by size class) ----
void *pool_alloc(size_t bytes);
void pool_free(void *p);
void pool_trim();

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) : p(bytes ? pool_alloc(bytes) : nullptr) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) pool_free(p); }
    template <class T> T *as() const { return (T *)p; }
};

hipStream_t thread_stream();

// Scratch words owned by a stream: work enqueued on one stream runs in order,
// so successive calls on that stream can share the buffer without waiting.
// Growing it waits for the stream before releasing the old buffer.
u64 *stream_scratch(hipStream_t s, size_t words, int slot = 0);

// ---- twiddles ----
struct PowTable {  // w^e = lo[e & (2^lo_bits - 1)] * hi[e >> lo_bits], e < order
    const u64 *lo = nullptr, *hi = nullptr;
    u32 lo_bits = 0;
};
// Powers of w_M (root of order M).
PowTable root_table(u64 M);
// Powers of an arbitrary base g for exponents < count (e.g. coset shifts).
PowTable power_table(u64 base, u64 count);
// w_R^e for e < R (inner twiddles of a two-sub-step pass).
const u64 *inner_table(u64 R);
// Largest sub-problem length M whose output twiddles use a full M-word
// table; returns the previous value.
u64 set_full_table_max(u64 words);
const PassVariant *rader65537_variant();
// Tables of the radix-257 Rader pass (see gl_rader.hip).
const u64 *rader257_table();

// ---- planning ----
struct PassPlan {
    const PassVariant *v;
    u64 R;
};
struct Plan {
    u64 n = 0;
    std::vector<PassPlan> passes;
};
// Returns GL_OK and a pass list for length n, GL_E_SIZE or GL_E_UNSUPPORTED.
int get_plan(u64 n, const Plan **out);
// Schedule for column transforms (S0 > 1 interleaved transforms).
int get_plan_cols(u64 n, const Plan **out);
// Install an explicit pass list (variant indices as listed by gl_variants;
// index == variant count means the radix-257 Rader pass); empty restores
// the planner's default.
// cols: the schedule of column transforms (S0 > 1) instead of contiguous ones
int set_plan(u64 n, const u32 *ids, size_t nids, bool cols = false);

// Pass p of a plan as launch arguments.  `src_is_data` tracks the ping-pong
// state: pass 0 reads data; non-final passes alternate data <-> tmp; the
// final pass writes data (in place when it also reads data).
struct PassCursor {
    const Plan *plan;
    u64 *data, *tmp;
    u64 S0, count, bstride;
    bool inverse;
    bool scale_out = true;  // inverse: multiply the last pass by 1/n
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
    void begin() { p = 0; S = S0; M = plan->n; src = input ? input : data; scale_folded = false; }
    bool done() const { return p >= plan->passes.size(); }
    // Fills args/grid for the current pass and advances; returns GL_OK or an error.
    int next(PassArgs &a, unsigned &grid);
};

// Run `count` items of bstride words, each holding S0 interleaved transforms
// of length n (S0 = 1: contiguous transforms; S0 = cols: matrix columns).
// tmp must hold count*bstride words when the plan has more than one pass.
// input (optional): the first pass reads the transform from there and the
// result lands in data (same layout), leaving input untouched.
int run_plan(const Plan &plan, u64 *data, u64 *tmp, u64 S0, u64 count, u64 bstride, bool inverse,
             hipStream_t stream, bool scale_out = true, const u64 *input = nullptr, const u64 *coset = nullptr,
             u64 coset_rows = 0, const u64 *row_scale = nullptr);
// Montgomery 2717^j for j < rows (device, cached).
const u64 *coset_table(u64 rows);
// Montgomery 2717^j / rows for j < rows (device, cached): the intt scale and
// the coset shift of the commit's extension in one factor per row.
const u64 *scaled_coset_table(u64 rows);
// True when the plan contains a pass that is not a single kernel (the
// radix-65537 Rader pass), which grouped batch launches cannot include.
bool plan_is_composite(const Plan &plan);
size_t plan_tmp_words(const Plan &plan, u64 count, u64 bstride);

void clear_caches();

// Per-pass timing (gl_timing_enable / gl_timing_collect).
u32 variant_id(const PassVariant *v);
void timing_enable(bool on);
hipEvent_t timing_begin(hipStream_t s);
void timing_end(const PassVariant *v, hipEvent_t e0, hipStream_t s);
size_t timing_pending();
int timing_collect(uint32_t *ids, float *ms, size_t cap, size_t *count);

// Commit pieces shared by the C ABI and the server (gl_api.cpp).
int commit_check(size_t rows, size_t cols, unsigned blowup);
size_t commit_tree_size(size_t rows, size_t cols, unsigned blowup);
int commit_enqueue(const u64 *d_m, size_t rows, size_t cols, unsigned blowup, u64 *d_lde, uint8_t *d_layers,
                   u64 *coef, u64 *tmp, hipStream_t s);

}  // namespace gl
