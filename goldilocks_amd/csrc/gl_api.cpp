// C ABI (include/goldilocks_gpu.h): host-buffer and device-buffer entry
// points for the NTT, K12 layers, Merkle trees and the commit.
#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gl_field.hpp"
#include "gl_runtime.hpp"

using namespace gl;

namespace gl {
const char *last_error_cstr();
}

namespace {

constexpr size_t HASH = 32, BLOCK = 160;

size_t next_layer(size_t len) { return (len + BLOCK - 1) / BLOCK * HASH; }

size_t tree_size(size_t len) {
    size_t total = 0;
    while (len > HASH) {
        len = next_layer(len);
        total += len;
    }
    return total;
}

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

// Row tree over `leaves` row hashes: nodes of every layer, leaves included.
size_t row_tree_nodes(size_t leaves) {
    size_t total = leaves;
    while (leaves > 1) {
        leaves = (leaves + 1) / 2;
        total += leaves;
    }
    return total;
}

int row_geometry(size_t leaves, RowGeometry &g) {
    g.depth = 0;
    g.n[0] = leaves;
    g.off[0] = 0;
    u64 n = leaves, off = 0;
    while (n > 1) {
        if (g.depth == kMaxRowDepth) return set_error(GL_E_UNSUPPORTED, "row tree too deep");
        off += n;
        n = (n + 1) / 2;
        g.depth++;
        g.n[g.depth] = n;
        g.off[g.depth] = off;
    }
    return GL_OK;
}

int row_tree_check(size_t nrows, size_t cols) {
    if (nrows == 0 || cols == 0) return set_error(GL_E_ARG, "empty row tree");
    if (words_overflow(nrows, cols, 4)) return set_error(GL_E_ARG, "row tree shape overflows the address space");
    if (cols > kRowTreeMaxCols) return set_error(GL_E_SIZE, "row tree: rows of at most 262143 words");
    return GL_OK;
}

int row_tree_dev_impl(const u64 *d_rows, size_t nrows, size_t cols, uint8_t *d_out, size_t cap, size_t *out_len,
                      hipStream_t s) {
    int rc = row_tree_check(nrows, cols);
    if (rc) return rc;
    const size_t need = 32 * row_tree_nodes(nrows);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the row tree");
    if (!d_rows || !d_out) return set_error(GL_E_ARG, "null device pointer");
    if ((uintptr_t)d_rows % 8 || (uintptr_t)d_out % 8) return set_error(GL_E_ARG, "row tree buffers must be 8-byte aligned");
    // chaining values of K12's tree mode (rows over 1023 words), stream ordered
    const u64 cvw = row_tree_cv_words(nrows, (u32)cols);
    DevBuf cv(cvw * 8, s);
    if (cvw && !cv.p) return set_error(GL_E_NOMEM, "row tree scratch allocation failed");
    launch_row_tree(d_rows, nrows, (u32)cols, (u64 *)d_out, cv.as<u64>(), s);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

// Copy a host entry point's result back (synchronously: it completes after
// every earlier use of the call's blocks on s), then return those blocks to
// the pool idle, with no event to record.
template <class... Bufs>
int copy_back_idle(void *h_dst, const void *d_src, size_t bytes, hipStream_t s, Bufs &...bufs) {
    const int rc = copy_d2h(h_dst, d_src, bytes, s);
    if (rc == GL_OK) (bufs.release_idle(), ...);
    return rc;
}

// scratch: where the ping-pong buffer goes when the caller frees it (e.g. as
// idle after a synchronous copy); by default the call's own, stream ordered.
int ntt_dev_impl(u64 *d, u64 n, u64 S0, u64 count, bool inverse, hipStream_t s, DevBuf *scratch = nullptr) {
    int rc = ensure_init();
    if (rc) return rc;
    if (inverse && n <= 1) return GL_OK;  // mod.rs:105-107
    PlanRef plan;
    if ((rc = S0 > 1 ? get_plan_cols(n, &plan) : get_plan(n, &plan))) return rc;
    if (count == 0 || S0 == 0 || plan->passes.empty()) return GL_OK;
    if (!d) return set_error(GL_E_ARG, "null device pointer");
    if (words_overflow(n, S0, count)) return set_error(GL_E_ARG, "n * count words overflow the address space");
    const u64 bstride = n * S0;
    const size_t words = plan_tmp_words(*plan, count, bstride);
    DevBuf own;
    DevBuf &tmp = scratch ? *scratch : own;
    tmp.alloc(words * 8, s);  // this call's scratch
    if (words && !tmp.p) return set_error(GL_E_NOMEM, "scratch allocation failed");
    return run_plan(*plan, d, tmp.as<u64>(), S0, count, bstride, inverse, s);
}

// Device scratch of a tree or layer call: two 32-word halves, each holding
// a layer's padded partial last block (a layer reads one and stages the next
// layer's in the other).
constexpr size_t kStageBytes = 64 * 8;

int merkle_dev_impl(const uint8_t *d_in, size_t len, uint8_t *d_out, size_t cap, size_t *out_len,
                    hipStream_t s) {
    size_t need = tree_size(len);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the Merkle tree");
    const uint8_t *layer = d_in;
    uint8_t *dst = d_out;
    if (need == 0) return GL_OK;
    DevBuf stage_buf(kStageBytes, s);
    u64 *stage = stage_buf.as<u64>();
    if (!stage) return set_error(GL_E_NOMEM, "scratch allocation failed");
    int cur = 0;
    bool ready = false;
    while (len > HASH) {
        ready = launch_k12_layer(layer, len, dst, stage + 32 * cur, ready, stage + 32 * (1 - cur), s);
        GL_HIP(hipGetLastError());
        cur = 1 - cur;
        layer = dst;
        len = next_layer(len);
        dst += len;
    }
    return GL_OK;
}

}  // namespace

namespace gl {

int commit_check(size_t rows, size_t cols, unsigned blowup, CommitTree tree) {
    const u64 ext = (u64)rows * blowup;
    if (rows == 0 || cols == 0 || blowup == 0) return set_error(GL_E_ARG, "empty commit shape");
    if (words_overflow(rows, blowup, cols)) return set_error(GL_E_ARG, "commit shape overflows the address space");
    if (!((P - 1) % rows == 0) || !((P - 1) % ext == 0))
        return set_error(GL_E_SIZE, "commit: rows and blowup*rows must divide p - 1");
    if (tree == TREE_ROWS && cols > kRowTreeMaxCols)
        return set_error(GL_E_SIZE, "row-tree commit: rows of at most 262143 words");
    return GL_OK;
}

size_t commit_tree_size(size_t rows, size_t cols, unsigned blowup, CommitTree tree) {
    if (tree == TREE_ROWS) return 32 * row_tree_nodes((u64)rows * blowup);
    return tree_size((u64)rows * blowup * cols * 8);
}

// The commit's two column plans (inverse over rows, forward over the
// extension).
static int commit_plans(size_t rows, size_t cols, unsigned blowup, PlanRef &prow, PlanRef &pext) {
    const u64 ext = (u64)rows * blowup;
    int rc;
    if ((rc = cols > 1 ? get_plan_cols(rows, &prow) : get_plan(rows, &prow))) return rc;
    return cols > 1 ? get_plan_cols(ext, &pext) : get_plan(ext, &pext);
}

int commit_prepare(size_t rows, size_t cols, unsigned blowup, Keep *keep) {
    int rc = commit_check(rows, cols, blowup);
    if (rc) return rc;
    PlanRef prow, pext;
    if ((rc = commit_plans(rows, cols, blowup, prow, pext))) return rc;
    if ((rc = warm_plan(*prow, cols, keep)) || (rc = warm_plan(*pext, cols, keep))) return rc;
    TableRef ct = coset_table(rows), sct = scaled_coset_table(rows);
    PowTable g = power_table(2717, rows);
    if (!ct || !sct || !g.lo || !g.hi) return set_error(GL_E_NOMEM, "twiddle table allocation failed");
    if (keep) keep->insert(keep->end(), {ct, sct, g.lo, g.hi});
    return GL_OK;
}

// Enqueue the whole commit on stream s (no host synchronization):
// coef (rows*cols words) and tmp (blowup*rows*cols words) are scratch.
int commit_enqueue(const u64 *d_m, size_t rows, size_t cols, unsigned blowup, u64 *d_lde, uint8_t *d_layers,
                   u64 *coef, u64 *tmp, hipStream_t s, CommitTree tree) {
    const u64 ext = (u64)rows * blowup;
    PlanRef prow, pext;
    int rc;
    if ((rc = commit_plans(rows, cols, blowup, prow, pext))) return rc;
    // extension input (coef shifted by 2717^j, zero rows j >= rows): when the
    // intt's last pass is a single kernel, the shift joins its 1/rows scale
    // (one product per output either way) and the extension's first pass
    // (two-sub-step or three-stage) reads coef directly, treating the upper
    // rows as zero; otherwise a two-sub-step first extension pass applies
    // the shift itself; otherwise a separate kernel builds the input.
    const int first_kind = pext->passes.empty() ? (int)VK_COMPOSITE : pext->passes[0].v->kind;
    const int last_kind = prow->passes.empty() ? (int)VK_COMPOSITE : prow->passes.back().v->kind;
    const bool single = last_kind == VK_TWO_STAGE || last_kind == VK_THREE_STAGE;
    // zero extension alone: two-sub-step and three-stage first passes; the
    // shift as well: two-sub-step ones
    const bool fold_scale = single && (first_kind == VK_TWO_STAGE || first_kind == VK_THREE_STAGE);
    const bool fold = fold_scale || first_kind == VK_TWO_STAGE;
    // held until the passes reading them are enqueued
    TableRef ct = fold && !fold_scale ? coset_table(rows) : nullptr;
    TableRef sct = fold_scale ? scaled_coset_table(rows) : nullptr;
    if ((fold && !fold_scale && !ct) || (fold_scale && !sct))
        return set_error(GL_E_NOMEM, "twiddle table allocation failed");
    // intt of the columns, reading the input matrix in the first pass
    if (prow->passes.empty()) GL_HIP(hipMemcpyAsync(coef, d_m, rows * cols * 8, hipMemcpyDeviceToDevice, s));
    if ((rc = run_plan(*prow, coef, tmp, cols, 1, rows * cols, true, s, true, d_m, nullptr, 0, sct.get()))) return rc;
    if (fold_scale) {
        if ((rc = run_plan(*pext, d_lde, tmp, cols, 1, ext * cols, false, s, true, coef, nullptr, rows))) return rc;
    } else if (fold) {
        if ((rc = run_plan(*pext, d_lde, tmp, cols, 1, ext * cols, false, s, true, coef, ct.get(), rows))) return rc;
    } else {
        PowTable g = power_table(2717, rows);
        if (!g.lo || !g.hi) return set_error(GL_E_NOMEM, "twiddle table allocation failed");
        launch_coset_extend(coef, d_lde, rows, cols, blowup, g.lo.get(), g.hi.get(), g.lo_bits, s);
        GL_HIP(hipGetLastError());
        if ((rc = run_plan(*pext, d_lde, tmp, cols, 1, ext * cols, false, s))) return rc;
    }
    if (tree == TREE_ROWS) return row_tree_dev_impl(d_lde, ext, cols, d_layers, commit_tree_size(rows, cols, blowup, tree), nullptr, s);
    size_t need = tree_size(ext * cols * 8);
    return merkle_dev_impl((const uint8_t *)d_lde, ext * cols * 8, d_layers, need, nullptr, s);
}

// ---------------------------------------------------------------- pipelined host batch
constexpr size_t kPipeChunk = size_t(32) << 20;  // bytes per chunk (whole transforms, at least one)
constexpr int kPipeUnavailable = -1;              // no helper thread or streams: use the serial path

// A host batch of several chunks (a chunk: the whole transforms that fit in
// kPipeChunk bytes, or one larger transform) runs as a pipeline over two
// device buffers: a helper
// thread copies chunk i + 1 to the device on one stream while the device
// transforms chunk i on the calling thread's stream and the calling thread
// copies chunk i - 1 back on a third, so the two copy directions (separate
// DMA engines) and the transforms overlap.  A buffer is refilled only after
// its chunk is back in host memory.

struct PipeState {
    std::mutex mu;
    std::condition_variable cv;
    bool slot_free[2] = {true, true};
    size_t ready = 0;  // chunks whose copy to the device is enqueued
    bool abort = false;
    int in_rc = GL_OK;
    std::string in_msg;
};

struct PipeEvents {
    hipEvent_t e[5] = {};
    ~PipeEvents() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

int ntt_batch_pipelined(uint64_t *v, size_t n, size_t count, bool inverse) {
    const size_t per = std::max<size_t>(1, kPipeChunk / (n * 8));  // transforms per chunk
    const size_t nchunks = (count + per - 1) / per;
    hipStream_t s = thread_stream(), s_in = thread_aux_stream(0), s_out = thread_aux_stream(1);
    if (!s || !s_in || !s_out) return kPipeUnavailable;
    int dev = 0;
    GL_HIP(hipGetDevice(&dev));
    PlanRef plan;  // builds the plan (and reports an unsupported length) before any copy
    int rc = get_plan(n, &plan);
    if (rc) return rc;
    const size_t cbytes = per * n * 8;
    DevBuf b0(cbytes, s), b1(cbytes, s);
    if (!b0.p || !b1.p) return set_error(GL_E_NOMEM, "device allocation failed");
    u64 *buf[2] = {b0.as<u64>(), b1.as<u64>()};
    PipeEvents ev;  // e[0]: buffers allocated (s); e[1..2]: copy in done (s_in); e[3..4]: transformed (s)
    for (hipEvent_t &x : ev.e) GL_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    hipEvent_t ev_alloc = ev.e[0], *ev_in = ev.e + 1, *ev_done = ev.e + 3;
    GL_HIP(hipEventRecord(ev_alloc, s));
    PipeState st;
    auto copy_in = [&] {
        int irc = GL_OK;
        try {
            hipError_t e = hipSetDevice(dev);
            if (e == hipSuccess) e = hipStreamWaitEvent(s_in, ev_alloc, 0);
            if (e != hipSuccess) irc = hip_error(e, "pipelined batch: copy stream");
            for (size_t i = 0; i < nchunks && irc == GL_OK; i++) {
                const int slot = (int)(i & 1);
                {
                    std::unique_lock<std::mutex> lk(st.mu);
                    st.cv.wait(lk, [&] { return st.slot_free[slot] || st.abort; });
                    if (st.abort) return;
                    st.slot_free[slot] = false;
                }
                const size_t cnt = std::min(per, count - i * per);
                irc = copy_h2d(buf[slot], v + i * per * n, cnt * n * 8, s_in);
                if (irc == GL_OK && (e = hipEventRecord(ev_in[slot], s_in)) != hipSuccess)
                    irc = hip_error(e, "pipelined batch: event");
                if (irc == GL_OK) {
                    std::lock_guard<std::mutex> lk(st.mu);
                    st.ready = i + 1;
                    st.cv.notify_all();
                }
            }
        } catch (...) {
            irc = set_error(GL_E_NOMEM, "pipelined batch: host copy failed");
        }
        if (irc != GL_OK) {
            std::lock_guard<std::mutex> lk(st.mu);
            st.in_rc = irc;
            st.in_msg = last_error_cstr();
            st.abort = true;
            st.cv.notify_all();
        }
    };
    std::thread helper;
    try {
        helper = std::thread(copy_in);
    } catch (...) {
        return kPipeUnavailable;  // no thread: the serial path
    }
    for (size_t i = 0; i < nchunks && rc == GL_OK; i++) {
        const int slot = (int)(i & 1);
        {
            std::unique_lock<std::mutex> lk(st.mu);
            st.cv.wait(lk, [&] { return st.ready > i || st.abort; });
            if (st.ready <= i) break;  // the copy in failed
        }
        const size_t cnt = std::min(per, count - i * per);
        hipError_t e = hipStreamWaitEvent(s, ev_in[slot], 0);
        if (e != hipSuccess) {
            rc = hip_error(e, "pipelined batch: wait");
            break;
        }
        if ((rc = ntt_dev_impl(buf[slot], n, 1, cnt, inverse, s))) break;
        e = hipEventRecord(ev_done[slot], s);
        if (e == hipSuccess) e = hipStreamWaitEvent(s_out, ev_done[slot], 0);
        if (e != hipSuccess) {
            rc = hip_error(e, "pipelined batch: event");
            break;
        }
        if ((rc = copy_d2h(v + i * per * n, buf[slot], cnt * n * 8, s_out, true))) break;
        std::lock_guard<std::mutex> lk(st.mu);
        st.slot_free[slot] = true;
        st.cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(st.mu);
        st.abort = true;  // the helper stops waiting for a free buffer
        st.cv.notify_all();
    }
    helper.join();
    if (rc == GL_OK && st.in_rc != GL_OK) rc = set_error(st.in_rc, st.in_msg);
    if (rc != GL_OK) {
        // nothing may still read or write the buffers when they return to the pool
        (void)hipStreamSynchronize(s_in);
        (void)hipStreamSynchronize(s_out);
        (void)hipStreamSynchronize(s);
        (void)hipGetLastError();
    }
    return rc;
}

}  // namespace gl

extern "C" {

const char *gl_last_error(void) { return last_error_cstr(); }

int gl_init(int device) {
    if (device >= 0) GL_HIP(hipSetDevice(device));
    return ensure_init();
}

int gl_is_divisor(uint64_t n) { return n > 0 && (P - 1) % n == 0; }

uint64_t gl_next_size(uint64_t n) {
    // the 1056 divisors of p - 1 = 2^32 * 3 * 5 * 17 * 257 * 65537, sorted
    static const std::vector<uint64_t> divs = [] {
        std::vector<uint64_t> d;
        const uint64_t odd[] = {3, 5, 17, 257, 65537};
        for (int m = 0; m < 32; m++) {
            for (int k = 0; k <= 32; k++) {
                uint64_t v = 1ull << k;
                for (int q = 0; q < 5; q++)
                    if (m & (1 << q)) v *= odd[q];
                d.push_back(v);
            }
        }
        std::sort(d.begin(), d.end());
        return d;
    }();
    if (n == 0) return 0;
    auto it = std::lower_bound(divs.begin(), divs.end(), n);
    return it == divs.end() ? 0 : *it;
}

// ---------------------------------------------------------------- NTT

void gl_debug_fail_allocs(int count) { set_fail_allocs(0, count); }

void gl_debug_fail_allocs_after(int skip, int count) { set_fail_allocs(skip, count); }

int gl_ntt_dev(uint64_t *d, size_t n, size_t count, int inverse, void *stream) {
    return ntt_dev_impl(d, n, 1, count, inverse != 0, as_stream(stream));
}

int gl_ntt_cols_dev(uint64_t *d, size_t n, size_t cols, size_t count, int inverse, void *stream) {
    return ntt_dev_impl(d, n, cols, count, inverse != 0, as_stream(stream));
}

int gl_ntt_batch(uint64_t *v, size_t n, size_t count, int inverse) {
    int rc = ensure_init();
    if (rc) return rc;
    if (inverse && n <= 1) return GL_OK;
    if (!gl_is_divisor(n)) return get_plan(n, nullptr);
    if (count == 0) return GL_OK;
    if (!v) return set_error(GL_E_ARG, "null pointer");
    if (words_overflow(n, count)) return set_error(GL_E_ARG, "n * count words overflow the address space");
    if (count >= 2 && n * count * 8 >= 2 * kPipeChunk) {
        int prc = ntt_batch_pipelined(v, n, count, inverse != 0);
        if (prc != kPipeUnavailable) return prc;
    }
    hipStream_t s = thread_stream();
    size_t bytes = n * count * 8;
    DevBuf d(bytes, s), tmp;
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    if ((rc = copy_h2d(d.p, v, bytes, s))) return rc;
    if ((rc = ntt_dev_impl(d.as<u64>(), n, 1, count, inverse != 0, s, &tmp))) return rc;
    return copy_back_idle(v, d.p, bytes, s, d, tmp);
}

int gl_ntt(uint64_t *v, size_t n) { return gl_ntt_batch(v, n, 1, 0); }
int gl_intt(uint64_t *v, size_t n) { return gl_ntt_batch(v, n, 1, 1); }

int gl_fill_random_dev(uint64_t *d, size_t n, uint64_t seed, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n && !d) return set_error(GL_E_ARG, "null device pointer");
    launch_fill_random(d, n, seed, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

void gl_clear_cache(void) { clear_caches(); }

size_t gl_variants(uint32_t *r1, uint32_t *r2, uint32_t *t, size_t cap) {
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    size_t total = (size_t)cnt + 1;
    for (size_t i = 0; i < total && i < cap; i++) {
        const PassVariant *v = i < (size_t)cnt ? &vs[i] : rader257_variant();
        if (r1) r1[i] = (uint32_t)v->R1;
        if (r2) r2[i] = (uint32_t)v->R2;
        if (t) t[i] = (uint32_t)v->T;
    }
    return total;
}

// Variant ids: [0, count) compiled pass kernels, count = radix-257 Rader
// kernel, count + 1 = radix-65537 Rader composite (several kernels).
static const PassVariant *variant_by_id(uint32_t id, size_t *count = nullptr) {
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    if (count) *count = (size_t)cnt;
    if (id < (uint32_t)cnt) return &vs[id];
    if (id == (uint32_t)cnt) return rader257_variant();
    if (id == (uint32_t)cnt + 1) return rader65537_variant();
    return nullptr;
}

const char *gl_variant_name(uint32_t id) {
    static std::mutex mu;
    static std::map<uint32_t, std::string> names;
    const PassVariant *v = variant_by_id(id);
    if (!v) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = names.find(id);
    if (it != names.end()) return it->second.c_str();
    char buf[96];
    switch (v->kind) {
    case VK_THREE_STAGE:
        snprintf(buf, sizeof buf, "ntt_pass3_kernel<%d, %d, %d, %d, %d>", v->R1, v->R2 / v->R3, v->R3, v->NT, v->T);
        break;
    case VK_RADER257: snprintf(buf, sizeof buf, "rader257_kernel"); break;
    case VK_COMPOSITE: snprintf(buf, sizeof buf, "rader_big_gather/pointwise/scatter + 2 x NTT_65536"); break;
    default: snprintf(buf, sizeof buf, "ntt_pass_kernel<%d, %d, %d, %d>", v->R1, v->R2, v->NT, v->T);
    }
    return names.emplace(id, buf).first->second.c_str();
}

static int plan_variant_ids(const Plan *plan, uint32_t *ids, size_t cap, size_t *npasses);

int gl_plan_variants(uint64_t n, uint32_t *ids, size_t cap, size_t *npasses) {
    PlanRef plan;
    int rc = get_plan(n, &plan);
    if (rc) return rc;
    return plan_variant_ids(plan.get(), ids, cap, npasses);
}

int gl_plan_variants_cols(uint64_t n, uint32_t *ids, size_t cap, size_t *npasses) {
    PlanRef plan;
    int rc = get_plan_cols(n, &plan);
    if (rc) return rc;
    return plan_variant_ids(plan.get(), ids, cap, npasses);
}

static int plan_variant_ids(const Plan *plan, uint32_t *ids, size_t cap, size_t *npasses) {
    if (npasses) *npasses = plan->passes.size();
    size_t cnt = 0;
    (void)variant_by_id(0, &cnt);
    for (size_t i = 0; i < plan->passes.size() && i < cap; i++) {
        const PassVariant *v = plan->passes[i].v;
        uint32_t id = UINT32_MAX;
        for (uint32_t j = 0; j < cnt + 2; j++)
            if (variant_by_id(j) == v) id = j;
        if (ids) ids[i] = id;
    }
    return GL_OK;
}

uint64_t gl_set_twiddle_full_max(uint64_t words) { return set_full_table_max(words); }

void gl_timing_enable(int on) { timing_enable(on != 0); }

size_t gl_timing_pending(void) { return timing_pending(); }

int gl_timing_collect(uint32_t *variant_ids, float *ms, size_t cap, size_t *count) {
    return timing_collect(variant_ids, ms, cap, count);
}

int gl_plan_set(uint64_t n, const uint32_t *variant_ids, size_t nvariants) {
    if (nvariants && !variant_ids) return set_error(GL_E_ARG, "null pointer");
    return set_plan(n, variant_ids, nvariants);
}

int gl_plan_set_cols(uint64_t n, const uint32_t *variant_ids, size_t nvariants) {
    if (nvariants && !variant_ids) return set_error(GL_E_ARG, "null pointer");
    return set_plan(n, variant_ids, nvariants, true);
}

int gl_plan_passes(uint64_t n, uint32_t *r1, uint32_t *r2, uint32_t *t, size_t cap, size_t *npasses) {
    PlanRef plan;
    int rc = get_plan(n, &plan);
    if (rc) return rc;
    if (npasses) *npasses = plan->passes.size();
    for (size_t i = 0; i < plan->passes.size() && i < cap; i++) {
        if (r1) r1[i] = (uint32_t)plan->passes[i].v->R1;
        if (r2) r2[i] = (uint32_t)plan->passes[i].v->R2;
        if (t) t[i] = (uint32_t)plan->passes[i].v->T;
    }
    return GL_OK;
}

// ---------------------------------------------------------------- K12 / Merkle

size_t gl_next_layer_size(size_t len) { return next_layer(len); }
size_t gl_merkle_tree_size(size_t len) { return tree_size(len); }

int gl_k12(const uint8_t *in, size_t len, uint8_t *out, size_t outlen) {
    int rc = ensure_init();
    if (rc) return rc;
    if (len > 166 || outlen > 168) return set_error(GL_E_ARG, "k12: input > 166 bytes or output > 168 bytes");
    if ((len && !in) || (outlen && !out)) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    DevBuf d(256 + 256, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    uint8_t *din = d.as<uint8_t>(), *dout = din + 256;
    if (len) GL_HIP(hipMemcpyAsync(din, in, len, hipMemcpyHostToDevice, s));
    launch_k12_single(din, len, dout, outlen, s);
    GL_HIP(hipGetLastError());
    if (outlen) GL_HIP(hipMemcpyAsync(out, dout, outlen, hipMemcpyDeviceToHost, s));
    GL_HIP(hipStreamSynchronize(s));
    return GL_OK;
}

int gl_k12_layer_dev(const uint8_t *d_in, size_t len, uint8_t *d_out, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (len && (!d_in || !d_out)) return set_error(GL_E_ARG, "null device pointer");
    if (len == 0) return GL_OK;
    DevBuf stage(kStageBytes, as_stream(stream));
    if (!stage.p) return set_error(GL_E_NOMEM, "scratch allocation failed");
    launch_k12_layer(d_in, len, d_out, stage.as<u64>(), false, nullptr, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_k12_layer(const uint8_t *in, size_t len, uint8_t *out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (len == 0) return GL_OK;
    if (!in || !out) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    size_t olen = next_layer(len);
    const size_t in_pad = (len + 255) / 256 * 256, out_pad = (olen + 255) / 256 * 256;
    DevBuf d(in_pad + out_pad + kStageBytes, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    uint8_t *din = d.as<uint8_t>(), *dout = din + in_pad;
    u64 *stage = (u64 *)(dout + out_pad);
    if ((rc = copy_h2d(din, in, len, s))) return rc;
    launch_k12_layer(din, len, dout, stage, false, nullptr, s);
    GL_HIP(hipGetLastError());
    return copy_back_idle(out, dout, olen, s, d);
}

// ---------------------------------------------------------------- openings

static int merkle_geometry(const uint8_t *leaves, size_t leaf_len, const uint8_t *layers, MerkleGeometry &g) {
    g.depth = 0;
    size_t len = leaf_len, off = 0;
    const uint8_t *p = leaves;
    while (len > HASH) {
        if (g.depth == kMaxMerkleDepth) return set_error(GL_E_UNSUPPORTED, "Merkle tree too deep");
        g.base[g.depth] = p;
        g.len[g.depth] = len;
        g.depth++;
        len = next_layer(len);
        p = layers ? layers + off : nullptr;
        off += len;
    }
    g.base[g.depth] = p;
    g.len[g.depth] = len;
    return GL_OK;
}

size_t gl_merkle_depth(size_t leaf_len) {
    size_t d = 0;
    for (size_t len = leaf_len; len > HASH; len = next_layer(len)) d++;
    return d;
}

int gl_merkle_open_dev(const uint8_t *d_leaves, size_t leaf_len, const uint8_t *d_layers, const uint64_t *d_blocks,
                       size_t k, uint8_t *d_out, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    MerkleGeometry g;
    if ((rc = merkle_geometry(d_leaves, leaf_len, d_layers, g))) return rc;
    if (k == 0 || g.depth == 0) return GL_OK;
    if (!d_leaves || !d_blocks || !d_out || (g.depth > 1 && !d_layers)) return set_error(GL_E_ARG, "null device pointer");
    launch_merkle_open(g, d_blocks, k, d_out, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_merkle_verify_dev(const uint8_t *d_paths, size_t leaf_len, const uint64_t *d_blocks, size_t k,
                         const uint8_t *d_root32, int32_t *d_ok, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    MerkleGeometry g;
    if ((rc = merkle_geometry(nullptr, leaf_len, nullptr, g))) return rc;
    if (k == 0) return GL_OK;
    if (g.depth == 0) return set_error(GL_E_ARG, "a tree of at most 32 bytes has no openings");
    if (!d_paths || !d_blocks || !d_root32 || !d_ok) return set_error(GL_E_ARG, "null device pointer");
    launch_merkle_verify(g, d_paths, d_blocks, k, d_root32, (int *)d_ok, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_merkle_tree_dev(const uint8_t *d_in, size_t len, uint8_t *d_out, size_t cap, size_t *out_len,
                       void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (len > HASH && (!d_in || !d_out)) return set_error(GL_E_ARG, "null device pointer");
    return merkle_dev_impl(d_in, len, d_out, cap, out_len, as_stream(stream));
}

int gl_merkle_tree(const uint8_t *in, size_t len, uint8_t *out, size_t cap, size_t *out_len) {
    int rc = ensure_init();
    if (rc) return rc;
    size_t need = tree_size(len);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the Merkle tree");
    if (need == 0) return GL_OK;
    if (!in || !out) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    size_t in_pad = (len + 255) / 256 * 256;
    DevBuf d(in_pad + need, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    uint8_t *din = d.as<uint8_t>(), *dout = din + in_pad;
    if ((rc = copy_h2d(din, in, len, s))) return rc;
    if ((rc = merkle_dev_impl(din, len, dout, need, nullptr, s))) return rc;
    return copy_back_idle(out, dout, need, s, d);
}

int gl_pow_solve(const uint8_t *challenge32, uint64_t threshold, uint64_t max_nonce, uint64_t *nonce) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!challenge32 || !nonce) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    DevBuf d(256, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    unsigned long long *best = (unsigned long long *)d.p;
    uint8_t *dch = (uint8_t *)d.p + 64;
    const unsigned long long none = ~0ull;
    GL_HIP(hipMemcpyAsync(best, &none, 8, hipMemcpyHostToDevice, s));
    GL_HIP(hipMemcpyAsync(dch, challenge32, 32, hipMemcpyHostToDevice, s));
    unsigned long long found = none;
    const u64 chunk = 1ull << 26;
    for (u64 start = 0; start < max_nonce; start += chunk) {
        u64 cnt = max_nonce - start < chunk ? max_nonce - start : chunk;
        launch_pow_search(dch, threshold, start, cnt, best, s);
        GL_HIP(hipGetLastError());
        GL_HIP(hipMemcpyAsync(&found, best, 8, hipMemcpyDeviceToHost, s));
        GL_HIP(hipStreamSynchronize(s));
        if (found != none) break;  // every smaller nonce was tested in earlier chunks
    }
    if (found == none) return set_error(GL_E_ARG, "pow: no nonce below max_nonce");
    *nonce = found;
    return GL_OK;
}

int gl_convolve_dev(uint64_t *d_a, const uint64_t *d_b_ntt, size_t n, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n && (!d_a || !d_b_ntt)) return set_error(GL_E_ARG, "null device pointer");
    hipStream_t s = as_stream(stream);
    PlanRef plan;
    if ((rc = get_plan(n, &plan))) return rc;
    const int last_kind = plan->passes.empty() ? (int)VK_COMPOSITE : plan->passes.back().v->kind;
    if (last_kind == VK_TWO_STAGE || last_kind == VK_THREE_STAGE) {
        // the forward transform's last pass multiplies output k by b[k] as
        // it stores it (the pointwise step, fused: no extra pass over a)
        const size_t words = plan_tmp_words(*plan, 1, n);
        DevBuf tmp(words * 8, s);
        if (words && !tmp.p) return set_error(GL_E_NOMEM, "scratch allocation failed");
        if ((rc = run_plan(*plan, d_a, tmp.as<u64>(), 1, 1, n, false, s, true, nullptr, nullptr, 0, d_b_ntt)))
            return rc;
    } else {
        // n = 1, or a radix-257 / radix-65537 last pass: separate pointwise kernel
        if ((rc = ntt_dev_impl(d_a, n, 1, 1, false, s))) return rc;
        launch_pointwise_mont(d_a, d_b_ntt, n, s);
        GL_HIP(hipGetLastError());
    }
    return ntt_dev_impl(d_a, n, 1, 1, true, s);
}

// ---------------------------------------------------------------- permutations

// Shape checks shared by the transposes: GL_OK for an empty or valid shape.
// One workgroup per 32 x 32 tile, at most 2^24 - 1 workgroups per launch.
static int transpose_check(size_t rows, size_t cols) {
    if (words_overflow(rows, cols)) return set_error(GL_E_ARG, "transpose shape overflows the address space");
    if (((u64)rows + 31) / 32 * (((u64)cols + 31) / 32) >= (1ull << 24))
        return set_error(GL_E_SIZE, "transpose: more than 2^34 words");
    return GL_OK;
}

int gl_transpose_oop_dev(const uint64_t *d_src, uint64_t *d_dst, size_t rows, size_t cols, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = transpose_check(rows, cols))) return rc;
    if (rows == 0 || cols == 0) return GL_OK;
    if (!d_src || !d_dst) return set_error(GL_E_ARG, "null device pointer");
    if (d_src == d_dst) return set_error(GL_E_ARG, "out-of-place transpose needs distinct buffers");
    launch_transpose(d_src, d_dst, rows, cols, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_transpose_dev(uint64_t *d, size_t rows, size_t cols, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = transpose_check(rows, cols))) return rc;
    if (rows == 0 || cols == 0 || rows == 1 || cols == 1) return GL_OK;  // a vector is its own transpose
    if (!d) return set_error(GL_E_ARG, "null device pointer");
    hipStream_t s = as_stream(stream);
    if (rows == cols) {
        launch_transpose_square(d, rows, s);
        GL_HIP(hipGetLastError());
        return GL_OK;
    }
    const size_t bytes = rows * cols * 8;
    DevBuf tmp(bytes, s);  // stream ordered: back to the pool after the copy
    if (!tmp.p) return set_error(GL_E_NOMEM, "transpose scratch allocation failed");
    launch_transpose(d, tmp.as<u64>(), rows, cols, s);
    GL_HIP(hipGetLastError());
    GL_HIP(hipMemcpyAsync(d, tmp.p, bytes, hipMemcpyDeviceToDevice, s));
    return GL_OK;
}

int gl_transpose(uint64_t *m, size_t rows, size_t cols) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = transpose_check(rows, cols))) return rc;
    if (rows == 0 || cols == 0 || rows == 1 || cols == 1) return GL_OK;
    if (!m) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    const size_t bytes = rows * cols * 8;
    DevBuf a(bytes, s), b(rows == cols ? 0 : bytes, s);
    if (!a.p || (rows != cols && !b.p)) return set_error(GL_E_NOMEM, "device allocation failed");
    if ((rc = copy_h2d(a.p, m, bytes, s))) return rc;
    if (rows == cols) {
        launch_transpose_square(a.as<u64>(), rows, s);
        GL_HIP(hipGetLastError());
        return copy_back_idle(m, a.p, bytes, s, a);
    }
    launch_transpose(a.as<u64>(), b.as<u64>(), rows, cols, s);
    GL_HIP(hipGetLastError());
    return copy_back_idle(m, b.p, bytes, s, a, b);
}

static int bit_reverse_log2(size_t n, int *k) {
    if (n == 0 || (n & (n - 1))) return set_error(GL_E_SIZE, "bit_reverse: length must be a power of two");
    if (n > (size_t(1) << 32)) return set_error(GL_E_SIZE, "bit_reverse: length above 2^32");
    int b = 0;
    while ((size_t(1) << b) < n) b++;
    *k = b;
    return GL_OK;
}

int gl_bit_reverse_dev(uint64_t *d, size_t n, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    int k = 0;
    if ((rc = bit_reverse_log2(n, &k))) return rc;
    if (n == 1) return GL_OK;
    if (!d) return set_error(GL_E_ARG, "null device pointer");
    launch_bit_reverse(d, n, k, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_bit_reverse(uint64_t *v, size_t n) {
    int rc = ensure_init();
    if (rc) return rc;
    int k = 0;
    if ((rc = bit_reverse_log2(n, &k))) return rc;
    if (n == 1) return GL_OK;
    if (!v) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    DevBuf d(n * 8, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    if ((rc = copy_h2d(d.p, v, n * 8, s))) return rc;
    launch_bit_reverse(d.as<u64>(), n, k, s);
    GL_HIP(hipGetLastError());
    return copy_back_idle(v, d.p, n * 8, s, d);
}

int gl_convolve(uint64_t *a, const uint64_t *b_ntt, size_t n) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!gl_is_divisor(n)) return set_error(GL_E_SIZE, "convolve: length does not divide p - 1");
    if (!a || !b_ntt) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    DevBuf d(2 * n * 8, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    u64 *da = d.as<u64>(), *db = da + n;
    if ((rc = copy_h2d(da, a, n * 8, s))) return rc;
    if ((rc = copy_h2d(db, b_ntt, n * 8, s))) return rc;
    if ((rc = gl_convolve_dev(da, db, n, s))) return rc;
    return copy_back_idle(a, da, n * 8, s, d);
}

}  // extern "C"

// ---------------------------------------------------------------- commit

namespace {

int commit_dev(const uint64_t *d_m, size_t rows, size_t cols, unsigned blowup, uint64_t *d_lde, uint8_t *d_layers,
               size_t cap, size_t *out_len, hipStream_t s, CommitTree tree) {
    int rc;
    if ((rc = commit_check(rows, cols, blowup, tree))) return rc;
    const u64 ext = (u64)rows * blowup;
    size_t need = commit_tree_size(rows, cols, blowup, tree);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the commit's tree");
    if (!d_m || !d_lde || !d_layers) return set_error(GL_E_ARG, "null device pointer");
    if (tree == TREE_ROWS && (uintptr_t)d_layers % 8) return set_error(GL_E_ARG, "row tree output must be 8-byte aligned");
    // scratch returns to the pool stream ordered: no wait needed
    DevBuf coef(rows * cols * 8, s);
    DevBuf tmp(ext * cols * 8, s);
    if (!coef.p || !tmp.p) return set_error(GL_E_NOMEM, "commit scratch allocation failed");
    return commit_enqueue(d_m, rows, cols, blowup, d_lde, d_layers, coef.as<u64>(), tmp.as<u64>(), s, tree);
}

int commit_host(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out, uint8_t *layers_out,
                size_t cap, size_t *out_len, CommitTree tree) {
    int rc;
    if ((rc = commit_check(rows, cols, blowup, tree))) return rc;
    const u64 ext = (u64)rows * blowup;
    size_t need = commit_tree_size(rows, cols, blowup, tree);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the commit's tree");
    if (!m || !lde_out || !layers_out) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    DevBuf dm(rows * cols * 8, s), dl(ext * cols * 8, s), dt(need + 256, s);
    if (!dm.p || !dl.p || !dt.p) return set_error(GL_E_NOMEM, "device allocation failed");
    if ((rc = copy_h2d(dm.p, m, rows * cols * 8, s))) return rc;
    if ((rc = commit_dev(dm.as<u64>(), rows, cols, blowup, dl.as<u64>(), dt.as<uint8_t>(), need, nullptr, s, tree)))
        return rc;
    if ((rc = copy_d2h(lde_out, dl.p, ext * cols * 8, s))) return rc;
    return copy_back_idle(layers_out, dt.p, need, s, dm, dl, dt);
}

}  // namespace

extern "C" {

int gl_commit_dev(const uint64_t *d_m, size_t rows, size_t cols, unsigned blowup, uint64_t *d_lde,
                  uint8_t *d_layers, size_t cap, size_t *out_len, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    return commit_dev(d_m, rows, cols, blowup, d_lde, d_layers, cap, out_len, as_stream(stream), TREE_BYTES);
}

int gl_commit(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out,
              uint8_t *layers_out, size_t cap, size_t *out_len) {
    int rc = ensure_init();
    if (rc) return rc;
    return commit_host(m, rows, cols, blowup, lde_out, layers_out, cap, out_len, TREE_BYTES);
}

int gl_commit_rows_dev(const uint64_t *d_m, size_t rows, size_t cols, unsigned blowup, uint64_t *d_lde,
                       uint8_t *d_layers, size_t cap, size_t *out_len, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    return commit_dev(d_m, rows, cols, blowup, d_lde, d_layers, cap, out_len, as_stream(stream), TREE_ROWS);
}

int gl_commit_rows(const uint64_t *m, size_t rows, size_t cols, unsigned blowup, uint64_t *lde_out,
                   uint8_t *layers_out, size_t cap, size_t *out_len) {
    int rc = ensure_init();
    if (rc) return rc;
    return commit_host(m, rows, cols, blowup, lde_out, layers_out, cap, out_len, TREE_ROWS);
}

// ---------------------------------------------------------------- row tree

size_t gl_row_tree_size(size_t leaves) { return 32 * row_tree_nodes(leaves); }

int gl_row_tree_dev(const uint64_t *d_rows, size_t nrows, size_t cols, uint8_t *d_layers, size_t cap,
                    size_t *out_len, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    return row_tree_dev_impl(d_rows, nrows, cols, d_layers, cap, out_len, as_stream(stream));
}

int gl_row_tree(const uint64_t *rows, size_t nrows, size_t cols, uint8_t *layers_out, size_t cap, size_t *out_len) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = row_tree_check(nrows, cols))) return rc;
    const size_t need = gl_row_tree_size(nrows);
    if (out_len) *out_len = need;
    if (cap < need) return set_error(GL_E_ARG, "output buffer smaller than the row tree");
    if (!rows || !layers_out) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = thread_stream();
    const size_t in_bytes = nrows * cols * 8, in_pad = (in_bytes + 255) / 256 * 256;
    DevBuf d(in_pad + need, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    uint8_t *din = d.as<uint8_t>(), *dout = din + in_pad;
    if ((rc = copy_h2d(din, rows, in_bytes, s))) return rc;
    if ((rc = row_tree_dev_impl((const u64 *)din, nrows, cols, dout, need, nullptr, s))) return rc;
    return copy_back_idle(layers_out, dout, need, s, d);
}

size_t gl_row_tree_depth(size_t leaves) {
    size_t d = 0;
    for (size_t n = leaves; n > 1; n = (n + 1) / 2) d++;
    return d;
}

int gl_row_open_dev(const uint8_t *d_layers, size_t leaves, const uint64_t *d_idx, size_t k, uint8_t *d_out,
                    void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    RowGeometry g;
    if (leaves == 0) return set_error(GL_E_ARG, "empty row tree");
    if ((rc = row_geometry(leaves, g))) return rc;
    if (k == 0 || g.depth == 0) return GL_OK;
    if (!d_layers || !d_idx || !d_out) return set_error(GL_E_ARG, "null device pointer");
    if ((uintptr_t)d_layers % 8 || (uintptr_t)d_out % 8) return set_error(GL_E_ARG, "row tree buffers must be 8-byte aligned");
    launch_row_open(g, (const u64 *)d_layers, d_idx, k, (u64 *)d_out, as_stream(stream));
    GL_HIP(hipGetLastError());
    return GL_OK;
}

int gl_row_verify_dev(const uint64_t *d_rows, size_t cols, size_t leaves, const uint64_t *d_idx, size_t k,
                      const uint8_t *d_paths, const uint8_t *d_root32, int32_t *d_ok, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = row_tree_check(leaves, cols))) return rc;
    RowGeometry g;
    if ((rc = row_geometry(leaves, g))) return rc;
    if (k == 0) return GL_OK;
    if (!d_rows || !d_idx || !d_root32 || !d_ok || (g.depth && !d_paths))
        return set_error(GL_E_ARG, "null device pointer");
    if ((uintptr_t)d_rows % 8 || (uintptr_t)d_paths % 8 || (uintptr_t)d_root32 % 8)
        return set_error(GL_E_ARG, "row tree buffers must be 8-byte aligned");
    hipStream_t s = as_stream(stream);
    // the opened rows' leaf hashes, then the walks up to the root
    const u64 cvw = row_tree_cv_words(k, (u32)cols);
    DevBuf hashes(k * 32, s), cv(cvw * 8, s);
    if (!hashes.p || (cvw && !cv.p)) return set_error(GL_E_NOMEM, "row verify scratch allocation failed");
    launch_row_leaves(d_rows, k, (u32)cols, hashes.as<u64>(), cv.as<u64>(), s);
    launch_row_verify(g, hashes.as<u64>(), d_idx, k, (const u64 *)d_paths, (const u64 *)d_root32, (int *)d_ok, s);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

}  // extern "C"
