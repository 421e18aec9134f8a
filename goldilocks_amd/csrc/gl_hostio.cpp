// Host <-> device copies of the host-buffer entry points (gl_ntt, gl_intt,
// gl_ntt_batch, gl_ntt_batch_mixed, gl_merkle_tree, gl_k12_layer,
// gl_convolve, gl_commit, gl_row_tree), the reference's slice-based API.
//
// On the MI355X box a pageable hipMemcpy of 128 MiB takes about 5 ms to the
// device and 8 ms back (26 and 17 GB/s), DMA from pinned host memory about
// 2.4 ms either way (57 GB/s), and a memcpy from pageable into pinned memory
// split over eight host threads about 1.3 ms (100 GB/s).  Copies above a few
// MiB therefore go through pinned staging chunks: host threads fill chunk
// i + 1 while the DMA engine moves chunk i, so a large copy runs at the DMA
// rate.  Small copies keep the plain pageable path (lower latency).
//
// Staging chunk sets are pooled (one set per concurrent large copy, up to
// kMaxSets; beyond that a copy falls back to the pageable path), and a
// chunk is overwritten only after the event recorded behind its last DMA
// has completed, so a set released with transfers in flight is safe to hand
// to the next caller.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

#include "gl_runtime.hpp"

namespace gl {
namespace {

constexpr size_t kChunk = size_t(8) << 20;      // bytes per staging chunk
constexpr int kDepth = 3;                       // chunks per set (in flight + being filled)
constexpr size_t kDirectMax = size_t(4) << 20;  // at most this many bytes: pageable copy
constexpr int kMaxSets = 8;
constexpr size_t kPart = size_t(1) << 20;       // least bytes per memcpy task

struct Staging {
    char *buf[kDepth] = {};
    hipEvent_t ev[kDepth] = {};
    bool busy[kDepth] = {};  // ev[i] is recorded behind the last DMA that used buf[i]
};

std::mutex g_stage_mu;
std::vector<Staging *> g_stage_free;
int g_stage_sets = 0;

void stage_destroy(Staging *st) {
    for (int i = 0; i < kDepth; i++) {
        if (st->buf[i]) (void)hipHostFree(st->buf[i]);
        if (st->ev[i]) (void)hipEventDestroy(st->ev[i]);
    }
    delete st;
}

// A set for one copy, or nullptr (use the pageable path).
Staging *stage_acquire() {
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        if (!g_stage_free.empty()) {
            Staging *st = g_stage_free.back();
            g_stage_free.pop_back();
            return st;
        }
        if (g_stage_sets >= kMaxSets) return nullptr;
        g_stage_sets++;
    }
    Staging *st = new Staging();
    for (int i = 0; i < kDepth; i++) {
        if (hipHostMalloc((void **)&st->buf[i], kChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&st->ev[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            stage_destroy(st);
            std::lock_guard<std::mutex> lk(g_stage_mu);
            g_stage_sets--;
            return nullptr;
        }
    }
    return st;
}

void stage_release(Staging *st) {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    g_stage_free.push_back(st);
}

// Block until buf[i] may be overwritten.
int stage_wait(Staging *st, int i) {
    if (!st->busy[i]) return GL_OK;
    GL_HIP(hipEventSynchronize(st->ev[i]));
    st->busy[i] = false;
    return GL_OK;
}

// After a failed enqueue: wait for the stream so that no chunk is still
// read or written by a transfer without an event behind it.
void stage_settle(Staging *st, hipStream_t s) {
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    for (int i = 0; i < kDepth; i++) st->busy[i] = false;
}

// Host threads for the memcpy halves of staged copies.  The pool object is
// never destroyed: its threads wait on it until the process exits.
struct CopyPool {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    int threads = 0;
    pid_t pid = 0;  // the process whose threads these are (a forked child has none)
};

CopyPool *copy_pool() {
    static CopyPool *pool = [] {
        CopyPool *c = new CopyPool();
        c->pid = getpid();
        const unsigned hw = std::thread::hardware_concurrency();
        const int want = hw > 1 ? (int)std::min(hw - 1, 7u) : 0;
        for (int i = 0; i < want; i++) {
            try {
                std::thread([c] {
                    for (;;) {
                        std::function<void()> f;
                        {
                            std::unique_lock<std::mutex> lk(c->mu);
                            c->cv.wait(lk, [c] { return !c->q.empty(); });
                            f = std::move(c->q.front());
                            c->q.pop_front();
                        }
                        f();
                    }
                }).detach();
                c->threads++;
            } catch (...) {
                break;  // fewer helper threads
            }
        }
        return c;
    }();
    return pool;
}

// memcpy split over the calling thread and the pool's threads.
void par_memcpy(void *dst, const void *src, size_t n) {
    CopyPool *c = copy_pool();
    const int helpers = c->pid == getpid() ? c->threads : 0;
    const int parts = (int)std::min<size_t>((size_t)helpers + 1, n / kPart);
    if (parts <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    // parts of `per` bytes (a multiple of 64); n / parts >= kPart keeps the
    // last part non-empty
    const size_t per = (n / parts + 63) / 64 * 64;
    struct Latch {
        std::mutex mu;
        std::condition_variable cv;
        int left = 0;
    };
    // shared with the tasks, so it outlives the last one's notify
    auto latch = std::make_shared<Latch>();
    latch->left = parts - 1;
    char *d = (char *)dst;
    const char *s = (const char *)src;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        for (int k = 1; k < parts; k++) {
            const size_t a = per * k, b = k == parts - 1 ? n : per * (k + 1);
            c->q.emplace_back([d, s, a, b, latch] {
                std::memcpy(d + a, s + a, b - a);
                std::lock_guard<std::mutex> g(latch->mu);
                if (--latch->left == 0) latch->cv.notify_one();
            });
        }
    }
    c->cv.notify_all();
    std::memcpy(d, s, per);
    std::unique_lock<std::mutex> lk(latch->mu);
    latch->cv.wait(lk, [&latch] { return latch->left == 0; });
}

// Pinned host slabs for small uploads that must stay asynchronous (launch
// descriptors): a slab returns to the pool with an event recorded behind the
// uploads that read it, and is handed out again only once that event has
// completed.
constexpr int kMaxSlabs = 8;
std::mutex g_slab_mu;
std::vector<PinnedSlab> g_slab_free;
int g_slabs = 0;

bool slab_ready(const PinnedSlab &sl) {
    const hipError_t e = hipEventQuery(sl.ev);
    if (e == hipSuccess) return true;
    (void)hipGetLastError();  // hipErrorNotReady
    return false;
}

}  // namespace

int pinned_acquire(size_t bytes, PinnedSlab &out) {
    out = PinnedSlab();
    PinnedSlab old;  // a free slab taken over: reused once ready, or freed if too small
    {
        std::lock_guard<std::mutex> lk(g_slab_mu);
        for (size_t i = 0; i < g_slab_free.size(); i++) {
            if (g_slab_free[i].bytes >= bytes && slab_ready(g_slab_free[i])) {
                out = g_slab_free[i];
                g_slab_free.erase(g_slab_free.begin() + (long)i);
                return GL_OK;
            }
        }
        if (g_slabs < kMaxSlabs || g_slab_free.empty()) {
            // a new slab, allocated below (beyond kMaxSlabs only while more
            // callers than that hold one at once)
            g_slabs++;
        } else {
            size_t best = 0;  // the largest free slab, still in use by uploads
            for (size_t i = 1; i < g_slab_free.size(); i++)
                if (g_slab_free[i].bytes > g_slab_free[best].bytes) best = i;
            old = g_slab_free[best];
            g_slab_free.erase(g_slab_free.begin() + (long)best);
        }
    }
    if (old.p) {
        (void)hipEventSynchronize(old.ev);  // its uploads have to finish first
        (void)hipGetLastError();
        if (old.bytes >= bytes) {
            out = old;
            return GL_OK;
        }
        (void)hipEventDestroy(old.ev);
        (void)hipHostFree(old.p);
        (void)hipGetLastError();
    }
    size_t sz = size_t(64) << 10;
    while (sz < bytes) sz *= 2;
    PinnedSlab sl;
    sl.bytes = sz;
    if (hipHostMalloc(&sl.p, sz, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (sl.p) (void)hipHostFree(sl.p);
        std::lock_guard<std::mutex> lk(g_slab_mu);
        g_slabs--;
        return set_error(GL_E_NOMEM, "pinned descriptor slab allocation failed");
    }
    out = sl;
    return GL_OK;
}

void pinned_release(PinnedSlab &slab, hipStream_t s) {
    if (!slab.p) return;
    if (hipEventRecord(slab.ev, s) != hipSuccess) {
        // no event behind the uploads: wait for them before the slab is reused
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        (void)hipGetLastError();
        (void)hipEventRecord(slab.ev, nullptr);
        (void)hipGetLastError();
    }
    std::lock_guard<std::mutex> lk(g_slab_mu);
    g_slab_free.push_back(slab);
    slab = PinnedSlab();
}

int copy_h2d(void *d_dst, const void *h_src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return GL_OK;
    Staging *st = bytes > kDirectMax ? stage_acquire() : nullptr;
    if (!st) {
        GL_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, s));
        return GL_OK;
    }
    int rc = GL_OK;
    int i = 0;
    for (size_t off = 0; off < bytes; off += kChunk, i = (i + 1) % kDepth) {
        const size_t len = std::min(kChunk, bytes - off);
        if ((rc = stage_wait(st, i))) break;
        par_memcpy(st->buf[i], (const char *)h_src + off, len);
        hipError_t e = hipMemcpyAsync((char *)d_dst + off, st->buf[i], len, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(st->ev[i], s);
        if (e != hipSuccess) {
            rc = hip_error(e, "staged copy to the device");
            break;
        }
        st->busy[i] = true;
    }
    if (rc) stage_settle(st, s);
    stage_release(st);
    return rc;
}

int copy_d2h(void *h_dst, const void *d_src, size_t bytes, hipStream_t s, bool sync) {
    if (bytes == 0) {
        if (sync) GL_HIP(hipStreamSynchronize(s));
        return GL_OK;
    }
    Staging *st = bytes > kDirectMax ? stage_acquire() : nullptr;
    if (!st) {
        GL_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
        if (sync) GL_HIP(hipStreamSynchronize(s));
        return GL_OK;
    }
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    int rc = GL_OK;
    size_t issued = 0, drained = 0;
    while (drained < chunks && rc == GL_OK) {
        // keep up to kDepth chunks in flight
        while (issued < chunks && issued < drained + kDepth) {
            const int i = (int)(issued % kDepth);
            if ((rc = stage_wait(st, i))) break;  // a previous caller's transfer
            const size_t off = issued * kChunk, len = std::min(kChunk, bytes - off);
            hipError_t e = hipMemcpyAsync(st->buf[i], (const char *)d_src + off, len, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipEventRecord(st->ev[i], s);
            if (e != hipSuccess) {
                rc = hip_error(e, "staged copy to the host");
                break;
            }
            st->busy[i] = true;
            issued++;
        }
        if (rc) break;
        const int i = (int)(drained % kDepth);
        if ((rc = stage_wait(st, i))) break;
        const size_t off = drained * kChunk, len = std::min(kChunk, bytes - off);
        par_memcpy((char *)h_dst + off, st->buf[i], len);
        drained++;
    }
    if (rc) stage_settle(st, s);
    stage_release(st);
    return rc;
}

}  // namespace gl
