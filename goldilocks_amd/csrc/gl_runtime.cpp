// Host runtime for the gfx950 Goldilocks library: errors, device pool,
// twiddle tables, the pass planner and plan execution.
#include "gl_runtime.hpp"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <dlfcn.h>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "gl_field.hpp"

namespace gl {

// ---------------------------------------------------------------- errors
static thread_local std::string t_error;

int set_error(int code, const std::string &msg) {
    t_error = msg;
    return code;
}

int hip_error(hipError_t e, const char *what) {
    int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GL_E_NOMEM : GL_E_DEVICE;
    return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

const char *last_error_cstr() { return t_error.c_str(); }

static std::once_flag g_init_once;
static int g_init_rc = GL_OK;
static std::string g_init_msg;

static void load_tune_file();
static std::once_flag g_tune_once;

int ensure_init() {
    std::call_once(g_init_once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            g_init_rc = GL_E_DEVICE;
            g_init_msg = "no HIP device available";
            return;
        }
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            g_init_rc = GL_E_DEVICE;
            g_init_msg = "hipGetDeviceProperties failed";
            return;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            g_init_rc = GL_E_DEVICE;
            g_init_msg = std::string("library is built for gfx950, device is ") + prop.gcnArchName;
            return;
        }
        std::call_once(g_tune_once, load_tune_file);
    });
    if (g_init_rc != GL_OK) return set_error(g_init_rc, g_init_msg);
    return GL_OK;
}

// ---------------------------------------------------------------- pool
// Free blocks remember the stream that last used them and an event recorded
// there at the free.  A block is handed out again at once for the same
// stream (stream order runs the earlier user first), to any other stream
// once its event has completed.  Blocks freed on a capturing stream become
// part of a graph and are never reused.  On the null stream, which is never
// destroyed, the event is recorded only when another stream asks for the
// block (it then marks all null-stream work enqueued so far, the block's
// last use included): back-to-back calls on the null stream record none.
namespace {
struct FreeBlock {
    void *p;
    hipStream_t stream;
    hipEvent_t ev;  // nullptr and !pending: no pending use
    bool pending;   // last used on the null stream; no event recorded yet
};
std::mutex g_pool_mu;
std::multimap<size_t, FreeBlock> g_pool_free;
std::map<void *, size_t> g_pool_size;
std::vector<hipEvent_t> g_pool_events;

size_t round_size(size_t b) {
    const size_t MB = 1 << 20;
    if (b >= MB) return (b + MB - 1) / MB * MB;
    return (b + 4095) / 4096 * 4096;
}

// Work enqueued on s runs after earlier work on s.  hipStreamPerThread is
// one handle for a different stream in every thread, so it never counts.
bool same_stream(hipStream_t a, hipStream_t b) { return a == b && a != hipStreamPerThread; }

// caller holds g_pool_mu
hipEvent_t take_event() {
    hipEvent_t e = nullptr;
    if (!g_pool_events.empty()) {
        e = g_pool_events.back();
        g_pool_events.pop_back();
    } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        e = nullptr;
    }
    return e;
}

// caller holds g_pool_mu
void recycle_event(hipEvent_t e) {
    if (e) g_pool_events.push_back(e);
}

// caller holds g_pool_mu.  A pending block gets its event now; without one
// (no event could be made) the null stream is synchronized instead.
void settle_pending(FreeBlock &fb) {
    if (!fb.pending) return;
    fb.pending = false;
    hipEvent_t e = take_event();
    if (e && hipEventRecord(e, nullptr) == hipSuccess) {
        fb.ev = e;
        return;
    }
    (void)hipGetLastError();
    if (e) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(nullptr);
}

// caller holds g_pool_mu
bool block_ready(FreeBlock &fb, hipStream_t s) {
    if (same_stream(fb.stream, s)) return true;
    settle_pending(fb);
    if (!fb.ev) return true;
    hipError_t q = hipEventQuery(fb.ev);
    if (q == hipSuccess) return true;
    if (q != hipErrorNotReady) (void)hipGetLastError();
    return false;
}
}  // namespace

// Allocations to fail on purpose (gl_debug_fail_allocs[_after]): after
// `skip` more allocations succeed, the next `count` fail.  A host-side switch
// for tests of the error paths; no device work is involved.
static std::mutex g_fail_mu;
static std::atomic<bool> g_fail_active{false};
static int g_fail_skip = 0, g_fail_count = 0;

void set_fail_allocs(int skip, int count) {
    std::lock_guard<std::mutex> lk(g_fail_mu);
    g_fail_skip = skip > 0 ? skip : 0;
    g_fail_count = count > 0 ? count : 0;
    g_fail_active.store(g_fail_count > 0);
}

static bool fail_this_alloc() {
    if (!g_fail_active.load()) return false;
    std::lock_guard<std::mutex> lk(g_fail_mu);
    if (g_fail_count == 0) return false;
    if (g_fail_skip > 0) {
        g_fail_skip--;
        return false;
    }
    if (--g_fail_count == 0) g_fail_active.store(false);
    return true;
}

void *pool_alloc(size_t bytes, hipStream_t s) {
    if (fail_this_alloc()) return nullptr;
    const size_t sz = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto lo = g_pool_free.lower_bound(sz), hi = g_pool_free.upper_bound(sz + sz / 4);
        for (int pass = 0; pass < 2; pass++) {
            for (auto it = lo; it != hi; ++it) {
                FreeBlock &fb = it->second;
                const bool ok = pass == 0 ? ((!fb.ev && !fb.pending) || same_stream(fb.stream, s))
                                          : block_ready(fb, s);
                if (!ok) continue;
                void *p = it->second.p;
                recycle_event(it->second.ev);
                g_pool_free.erase(it);
                return p;
            }
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        pool_release_idle();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool_size[p] = sz;
    return p;
}

void pool_free(void *p, hipStream_t s) {
    if (!p) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        cs = hipStreamCaptureStatusNone;
    }
    if (cs != hipStreamCaptureStatusNone) return;  // owned by the captured graph from now on
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_pool_size.find(p);
    if (it == g_pool_size.end()) return;
    if (s == nullptr) {
        g_pool_free.emplace(it->second, FreeBlock{p, s, nullptr, true});
        return;
    }
    hipEvent_t e = take_event();
    if (e && hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(e);
        e = nullptr;
    }
    // no event: only safe on the same stream; wait for it instead
    if (!e) (void)hipStreamSynchronize(s);
    g_pool_free.emplace(it->second, FreeBlock{p, s, e, false});
}

void pool_free_idle(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_pool_size.find(p);
    if (it == g_pool_size.end()) return;
    g_pool_free.emplace(it->second, FreeBlock{p, nullptr, nullptr, false});
}

void pool_release_idle() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (auto it = g_pool_free.begin(); it != g_pool_free.end();) {
        FreeBlock &fb = it->second;
        bool done = true;
        if (fb.pending) {
            settle_pending(fb);
            if (fb.ev) (void)hipEventSynchronize(fb.ev);  // its null-stream work first
        }
        if (fb.ev) {
            hipError_t q = hipEventQuery(fb.ev);
            if (q != hipSuccess) {
                if (q != hipErrorNotReady) (void)hipGetLastError();
                done = false;
            }
        }
        if (!done) {
            ++it;
            continue;
        }
        recycle_event(fb.ev);
        g_pool_size.erase(fb.p);
        (void)hipFree(fb.p);
        it = g_pool_free.erase(it);
    }
}

struct ThreadStream {
    hipStream_t s = nullptr;
    ~ThreadStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
static thread_local ThreadStream t_stream;

hipStream_t thread_stream() {
    if (!t_stream.s) {
        if (hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            t_stream.s = nullptr;
        }
    }
    return t_stream.s;
}

static thread_local ThreadStream t_aux[2];

hipStream_t thread_aux_stream(int i) {
    ThreadStream &t = t_aux[i & 1];
    if (!t.s) {
        if (hipStreamCreateWithFlags(&t.s, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            t.s = nullptr;
        }
    }
    return t.s;
}

// ---------------------------------------------------------------- twiddles
// Retired tables (last reference dropped) wait here until clear_caches()
// frees them after a device synchronization.
static std::mutex g_grave_mu;
static std::vector<void *> g_grave;

static TableRef adopt(void *d) {
    return TableRef((const u64 *)d, [](const u64 *p) {
        std::lock_guard<std::mutex> lk(g_grave_mu);
        g_grave.push_back((void *)p);
    });
}

// Table generators run on a private stream, so building a table waits for
// that stream only, never for work the callers have in flight.
static hipStream_t table_stream() {
    static std::once_flag once;
    static hipStream_t s = nullptr;
    std::call_once(once, [] {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            s = nullptr;
        }
    });
    return s;
}

static std::mutex g_tw_mu;
static std::map<std::pair<u64, u64>, PowTable> g_pow_tables;  // (base, count)
static std::map<u64, TableRef> g_inner_tables;
static std::map<std::pair<u64, u64>, TableRef> g_inner_matrices;
static std::map<std::pair<u64, u64>, TableRef> g_full_tables;
static std::map<u64, TableRef> g_coset_tables, g_scaled_coset_tables;
static TableRef g_rader257;

static int bit_length(u64 x) { return x ? 64 - __builtin_clzll(x) : 0; }

// Copy a host table to a new device allocation on the table stream (a
// plain hipMemcpy would queue behind the caller's work on the NULL stream).
static TableRef upload(const std::vector<u64> &h) {
    hipStream_t ts = table_stream();
    void *d = nullptr;
    if (!ts || hipMalloc(&d, h.size() * 8) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, ts) != hipSuccess ||
        hipStreamSynchronize(ts) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(d);
        return nullptr;
    }
    return adopt(d);
}

// Allocate `words`, run `gen(ptr, stream)` on the table stream and wait
// for it; nullptr on failure.
template <class F>
static TableRef generate(u64 words, F gen) {
    hipStream_t ts = table_stream();
    void *d = nullptr;
    if (!ts || hipMalloc(&d, words * 8) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    gen((u64 *)d, ts);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ts) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(d);
        return nullptr;
    }
    return adopt(d);
}

PowTable power_table(u64 base, u64 count, bool create) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(base, count);
    auto it = g_pow_tables.find(key);
    if (it != g_pow_tables.end()) return it->second;
    if (!create) return PowTable();
    u32 lo_bits = (u32)((bit_length(count > 1 ? count - 1 : 1) + 1) / 2);
    u64 L = 1ull << lo_bits;
    u64 H = (count + L - 1) / L;
    std::vector<u64> lo(L), hi(H ? H : 1);
    u64 w = 1;
    for (u64 e = 0; e < L; e++) {
        lo[e] = to_mont(w);
        w = mul(w, base);
    }
    u64 wl = w, v = 1;  // wl = base^L
    for (u64 j = 0; j < hi.size(); j++) {
        hi[j] = to_mont(v);
        v = mul(v, wl);
    }
    PowTable t;
    t.lo = upload(lo);
    t.hi = upload(hi);
    t.lo_bits = lo_bits;
    if (t.lo && t.hi) g_pow_tables[key] = t;
    return t;
}

// root(M) memoized: planning looks up the tables of every pass of every
// transform of a batch, and the exponentiation costs more than the lookup.
static u64 root_cached(u64 M) {
    static std::mutex mu;
    static std::unordered_map<u64, u64> memo;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = memo.find(M);
        if (it != memo.end()) return it->second;
    }
    const u64 r = root(M);
    std::lock_guard<std::mutex> lk(mu);
    memo.emplace(M, r);
    return r;
}

PowTable root_table(u64 M, bool create) { return power_table(root_cached(M), M, create); }

// Montgomery form of 1/n (the inverse transform's scale), memoized likewise.
static u64 inv_scale_cached(u64 n) {
    static std::mutex mu;
    static std::unordered_map<u64, u64> memo;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = memo.find(n);
        if (it != memo.end()) return it->second;
    }
    const u64 r = to_mont(inv(n % P));
    std::lock_guard<std::mutex> lk(mu);
    memo.emplace(n, r);
    return r;
}

TableRef twiddle_matrix(u64 M, u64 R, bool create) {
    {
        std::lock_guard<std::mutex> lk(g_tw_mu);
        auto it = g_full_tables.find(std::make_pair(M, R));
        if (it != g_full_tables.end()) return it->second;
        if (!create) return nullptr;
    }
    PowTable t = root_table(M);
    if (!t.lo || !t.hi) return nullptr;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(M, R);
    auto it = g_full_tables.find(key);
    if (it != g_full_tables.end()) return it->second;
    TableRef d = generate(M, [&](u64 *out, hipStream_t ts) {
        launch_twiddle_matrix(out, M, R, t.lo.get(), t.hi.get(), t.lo_bits, ts);
    });
    if (d) g_full_tables[key] = d;
    return d;
}

// A full twiddle matrix costs M words and one read per output; it replaces
// the two-level product (one multiply per output).  It pays off while the
// matrix stays cache resident (M up to 2^20 words = 8 MiB by default, set by
// GOLDILOCKS_TWIDDLE_FULL_MAX), or while adjacent instances share its words
// (see PassCursor::next); other passes use the two-level tables.
static std::atomic<u64> &full_table_max_ref() {
    static std::atomic<u64> v{[] {
        const char *e = getenv("GOLDILOCKS_TWIDDLE_FULL_MAX");
        return e ? (u64)strtoull(e, nullptr, 10) : (u64)1 << 20;
    }()};
    return v;
}
static u64 full_table_max() { return full_table_max_ref().load(); }
u64 set_full_table_max(u64 words) { return full_table_max_ref().exchange(words); }

TableRef inner_table(u64 R, bool create) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_inner_tables.find(R);
    if (it != g_inner_tables.end()) return it->second;
    if (!create) return nullptr;
    std::vector<u64> t(R);
    u64 w = root(R), v = 1;
    for (u64 e = 0; e < R; e++) {
        t[e] = to_mont(v);
        v = mul(v, w);
    }
    TableRef d = upload(t);
    if (d) g_inner_tables[R] = d;
    return d;
}

TableRef coset_table(u64 rows) {
    PowTable t = power_table(2717, rows);
    if (!t.lo || !t.hi) return nullptr;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_coset_tables.find(rows);
    if (it != g_coset_tables.end()) return it->second;
    TableRef d = generate(rows, [&](u64 *out, hipStream_t ts) {
        launch_power_series(out, rows, t.lo.get(), t.hi.get(), t.lo_bits, ts);
    });
    if (d) g_coset_tables[rows] = d;
    return d;
}

TableRef scaled_coset_table(u64 rows) {
    TableRef ct = coset_table(rows);
    if (!ct) return nullptr;
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_scaled_coset_tables.find(rows);
    if (it != g_scaled_coset_tables.end()) return it->second;
    TableRef d = generate(rows, [&](u64 *out, hipStream_t ts) {
        launch_scale_table(out, ct.get(), rows, to_mont(inv(rows % P)), ts);
    });
    if (d) g_scaled_coset_tables[rows] = d;
    return d;
}

TableRef inner_matrix(u64 R1, u64 R2, bool create) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto key = std::make_pair(R1, R2);
    auto it = g_inner_matrices.find(key);
    if (it != g_inner_matrices.end()) return it->second;
    if (!create) return nullptr;
    const u64 R = R1 * R2;
    std::vector<u64> pw(R), t(R);
    u64 w = root(R), v = 1;
    for (u64 e = 0; e < R; e++) {
        pw[e] = to_mont(v);
        v = mul(v, w);
    }
    for (u64 rb = 0; rb < R2; rb++)
        for (u64 ka = 0; ka < R1; ka++) t[rb * R1 + ka] = pw[(ka * rb) % R];
    TableRef d = upload(t);
    if (d) g_inner_matrices[key] = d;
    return d;
}

// Tables for the radix-257 Rader pass (rader.rs:19-57), Montgomery form:
// [0, 256) w_256^(k1*j) at entry j*16 + k1 (a thread's inner twiddles are
// consecutive), [256, 512) the convolution spectrum DFT_256(w_257^(3^q)/256),
// [512, 768) the conjugate inner twiddles of the second (inverse) DFT_256.
TableRef rader257_table(bool create) {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (g_rader257 || !create) return g_rader257;
    std::vector<u64> pw(256), t(768);
    u64 w256 = root(256), w257 = root(257), s256 = inv(256);
    u64 v = 1;
    for (int e = 0; e < 256; e++) {
        pw[e] = v;
        v = mul(v, w256);
    }
    for (int j = 0; j < 16; j++)
        for (int k1 = 0; k1 < 16; k1++) t[j * 16 + k1] = pw[(k1 * j) & 255];
    std::vector<u64> c(256);
    u64 gq = 1;
    for (int q = 0; q < 256; q++) {
        c[q] = mul(pow(w257, gq), s256);
        gq = gq * 3 % 257;
    }
    for (int k = 0; k < 256; k++) {
        u64 acc = 0;
        for (int q = 0; q < 256; q++) acc = add(acc, mul(c[q], pw[(q * k) & 255]));
        t[256 + k] = acc;
    }
    for (int j = 0; j < 16; j++)
        for (int q1 = 0; q1 < 16; q1++) t[512 + j * 16 + q1] = pw[(256 - ((q1 * j) & 255)) & 255];
    for (int e = 0; e < 768; e++) t[e] = to_mont(t[e]);
    g_rader257 = upload(t);
    return g_rader257;
}

// Radix-65537 Rader pass: no single kernel (launch == nullptr); run_plan
// composes gather, two batched 65536-point transforms, a pointwise product
// with the spectrum and a scatter.
static const PassVariant kRader65537 = {65537, 1, 1, 256, 0, nullptr, nullptr, VK_COMPOSITE, 1};
const PassVariant *rader65537_variant() { return &kRader65537; }

struct BigRaderTables {
    TableRef permk, permi, spec;  // permk/permi hold u32 entries
};
static BigRaderTables g_big_rader;

static int big_rader_tables(BigRaderTables &out) {
    {
        std::lock_guard<std::mutex> lk(g_tw_mu);
        if (g_big_rader.spec) {
            out = g_big_rader;
            return GL_OK;
        }
    }
    const u64 P2 = 65537, M2 = 65536;
    std::vector<u32> pk(M2), pi(M2);
    std::vector<u64> c(M2);
    u64 w = root(P2), sc = inv(M2);
    u64 a = 1, b = 1;
    for (u64 q = 0; q < M2; q++) {
        pk[q] = (u32)a;
        pi[q] = (u32)b;
        c[q] = to_mont(mul(pow(w, b), sc));
        a = a * 21846 % P2;
        b = b * 3 % P2;
    }
    std::vector<u64> pk64(M2 / 2), pi64(M2 / 2);
    std::memcpy(pk64.data(), pk.data(), M2 * 4);
    std::memcpy(pi64.data(), pi.data(), M2 * 4);
    BigRaderTables t;
    t.permk = upload(pk64);
    t.permi = upload(pi64);
    t.spec = upload(c);
    if (!t.permk || !t.permi || !t.spec) return set_error(GL_E_NOMEM, "radix-65537 table allocation failed");
    // spectrum = DFT_65536(c) (linear, so Montgomery form is preserved),
    // computed on the table stream
    PlanRef p65536;
    int rc = get_plan(M2, &p65536);
    if (rc) return rc;
    hipStream_t ts = table_stream();
    if (!ts) return set_error(GL_E_DEVICE, "table stream creation failed");
    {
        DevBuf tmp(M2 * 8, ts);
        if (!tmp.p) return set_error(GL_E_NOMEM, "radix-65537 table allocation failed");
        if ((rc = run_plan(*p65536, (u64 *)t.spec.get(), tmp.as<u64>(), 1, 1, M2, false, ts))) return rc;
    }
    GL_HIP(hipStreamSynchronize(ts));
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (!g_big_rader.spec) g_big_rader = t;
    out = g_big_rader;
    return GL_OK;
}

static int run_rader_big(const PassArgs &a, hipStream_t stream, Keep &keep) {
    BigRaderTables t;
    int rc = big_rader_tables(t);
    if (rc) return rc;
    keep.push_back(t.permk);
    keep.push_back(t.permi);
    keep.push_back(t.spec);
    PlanRef p65536;
    if ((rc = get_plan(65536, &p65536))) return rc;
    const u64 ni = a.total_inst, words = ni * 65536;
    DevBuf C(words * 8, stream), Ct(words * 8, stream), x0(ni * 8, stream);
    if (!C.p || !Ct.p || !x0.p) return set_error(GL_E_NOMEM, "radix-65537 scratch allocation failed");
    launch_rader_big_gather(a, (const u32 *)t.permk.get(), C.as<u64>(), x0.as<u64>(), stream);
    GL_HIP(hipGetLastError());
    if ((rc = run_plan(*p65536, C.as<u64>(), Ct.as<u64>(), 1, ni, 65536, false, stream))) return rc;
    launch_rader_big_pointwise(C.as<u64>(), x0.as<u64>(), t.spec.get(), ni, stream);
    GL_HIP(hipGetLastError());
    // reverse C[1..] and transform again (unscaled: the spectrum holds 1/65536)
    if ((rc = run_plan(*p65536, C.as<u64>(), Ct.as<u64>(), 1, ni, 65536, true, stream, false))) return rc;
    launch_rader_big_scatter(a, (const u32 *)t.permi.get(), C.as<u64>(), x0.as<u64>(), stream);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

bool plan_is_composite(const Plan &plan) {
    for (const PassPlan &p : plan.passes)
        if (!p.v->launch) return true;
    return false;
}

// ---------------------------------------------------------------- planner
// Built plans (a cache), and the schedules that decide them: explicit ones
// (set_plan) first, then the tune file's, then the planner's default.
// Index 0: contiguous transforms, 1: column transforms (S0 > 1).
static std::mutex g_plan_mu;
static std::map<u64, PlanRef> g_plans[2];
static std::map<u64, std::vector<u32>> g_explicit[2], g_tuned[2];

static const PassVariant *find_variant(int R1, int R2) {
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    for (int i = 0; i < cnt; i++)
        if (vs[i].R1 == R1 && vs[i].R2 == R2) return &vs[i];
    return nullptr;
}

// Pass with `bits` power-of-two bits and odd codelet g (g = 1: none).
// Column transforms (cols) use two-sub-step variants only: their instances
// are adjacent columns, and those variants load 16-32 of them per row
// segment where the three-stage ones load 2-8.
static const PassVariant *variant_for(int bits, int g, bool cols = false) {
    if (g == 1 && cols && bits > 8) return nullptr;
    if (g == 1) {
        // Power-of-two passes: two register stages up to 2^8, three stages
        // (16-register codelets, four waves per SIMD) above; chosen from the
        // autotuner's sweeps on MI355X.
        static const int pref[13][2] = {{0, 0},  {2, 1},  {2, 2},  {4, 2},   {4, 4},   {8, 4},  {8, 8},
                                        {16, 8}, {16, 16}, {8, 64}, {16, 64}, {16, 128}, {16, 256}};
        if (bits < 1 || bits > 12) return nullptr;
        return find_variant(pref[bits][0], pref[bits][1]);
    }
    if (bits == 0) return find_variant(g, 1);
    if (bits > 6) return nullptr;
    return find_variant(1 << bits, g);
}

static bool build_plan(u64 n, Plan &plan, bool cols = false) {
    plan.n = n;
    plan.passes.clear();
    if (n == 1) return true;
    int k = __builtin_ctzll(n);
    u64 f = n >> k;
    const bool rader257 = f % 257 == 0;
    if (rader257) f /= 257;
    const bool rader65537 = f % 65537 == 0;
    if (rader65537) f /= 65537;
    std::vector<int> primes;
    for (u64 q : {3ull, 5ull, 17ull})
        if (f % q == 0) {
            primes.push_back((int)q);
            f /= q;
        }
    if (f != 1) return false;
    if (rader65537) plan.passes.push_back({&kRader65537, 65537});
    if (rader257) plan.passes.push_back({rader257_variant(), 257});
    if (k == 0 && primes.empty()) return true;
    // Candidate groupings of the odd primes into single-codelet groups.
    std::vector<std::vector<int>> groupings;
    bool has3 = std::count(primes.begin(), primes.end(), 3) > 0;
    bool has5 = std::count(primes.begin(), primes.end(), 5) > 0;
    bool has17 = std::count(primes.begin(), primes.end(), 17) > 0;
    {
        std::vector<int> g;
        if (has3) g.push_back(3);
        if (has5) g.push_back(5);
        if (has17) g.push_back(17);
        groupings.push_back(g);
        if (has3 && has5) {
            std::vector<int> h{15};
            if (has17) h.push_back(17);
            groupings.push_back(h);
        }
    }
    double best_score = 1e30;
    std::vector<std::pair<int, int>> best;  // (bits, g)
    // Passes of at most ~2^8 points (two register stages) run most
    // efficiently; start from that pass count, at most three passes, and
    // add passes only when no split fits the compiled variants.
    double lg = (double)k;
    for (int q : primes) lg += std::log2((double)q);
    const int p_start = std::max(1, std::min(3, (int)std::ceil(lg / 8.0)));
    for (int P = p_start; P <= 5 && best.empty(); P++) {
        for (auto &grp : groupings) {
            if ((int)grp.size() > P) continue;
            std::vector<int> g(P, 1), bits(P, 0), cap(P, 12), lo(P, 1);
            for (size_t i = 0; i < grp.size(); i++) {
                g[i] = grp[i];
                cap[i] = 6;
                lo[i] = 0;
            }
            int need = k;
            for (int i = 0; i < P; i++) {
                bits[i] = lo[i];
                need -= lo[i];
            }
            int total_cap = 0;
            for (int i = 0; i < P; i++) total_cap += cap[i] - lo[i];
            if (need < 0 || need > total_cap) continue;
            while (need > 0) {
                int bi = -1;
                double bs = 1e30;
                for (int i = 0; i < P; i++) {
                    if (bits[i] >= cap[i]) continue;
                    double sz = bits[i] + std::log2((double)g[i]);
                    if (sz < bs) {
                        bs = sz;
                        bi = i;
                    }
                }
                bits[bi]++;
                need--;
            }
            bool ok = true;
            double mx = 0;
            for (int i = 0; i < P; i++) {
                if (!variant_for(bits[i], g[i], cols)) ok = false;
                mx = std::max(mx, bits[i] + std::log2((double)g[i]));
            }
            if (!ok) continue;
            if (mx < best_score) {
                best_score = mx;
                best.clear();
                for (int i = 0; i < P; i++) best.emplace_back(bits[i], g[i]);
            }
        }
    }
    if (best.empty()) return false;
    for (auto &bg : best) {
        const PassVariant *v = variant_for(bg.first, bg.second, cols);
        plan.passes.push_back({v, (u64)v->R1 * (u64)v->R2});
    }
    return true;
}

// Pass list from variant ids (index == compiled variant count: the
// radix-257 Rader pass); false when an id is unknown or the radices do not
// multiply to n.
static bool plan_from_ids(u64 n, const std::vector<u32> &ids, Plan &plan) {
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    plan.n = n;
    plan.passes.clear();
    u64 prod = 1;
    for (u32 id : ids) {
        const PassVariant *v = nullptr;
        if (id < (u32)cnt) v = &vs[id];
        else if (id == (u32)cnt) v = rader257_variant();
        if (!v) return false;
        u64 R = (u64)v->R1 * (u64)v->R2;
        plan.passes.push_back({v, R});
        prod *= R;
    }
    return prod == n;
}

static int get_plan_impl(u64 n, PlanRef *out, int cols) {
    if (n == 0 || (P - 1) % n != 0)
        return set_error(GL_E_SIZE, std::to_string(n) +
                                        " is not a supported NTT size (does not divide the multiplicative order)");
    std::call_once(g_tune_once, load_tune_file);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans[cols].find(n);
    if (it == g_plans[cols].end()) {
        std::shared_ptr<Plan> p(new Plan());
        auto ex = g_explicit[cols].find(n);
        auto tu = g_tuned[cols].find(n);
        bool ok;
        if (ex != g_explicit[cols].end()) ok = plan_from_ids(n, ex->second, *p);
        else if (tu != g_tuned[cols].end()) ok = plan_from_ids(n, tu->second, *p);
        else ok = build_plan(n, *p, cols != 0);
        if (!ok) return set_error(GL_E_UNSUPPORTED, "no device schedule for NTT size " + std::to_string(n));
        it = g_plans[cols].emplace(n, std::move(p)).first;
    }
    if (out) *out = it->second;
    return GL_OK;
}

int get_plan(u64 n, PlanRef *out) { return get_plan_impl(n, out, 0); }
int get_plan_cols(u64 n, PlanRef *out) { return get_plan_impl(n, out, 1); }

int set_plan(u64 n, const u32 *ids, size_t nids, bool cols) {
    if (n == 0 || (P - 1) % n != 0) return set_error(GL_E_SIZE, "plan_set: size does not divide p - 1");
    std::call_once(g_tune_once, load_tune_file);
    std::vector<u32> v(ids, ids + nids);
    if (nids) {
        Plan check;
        if (!plan_from_ids(n, v, check)) return set_error(GL_E_ARG, "plan_set: unknown variant id or radices do not multiply to n");
    }
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (nids) g_explicit[cols][n] = v;
    else g_explicit[cols].erase(n);
    g_plans[cols].erase(n);  // holders keep their PlanRef
    return GL_OK;
}

// Autotuned schedules: text lines "n R1xR2/T R1xR2/T ..." (variant radices
// and instances per workgroup, as gl_variants lists them); "cn ..." lines
// hold the schedule of column transforms of length n (matrices of several
// columns).  Read from GOLDILOCKS_TUNE_FILE, else tune_gfx950.txt next to
// this library.
static void load_tune_file() {
    std::string path;
    if (const char *e = getenv("GOLDILOCKS_TUNE_FILE")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr((void *)&load_tune_file, &info) && info.dli_fname) {
            path = info.dli_fname;
            size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tune_gfx950.txt";
        }
    }
    FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (!f) return;
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    char line[1024];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        char *save = nullptr;
        char *tok = strtok_r(line, " \t\r\n", &save);
        if (!tok) continue;
        const bool cols = tok[0] == 'c';
        u64 n = strtoull(tok + (cols ? 1 : 0), nullptr, 10);
        std::vector<u32> ids;
        bool ok = true;
        while ((tok = strtok_r(nullptr, " \t\r\n", &save))) {
            int r1 = 0, r2 = 0, t = 0;
            if (sscanf(tok, "%dx%d/%d", &r1, &r2, &t) != 3) {
                ok = false;
                break;
            }
            int id = -1;
            for (int i = 0; i < cnt; i++)
                if (vs[i].R1 == r1 && vs[i].R2 == r2 && vs[i].T == t) id = i;
            if (r1 == 257 && r2 == 1) id = cnt;
            if (id < 0) {
                ok = false;
                break;
            }
            ids.push_back((u32)id);
        }
        Plan check;
        if (ok && !ids.empty() && n && (P - 1) % n == 0 && plan_from_ids(n, ids, check)) {
            std::lock_guard<std::mutex> lk(g_plan_mu);
            g_tuned[cols ? 1 : 0][n] = ids;
            g_plans[cols ? 1 : 0].erase(n);
        }
    }
    fclose(f);
}

size_t plan_tmp_words(const Plan &plan, u64 count, u64 bstride) {
    return plan.passes.size() > 1 ? (size_t)(count * bstride) : 0;
}

int PassCursor::next(PassArgs &a, unsigned &grid) {
    const size_t np = plan->passes.size();
    const u64 n = plan->n;
    const PassPlan &pp = plan->passes[p];
    const u64 R = pp.R, Mp = M / R;
    const bool last = p + 1 == np;
    u64 *dst = last ? data : (src == data ? tmp : data);
    std::memset(&a, 0, sizeof a);
    a.in = src;
    a.out = dst;
    a.bstride = bstride;
    a.S = S;
    a.SMp = S * Mp;
    a.n_rev = n;
    u64 ninst = S * Mp, total = count * ninst;
    // kernels index instances with 31-bit multiply-shift division
    if (total >= (1ull << 31)) return set_error(GL_E_UNSUPPORTED, "batch too large for one launch");
    a.ninst = (u32)ninst;
    a.total_inst = (u32)total;
    a.Mp = (u32)Mp;
    fastdiv_params(ninst, a.ninst_mul, a.ninst_sh);
    fastdiv_params(S, a.S_mul, a.S_sh);
    fastdiv_params(S0, a.S0_mul, a.S0_sh);
    a.SS = S / S0;
    const bool mk = create_tables;
    static const char *kNotResident =
        "twiddle tables for this length are not resident and this call allocates nothing: "
        "run the allocating entry point on the same length once first";
    TableRef inner;
    if (pp.v->kind == VK_RADER257) {
        inner = rader257_table(mk);
    } else if (pp.v->kind == VK_THREE_STAGE) {
        inner = inner_table(R, mk);
    } else if (pp.v->kind == VK_TWO_STAGE && pp.v->R2 > 1) {
        inner = inner_matrix((u64)pp.v->R1, (u64)pp.v->R2, mk);
    }
    if (inner) keep.push_back(inner);
    a.tw_inner = inner.get();
    const bool needs_inner = pp.v->kind == VK_RADER257 || pp.v->kind == VK_THREE_STAGE ||
                             (pp.v->kind == VK_TWO_STAGE && pp.v->R2 > 1);
    if (needs_inner && !a.tw_inner)
        return mk ? set_error(GL_E_NOMEM, "twiddle table allocation failed") : set_error(GL_E_UNSUPPORTED, kNotResident);
    if (Mp > 1) {
        // Past the cache-resident size a full matrix still pays off when S >= 32:
        // the S adjacent instances s + S*m share each of its words, so every
        // workgroup fetches one run of R words for T*R data words.
        const u64 fmax = full_table_max();
        const bool use_full = M <= fmax || (S >= 32 && M <= 4 * fmax);
        TableRef full = use_full ? twiddle_matrix(M, R, mk) : nullptr;
        if (use_full && !full && !mk) return set_error(GL_E_UNSUPPORTED, kNotResident);
        if (full) {
            keep.push_back(full);
            a.tw_lo = full.get();
            a.flags |= FLAG_TWIDDLE_FULL;
        } else {
            PowTable t = root_table(M, mk);
            if (!t.lo || !t.hi)
                return mk ? set_error(GL_E_NOMEM, "twiddle table allocation failed")
                          : set_error(GL_E_UNSUPPORTED, kNotResident);
            keep.push_back(t.lo);
            keep.push_back(t.hi);
            a.tw_lo = t.lo.get();
            a.tw_hi = t.hi.get();
            a.lo_bits = t.lo_bits;
            a.flags |= FLAG_TWIDDLE;
            // inverse: 1/n multiplies the first output twiddle of each thread's
            // recurrence (one product per thread) instead of every output of
            // the last pass
            if (inverse && p == 0 && scale_out && !row_scale && np > 1 &&
                (pp.v->kind == VK_TWO_STAGE || pp.v->kind == VK_THREE_STAGE)) {
                a.flags |= FLAG_SCALE_TW;
                a.scale = inv_scale_cached(n);
                scale_folded = true;
            }
        }
    }
    if (inverse && p == 0) a.flags |= FLAG_REVERSE;
    if ((coset || coset_rows) && p == 0) {
        // coset shift: two-sub-step kernels; zero extension alone: also three-stage ones
        if (pp.v->kind != VK_TWO_STAGE && (coset || pp.v->kind != VK_THREE_STAGE))
            return set_error(GL_E_ARG, "coset input needs a two-sub-step first pass");
        a.flags |= coset ? FLAG_COSET : FLAG_ZEXT;
        if (!coset && 2 * coset_rows == M) a.flags |= FLAG_ZEXT_HALF;
        a.coset = coset;
        a.coset_rows = coset_rows;
    }
    if (last && row_scale) {
        if (pp.v->kind != VK_TWO_STAGE && pp.v->kind != VK_THREE_STAGE)
            return set_error(GL_E_ARG, "row scale needs a single-kernel last pass");
        a.flags |= FLAG_ROW_SCALE;
        a.row_scale = row_scale;
    } else if (inverse && last && scale_out && !scale_folded) {
        a.flags |= FLAG_SCALE;
        a.scale = inv_scale_cached(n);
    }
    grid = (unsigned)((total + pp.v->T - 1) / pp.v->T);
    src = dst;
    S *= R;
    M = Mp;
    p++;
    return GL_OK;
}

// ---------------------------------------------------------------- pass timing
// While enabled, every pass launch is bracketed by HIP events on its stream;
// timing_collect() waits for them and returns (variant id, milliseconds).
namespace {
struct TimingRec {
    u32 variant;
    hipEvent_t e0, e1;
};
std::mutex g_tm_mu;
std::atomic<bool> g_tm_on{false};
std::vector<TimingRec> g_tm_recs;
std::vector<hipEvent_t> g_tm_pool;

hipEvent_t tm_event() {
    std::lock_guard<std::mutex> lk(g_tm_mu);
    if (!g_tm_pool.empty()) {
        hipEvent_t e = g_tm_pool.back();
        g_tm_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}
}  // namespace

u32 variant_id(const PassVariant *v) {
    int cnt = 0;
    const PassVariant *vs = pass_variants(&cnt);
    if (v >= vs && v < vs + cnt) return (u32)(v - vs);
    if (v == rader257_variant()) return (u32)cnt;
    return (u32)cnt + 1;
}

void timing_enable(bool on) { g_tm_on.store(on); }

// Brackets one launch: call with e0 = timing_begin(s) before and
// timing_end(v, e0, s) after.
// Records kept until collected; beyond this many, launches go untimed.
constexpr size_t kMaxTimingRecords = size_t(1) << 20;

hipEvent_t timing_begin(hipStream_t s) {
    if (!g_tm_on.load()) return nullptr;
    {
        std::lock_guard<std::mutex> lk(g_tm_mu);
        if (g_tm_recs.size() >= kMaxTimingRecords) return nullptr;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (cs != hipStreamCaptureStatusNone) return nullptr;  // no timing inside a graph capture
    hipEvent_t e = tm_event();
    if (e && hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}
void timing_end(const PassVariant *v, hipEvent_t e0, hipStream_t s) {
    if (!e0) return;
    hipEvent_t e1 = tm_event();
    if (!e1 || hipEventRecord(e1, s) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    std::lock_guard<std::mutex> lk(g_tm_mu);
    g_tm_recs.push_back({variant_id(v), e0, e1});
}

size_t timing_pending() {
    std::lock_guard<std::mutex> lk(g_tm_mu);
    return g_tm_recs.size();
}

int timing_collect(uint32_t *ids, float *ms, size_t cap, size_t *count) {
    std::vector<TimingRec> recs;
    {
        std::lock_guard<std::mutex> lk(g_tm_mu);
        recs.swap(g_tm_recs);
    }
    if (count) *count = recs.size();
    int rc = GL_OK;
    for (size_t i = 0; i < recs.size(); i++) {
        float t = 0.f;
        hipError_t e = hipEventSynchronize(recs[i].e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, recs[i].e0, recs[i].e1);
        if (e != hipSuccess && !rc) rc = hip_error(e, "pass timing");
        if (i < cap) {
            if (ids) ids[i] = recs[i].variant;
            if (ms) ms[i] = t;
        }
    }
    std::lock_guard<std::mutex> lk(g_tm_mu);
    for (const TimingRec &r : recs) {
        g_tm_pool.push_back(r.e0);
        g_tm_pool.push_back(r.e1);
    }
    return rc;
}

int run_plan(const Plan &plan, u64 *data, u64 *tmp, u64 S0, u64 count, u64 bstride, bool inverse,
             hipStream_t stream, bool scale_out, const u64 *input, const u64 *coset, u64 coset_rows,
             const u64 *row_scale) {
    if (plan.passes.empty() || count == 0) return GL_OK;
    PassCursor c{&plan, data, tmp, S0, count, bstride, inverse};
    c.scale_out = scale_out;
    c.input = input;
    c.coset = coset;
    c.coset_rows = coset_rows;
    c.row_scale = row_scale;
    c.begin();
    while (!c.done()) {
        const PassVariant *v = plan.passes[c.p].v;
        PassArgs a;
        unsigned grid = 0;
        int rc = c.next(a, grid);
        if (rc) return rc;
        hipEvent_t e0 = timing_begin(stream);
        if (!v->launch) {
            if ((rc = run_rader_big(a, stream, c.keep))) return rc;
        } else {
            v->launch(a, grid, stream);
            GL_HIP(hipGetLastError());
        }
        timing_end(v, e0, stream);
    }
    return GL_OK;  // c.keep drops here: the passes are enqueued
}

int warm_plan(const Plan &plan, u64 S0, Keep *keep) {
    int rc = GL_OK;
    for (int inverse = 0; inverse < 2 && !rc; inverse++) {
        // null data/tmp: next() only fills arguments, nothing is launched
        PassCursor c{&plan, nullptr, nullptr, S0, 1, plan.n * S0, inverse != 0};
        c.begin();
        while (!c.done() && !rc) {
            PassArgs a;
            unsigned grid = 0;
            rc = c.next(a, grid);
        }
        if (!rc) {
            for (const PassPlan &pp : plan.passes) {
                if (pp.v->launch) continue;
                BigRaderTables t;
                if ((rc = big_rader_tables(t))) break;
                c.keep.push_back(t.permk);
                c.keep.push_back(t.permi);
                c.keep.push_back(t.spec);
            }
        }
        if (keep) keep->insert(keep->end(), c.keep.begin(), c.keep.end());
    }
    return rc;
}

size_t plan_extra_words(const Plan &plan, u64 count, u64 bstride) {
    // the radix-65537 pass takes C and its ping-pong copy (65536 words per
    // instance each) and one word per instance
    for (const PassPlan &pp : plan.passes)
        if (!pp.v->launch) {
            const u64 inst = count * bstride / 65537;
            return (size_t)(2 * inst * 65536 + inst);
        }
    return 0;
}

void clear_caches() {
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_plans[0].clear();
        g_plans[1].clear();
    }
    {
        std::lock_guard<std::mutex> lk(g_tw_mu);
        g_pow_tables.clear();
        g_inner_tables.clear();
        g_inner_matrices.clear();
        g_coset_tables.clear();
        g_scaled_coset_tables.clear();
        g_full_tables.clear();
        g_rader257.reset();
        g_big_rader = BigRaderTables();
    }
    // Tables retired up to now: their last user enqueued its work before
    // the retirement, so once the device has drained that work is done.
    // Tables still referenced (by a caller mid-enqueue or a captured batch
    // plan) are retired later and freed by a later call.
    std::vector<void *> dead;
    {
        std::lock_guard<std::mutex> lk(g_grave_mu);
        dead.swap(g_grave);
    }
    if (hipDeviceSynchronize() != hipSuccess) (void)hipGetLastError();
    for (void *p : dead) (void)hipFree(p);
    pool_release_idle();
}

}  // namespace gl
