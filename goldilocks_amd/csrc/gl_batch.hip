// Mixed-size batches: many transforms of different lengths and radices.
//
// Batched path: level-synchronous grouped launches over cache-sized chunks.
// Transforms are packed in order into chunks of at most chunk_words() words;
// within a chunk, for pass index L, all transforms whose plan has an L-th
// pass with the same kernel variant share one launch whose blocks are split
// across per-transform descriptors, so the launch count depends on the
// distinct (level, variant) pairs, not on the number of transforms, while a
// chunk's words stay in the last-level cache from one pass to the next.
// Transforms larger than a chunk (which fill the device by themselves) and
// radix-65537 ones run on their own.  gl_batch_plan_* keeps the descriptors resident and
// replays the launch sequence as a captured hipGraph.  The unbatched baseline
// issues one launch sequence per transform.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gl_field.hpp"
#include "gl_runtime.hpp"

using namespace gl;

namespace {

struct Group {
    const PassVariant *v;
    size_t desc_off, nd, starts_off;
    u32 blocks;
};

struct Schedule {
    std::vector<PlanRef> plans;
    Keep keep;                                // tables the grouped launches read
    std::vector<size_t> alone;                // run as single transforms (gl_ntt_dev)
    std::vector<std::vector<size_t>> chunks;  // grouped level by level
    std::vector<u64> tmp_off;                 // scratch offset; restarts in every chunk
    u64 tmp_words = 0;                        // largest chunk's scratch
    std::vector<Group> groups;
    std::vector<PassArgs> descs;
    std::vector<u32> starts;
    size_t desc_bytes() const { return descs.size() * sizeof(PassArgs); }
    size_t meta_bytes() const { return desc_bytes() + starts.size() * sizeof(u32); }
};

// Transforms are grouped in chunks of at most this many bytes; a transform
// larger than a chunk runs by itself.  Bigger chunks mean fewer, fuller
// grouped launches, smaller ones keep a chunk's words in the last-level
// cache between its passes.  1 GiB measured best on MI355X for both the
// 2^4..2^12 and the 2^10..2^20 mixed batches (64 MiB, 256 MiB, 512 MiB and
// 4 GiB were slower).  GOLDILOCKS_BATCH_CHUNK_BYTES overrides.
std::atomic<u64> &chunk_bytes_ref() {
    static std::atomic<u64> b{[] {
        const char *e = getenv("GOLDILOCKS_BATCH_CHUNK_BYTES");
        return e ? (u64)strtoull(e, nullptr, 10) : (u64)1 << 30;
    }()};
    return b;
}
u64 chunk_words() {
    u64 w = chunk_bytes_ref().load() / 8;
    return w ? w : 1;
}

int plan_batch(uint64_t *const *d, const size_t *n, size_t k, bool inverse, Schedule &sc) {
    sc.plans.assign(k, PlanRef());
    sc.tmp_off.assign(k, 0);
    sc.tmp_words = 0;
    const u64 cw = chunk_words();
    u64 cur = 0;
    int rc;
    for (size_t i = 0; i < k; i++) {
        if (inverse && n[i] <= 1) continue;
        if ((rc = get_plan(n[i], &sc.plans[i]))) return rc;
        if (sc.plans[i]->passes.empty()) {
            sc.plans[i].reset();
            continue;
        }
        if (!d[i]) return set_error(GL_E_ARG, "null device pointer");
        if (plan_is_composite(*sc.plans[i]) || n[i] > cw) {
            sc.alone.push_back(i);  // not groupable, or big enough to fill the device alone
            continue;
        }
        const u64 w = (n[i] + 31) / 32 * 32;
        if (sc.chunks.empty() || cur + w > cw) {
            sc.chunks.emplace_back();
            cur = 0;
        }
        sc.chunks.back().push_back(i);
        if (sc.plans[i]->passes.size() > 1) {
            sc.tmp_off[i] = cur;
            cur += w;
            sc.tmp_words = std::max(sc.tmp_words, cur);
        }
    }
    return GL_OK;
}

// Called after the groups of each (chunk, pass level) have been appended:
// groups [g0, end), their descriptors [d0, end) and block starts [s0, end).
using LevelFn = std::function<int(size_t g0, size_t d0, size_t s0)>;

int build_groups(uint64_t *const *d, const size_t *n, bool inverse, u64 *tmp, Schedule &sc,
                 const LevelFn &on_level = nullptr) {
    int rc;
    for (const std::vector<size_t> &chunk : sc.chunks) {
        size_t levels = 0;
        std::vector<PassCursor> cur;
        cur.reserve(chunk.size());
        for (size_t i : chunk) {
            PassCursor c{sc.plans[i].get(), d[i], tmp ? tmp + sc.tmp_off[i] : nullptr, 1, 1, n[i], inverse};
            c.begin();
            levels = std::max(levels, sc.plans[i]->passes.size());
            cur.push_back(c);
        }
        for (size_t L = 0; L < levels; L++) {
            const size_t g0 = sc.groups.size(), d0 = sc.descs.size(), s0 = sc.starts.size();
            std::vector<std::pair<const PassVariant *, size_t>> items;  // (variant, position in chunk)
            for (size_t j = 0; j < chunk.size(); j++)
                if (L < sc.plans[chunk[j]]->passes.size()) items.emplace_back(sc.plans[chunk[j]]->passes[L].v, j);
            std::stable_sort(items.begin(), items.end(),
                             [](const std::pair<const PassVariant *, size_t> &x,
                                const std::pair<const PassVariant *, size_t> &y) { return x.first < y.first; });
            for (size_t a = 0; a < items.size();) {
                size_t b = a;
                while (b < items.size() && items[b].first == items[a].first) b++;
                Group g{items[a].first, sc.descs.size(), b - a, sc.starts.size(), 0};
                u32 blocks = 0;
                for (size_t j = a; j < b; j++) {
                    PassArgs pa;
                    unsigned grid = 0;
                    if ((rc = cur[items[j].second].next(pa, grid))) return rc;
                    sc.descs.push_back(pa);
                    sc.starts.push_back(blocks);
                    if ((u64)blocks + grid >= (1ull << 31))
                        return set_error(GL_E_UNSUPPORTED, "grouped launch too large");
                    blocks += grid;
                }
                sc.starts.push_back(blocks);
                g.blocks = blocks;
                sc.groups.push_back(g);
                a = b;
            }
            if (on_level && sc.groups.size() > g0 && (rc = on_level(g0, d0, s0))) return rc;
        }
        for (PassCursor &c : cur) sc.keep.insert(sc.keep.end(), c.keep.begin(), c.keep.end());
    }
    return GL_OK;
}

void pack_meta(const Schedule &sc, std::vector<unsigned char> &host) {
    host.resize(sc.meta_bytes());
    std::memcpy(host.data(), sc.descs.data(), sc.desc_bytes());
    std::memcpy(host.data() + sc.desc_bytes(), sc.starts.data(), sc.starts.size() * sizeof(u32));
}

// timed: bracket launches for gl_timing_collect (never inside a graph capture)
int launch_groups(const Schedule &sc, const void *meta, hipStream_t s, bool timed) {
    const PassArgs *ddesc = (const PassArgs *)meta;
    const u32 *dstarts = (const u32 *)((const unsigned char *)meta + sc.desc_bytes());
    for (const Group &g : sc.groups) {
        if (g.blocks == 0) continue;
        hipEvent_t e0 = timed ? timing_begin(s) : nullptr;
        g.v->launch_grouped(ddesc + g.desc_off, dstarts + g.starts_off, (u32)g.nd, g.blocks, s);
        GL_HIP(hipGetLastError());
        timing_end(g.v, e0, s);
    }
    return GL_OK;
}

}  // namespace

extern "C" uint64_t gl_set_batch_chunk_bytes(uint64_t bytes) { return chunk_bytes_ref().exchange(bytes); }

extern "C" int gl_ntt_mixed_dev_unbatched(uint64_t *const *d, const size_t *n, size_t k, int inverse,
                                          void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (k && (!d || !n)) return set_error(GL_E_ARG, "null pointer");
    for (size_t i = 0; i < k; i++)
        if ((rc = gl_ntt_dev(d[i], n[i], 1, inverse, stream))) return rc;
    return GL_OK;
}

extern "C" int gl_ntt_mixed_dev(uint64_t *const *d, const size_t *n, size_t k, int inverse, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (k == 0) return GL_OK;
    if (!d || !n) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    // the launch descriptors come from a pinned slab that is reused once the
    // uploads have run: a captured graph would keep reading it later
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GL_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return set_error(GL_E_UNSUPPORTED, "gl_ntt_mixed_dev cannot be captured; use gl_batch_plan_create");
    Schedule sc;
    if ((rc = plan_batch(d, n, k, inverse != 0, sc))) return rc;
    // stand-alone transforms first (each takes its own scratch)
    for (size_t i : sc.alone)
        if ((rc = gl_ntt_dev(d[i], n[i], 1, inverse, stream))) return rc;
    DevBuf tmp(sc.tmp_words * 8, s);
    if (sc.tmp_words && !tmp.p) return set_error(GL_E_NOMEM, "scratch allocation failed");
    // Each (chunk, pass level) is uploaded and launched as soon as its
    // descriptors are built, so the device runs one level while the host
    // plans the next.  Level blocks [PassArgs x nd][u32 block starts],
    // 256-byte aligned, go through a pinned slab: at most one descriptor per
    // pass of every grouped transform, and one block-start terminator per
    // group (at most one group per descriptor).
    size_t dmax = 0;
    for (const std::vector<size_t> &chunk : sc.chunks)
        for (size_t i : chunk) dmax += sc.plans[i]->passes.size();
    if (dmax == 0) return GL_OK;
    const size_t cap = dmax * (sizeof(PassArgs) + 256) + 2 * dmax * sizeof(u32);
    DevBuf meta(cap, s);
    if (!meta.p) return set_error(GL_E_NOMEM, "descriptor allocation failed");
    PinnedSlab pin;
    if ((rc = pinned_acquire(cap, pin))) return rc;
    size_t off = 0;
    auto upload_and_launch = [&](size_t g0, size_t d0, size_t s0) -> int {
        const size_t nd = sc.descs.size() - d0, ns = sc.starts.size() - s0;
        const size_t dbytes = nd * sizeof(PassArgs), bytes = dbytes + ns * sizeof(u32);
        if (off + bytes > cap) return set_error(GL_E_UNSUPPORTED, "batch descriptors exceed their buffer");
        char *h = (char *)pin.p + off, *dv = (char *)meta.p + off;
        std::memcpy(h, sc.descs.data() + d0, dbytes);
        std::memcpy(h + dbytes, sc.starts.data() + s0, ns * sizeof(u32));
        GL_HIP(hipMemcpyAsync(dv, h, bytes, hipMemcpyHostToDevice, s));
        const PassArgs *ddesc = (const PassArgs *)dv;
        const u32 *dstarts = (const u32 *)(dv + dbytes);
        for (size_t gi = g0; gi < sc.groups.size(); gi++) {
            const Group &g = sc.groups[gi];
            if (g.blocks == 0) continue;
            hipEvent_t e0 = timing_begin(s);
            g.v->launch_grouped(ddesc + (g.desc_off - d0), dstarts + (g.starts_off - s0), (u32)g.nd, g.blocks, s);
            GL_HIP(hipGetLastError());
            timing_end(g.v, e0, s);
        }
        off += (bytes + 255) / 256 * 256;
        return GL_OK;
    };
    rc = build_groups(d, n, inverse != 0, tmp.as<u64>(), sc, upload_and_launch);
    pinned_release(pin, s);  // reused only after the uploads enqueued so far
    return rc;
}

extern "C" int gl_ntt_batch_mixed(uint64_t *const *v, const size_t *n, size_t k, int inverse) {
    int rc = ensure_init();
    if (rc) return rc;
    if (k == 0) return GL_OK;
    if (!v || !n) return set_error(GL_E_ARG, "null pointer");
    u64 total = 0;
    std::vector<u64> off(k);
    for (size_t i = 0; i < k; i++) {
        if (!gl_is_divisor(n[i]) && !(inverse && n[i] <= 1))
            return set_error(GL_E_SIZE, std::to_string(n[i]) + " is not a supported NTT size");
        off[i] = total;
        total += (n[i] + 31) / 32 * 32;
    }
    hipStream_t s = thread_stream();
    DevBuf d(total * 8 + 8, s);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    std::vector<uint64_t *> ptrs(k);
    for (size_t i = 0; i < k; i++) {
        ptrs[i] = d.as<u64>() + off[i];
        if (n[i] && (rc = copy_h2d(ptrs[i], v[i], n[i] * 8, s))) return rc;
    }
    if ((rc = gl_ntt_mixed_dev(ptrs.data(), n, k, inverse, s))) return rc;
    for (size_t i = 0; i < k; i++)
        if (n[i] && (rc = copy_d2h(v[i], ptrs[i], n[i] * 8, s, false))) return rc;
    GL_HIP(hipStreamSynchronize(s));
    return GL_OK;
}

// ---------------------------------------------------------------- captured plans

struct gl_batch_plan {
    void *meta = nullptr;  // descriptors + block starts (device, owned)
    void *tmp = nullptr;   // ping-pong scratch (device, owned)
    Keep keep;             // twiddle tables the captured launches read
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" void gl_batch_plan_destroy(gl_batch_plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->meta) (void)hipFree(p->meta);
    if (p->tmp) (void)hipFree(p->tmp);
    delete p;
}

extern "C" int gl_batch_plan_create(uint64_t *const *d, const size_t *n, size_t k, int inverse,
                                    gl_batch_plan **out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out || (k && (!d || !n))) return set_error(GL_E_ARG, "null pointer");
    *out = nullptr;
    gl_batch_plan *p = new gl_batch_plan();
    Schedule sc;
    if ((rc = plan_batch(d, n, k, inverse != 0, sc))) {
        gl_batch_plan_destroy(p);
        return rc;
    }
    // stand-alone transforms are captured too, one after another, sharing the
    // plan's scratch with the grouped chunks
    u64 tmp_words = sc.tmp_words;
    for (size_t i : sc.alone) {
        if (plan_is_composite(*sc.plans[i])) {
            gl_batch_plan_destroy(p);
            return set_error(GL_E_UNSUPPORTED, "batch plans cannot capture lengths with a 65537 factor");
        }
        tmp_words = std::max<u64>(tmp_words, plan_tmp_words(*sc.plans[i], 1, n[i]));
    }
    if (tmp_words && hipMalloc(&p->tmp, tmp_words * 8) != hipSuccess) {
        (void)hipGetLastError();
        gl_batch_plan_destroy(p);
        return set_error(GL_E_NOMEM, "batch plan scratch allocation failed");
    }
    if ((rc = build_groups(d, n, inverse != 0, (u64 *)p->tmp, sc))) {
        gl_batch_plan_destroy(p);
        return rc;
    }
    // build the stand-alone transforms' twiddle tables now: table setup
    // synchronizes the device, which a capturing thread may not do
    for (size_t i : sc.alone) {
        if ((rc = warm_plan(*sc.plans[i], 1, &p->keep))) {
            gl_batch_plan_destroy(p);
            return rc;
        }
    }
    p->keep.insert(p->keep.end(), sc.keep.begin(), sc.keep.end());
    std::vector<unsigned char> host;
    pack_meta(sc, host);
    hipStream_t cs = nullptr;
    hipError_t e = hipMalloc(&p->meta, host.size() + 256);
    if (e == hipSuccess) e = hipMemcpy(p->meta, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = hip_error(e, "batch plan setup");
        if (cs) (void)hipStreamDestroy(cs);
        gl_batch_plan_destroy(p);
        return rc;
    }
    for (size_t i : sc.alone)
        if (!rc) rc = run_plan(*sc.plans[i], d[i], (u64 *)p->tmp, 1, 1, n[i], inverse != 0, cs);
    if (!rc) rc = launch_groups(sc, p->meta, cs, false);
    e = hipStreamEndCapture(cs, &p->graph);
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (!rc && e != hipSuccess) rc = hip_error(e, "batch plan capture");
    if (rc) {
        gl_batch_plan_destroy(p);
        return rc;
    }
    *out = p;
    return GL_OK;
}

extern "C" int gl_batch_plan_launch(gl_batch_plan *p, void *stream) {
    if (!p) return set_error(GL_E_ARG, "null plan");
    if (!p->exec) return GL_OK;
    GL_HIP(hipGraphLaunch(p->exec, (hipStream_t)stream));
    return GL_OK;
}
