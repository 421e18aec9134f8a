// Mixed-size batches: many transforms of different lengths and radices.
//
// Batched path: level-synchronous grouped launches.  For pass index L, all
// transforms whose plan has an L-th pass with the same kernel variant share
// one launch whose blocks are split across per-transform descriptors, so the
// launch count depends on the distinct (level, variant) pairs, not on the
// number of transforms.  gl_batch_plan_* keeps the descriptors resident and
// replays the launch sequence as a captured hipGraph.  The unbatched baseline
// issues one launch sequence per transform.
#include <algorithm>
#include <cstring>
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
    std::vector<const Plan *> plans;
    std::vector<size_t> composite;
    std::vector<u64> tmp_off;
    u64 tmp_words = 0;
    std::vector<Group> groups;
    std::vector<PassArgs> descs;
    std::vector<u32> starts;
    size_t desc_bytes() const { return descs.size() * sizeof(PassArgs); }
    size_t meta_bytes() const { return desc_bytes() + starts.size() * sizeof(u32); }
};

int plan_batch(uint64_t *const *d, const size_t *n, size_t k, bool inverse, Schedule &sc) {
    sc.plans.assign(k, nullptr);
    sc.tmp_off.assign(k, 0);
    sc.tmp_words = 0;
    int rc;
    for (size_t i = 0; i < k; i++) {
        if (inverse && n[i] <= 1) continue;
        if ((rc = get_plan(n[i], &sc.plans[i]))) return rc;
        if (plan_is_composite(*sc.plans[i])) {
            sc.composite.push_back(i);  // not groupable: run on its own
            sc.plans[i] = nullptr;
            continue;
        }
        if (!d[i] && !sc.plans[i]->passes.empty()) return set_error(GL_E_ARG, "null device pointer");
        if (sc.plans[i]->passes.size() > 1) {
            sc.tmp_off[i] = sc.tmp_words;
            sc.tmp_words += (n[i] + 31) / 32 * 32;
        }
    }
    return GL_OK;
}

int build_groups(uint64_t *const *d, const size_t *n, size_t k, bool inverse, u64 *tmp, Schedule &sc) {
    size_t levels = 0;
    std::vector<PassCursor> cur;
    cur.reserve(k);
    for (size_t i = 0; i < k; i++) {
        PassCursor c{sc.plans[i], d[i], tmp ? tmp + sc.tmp_off[i] : nullptr, 1, 1, n[i], inverse};
        if (sc.plans[i]) {
            c.begin();
            levels = std::max(levels, sc.plans[i]->passes.size());
        }
        cur.push_back(c);
    }
    int rc;
    for (size_t L = 0; L < levels; L++) {
        std::vector<std::pair<const PassVariant *, size_t>> items;  // (variant, transform)
        for (size_t i = 0; i < k; i++)
            if (sc.plans[i] && L < sc.plans[i]->passes.size()) items.emplace_back(sc.plans[i]->passes[L].v, i);
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
    }
    return GL_OK;
}

void pack_meta(const Schedule &sc, std::vector<unsigned char> &host) {
    host.resize(sc.meta_bytes());
    std::memcpy(host.data(), sc.descs.data(), sc.desc_bytes());
    std::memcpy(host.data() + sc.desc_bytes(), sc.starts.data(), sc.starts.size() * sizeof(u32));
}

int launch_groups(const Schedule &sc, const void *meta, hipStream_t s) {
    const PassArgs *ddesc = (const PassArgs *)meta;
    const u32 *dstarts = (const u32 *)((const unsigned char *)meta + sc.desc_bytes());
    for (const Group &g : sc.groups) {
        if (g.blocks == 0) continue;
        g.v->launch_grouped(ddesc + g.desc_off, dstarts + g.starts_off, (u32)g.nd, g.blocks, s);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

}  // namespace

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
    Schedule sc;
    if ((rc = plan_batch(d, n, k, inverse != 0, sc))) return rc;
    for (size_t i : sc.composite)
        if ((rc = gl_ntt_dev(d[i], n[i], 1, inverse, stream))) return rc;
    u64 *tmp = stream_scratch(s, sc.tmp_words, 0);
    if (sc.tmp_words && !tmp) return set_error(GL_E_NOMEM, "scratch allocation failed");
    if ((rc = build_groups(d, n, k, inverse != 0, tmp, sc))) return rc;
    if (sc.groups.empty()) return GL_OK;
    u64 *meta = stream_scratch(s, sc.meta_bytes() / 8 + 64, 1);
    if (!meta) return set_error(GL_E_NOMEM, "descriptor allocation failed");
    std::vector<unsigned char> host;
    pack_meta(sc, host);
    GL_HIP(hipMemcpyAsync(meta, host.data(), host.size(), hipMemcpyHostToDevice, s));
    if ((rc = launch_groups(sc, meta, s))) return rc;
    // `host` is released on return: wait until the copy has consumed it.
    GL_HIP(hipStreamSynchronize(s));
    return GL_OK;
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
    DevBuf d(total * 8 + 8);
    if (!d.p) return set_error(GL_E_NOMEM, "device allocation failed");
    std::vector<uint64_t *> ptrs(k);
    for (size_t i = 0; i < k; i++) {
        ptrs[i] = d.as<u64>() + off[i];
        if (n[i]) GL_HIP(hipMemcpyAsync(ptrs[i], v[i], n[i] * 8, hipMemcpyHostToDevice, s));
    }
    if ((rc = gl_ntt_mixed_dev(ptrs.data(), n, k, inverse, s))) return rc;
    for (size_t i = 0; i < k; i++)
        if (n[i]) GL_HIP(hipMemcpyAsync(v[i], ptrs[i], n[i] * 8, hipMemcpyDeviceToHost, s));
    GL_HIP(hipStreamSynchronize(s));
    return GL_OK;
}

// ---------------------------------------------------------------- captured plans

struct gl_batch_plan {
    void *meta = nullptr;  // descriptors + block starts (device, owned)
    void *tmp = nullptr;   // ping-pong scratch (device, owned)
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
    if (!sc.composite.empty()) {
        gl_batch_plan_destroy(p);
        return set_error(GL_E_UNSUPPORTED, "batch plans cannot capture lengths with a 65537 factor");
    }
    if (sc.tmp_words && hipMalloc(&p->tmp, sc.tmp_words * 8) != hipSuccess) {
        (void)hipGetLastError();
        gl_batch_plan_destroy(p);
        return set_error(GL_E_NOMEM, "batch plan scratch allocation failed");
    }
    if ((rc = build_groups(d, n, k, inverse != 0, (u64 *)p->tmp, sc))) {
        gl_batch_plan_destroy(p);
        return rc;
    }
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
    rc = launch_groups(sc, p->meta, cs);
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
