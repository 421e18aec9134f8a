// Mixed-size batches: many transforms of different lengths and radices.
//
// Unbatched baseline: one launch sequence per transform.  The batched path
// is in gl_batch_grouped below: transforms that share a pass variant at the
// same pass index are issued as one grouped launch.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gl_field.hpp"
#include "gl_runtime.hpp"

using namespace gl;

extern "C" int gl_ntt_mixed_dev_unbatched(uint64_t *const *d, const size_t *n, size_t k, int inverse,
                                          void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (k && (!d || !n)) return set_error(GL_E_ARG, "null pointer");
    for (size_t i = 0; i < k; i++)
        if ((rc = gl_ntt_dev(d[i], n[i], 1, inverse, stream))) return rc;
    return GL_OK;
}

// Batched path: level-synchronous grouped launches.  For pass index L, all
// transforms whose plan has an L-th pass with the same kernel variant share
// one launch whose blocks are split across per-transform descriptors.  The
// number of launches depends on the distinct (level, variant) pairs, not on
// the number of transforms.
extern "C" int gl_ntt_mixed_dev(uint64_t *const *d, const size_t *n, size_t k, int inverse, void *stream) {
    int rc = ensure_init();
    if (rc) return rc;
    if (k == 0) return GL_OK;
    if (!d || !n) return set_error(GL_E_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    std::vector<const Plan *> plans(k, nullptr);
    std::vector<u64> tmp_off(k, 0);
    u64 tmp_words = 0;
    size_t levels = 0;
    for (size_t i = 0; i < k; i++) {
        if (inverse && n[i] <= 1) continue;
        if ((rc = get_plan(n[i], &plans[i]))) return rc;
        if (!d[i] && !plans[i]->passes.empty()) return set_error(GL_E_ARG, "null device pointer");
        if (plans[i]->passes.size() > 1) {
            tmp_off[i] = tmp_words;
            tmp_words += (n[i] + 31) / 32 * 32;
        }
        levels = std::max(levels, plans[i]->passes.size());
    }
    u64 *tmp = stream_scratch(s, tmp_words, 0);
    if (tmp_words && !tmp) return set_error(GL_E_NOMEM, "scratch allocation failed");
    std::vector<PassCursor> cur;
    cur.reserve(k);
    for (size_t i = 0; i < k; i++) {
        PassCursor c{plans[i], d[i], tmp ? tmp + tmp_off[i] : nullptr, 1, 1, n[i], inverse != 0};
        if (plans[i]) c.begin();
        cur.push_back(c);
    }
    // Build every level's groups on the host, then upload once.
    struct Group {
        const PassVariant *v;
        size_t desc_off, nd;
        u32 blocks;
    };
    std::vector<Group> groups;
    std::vector<size_t> level_end;
    std::vector<PassArgs> descs;
    std::vector<u32> starts;  // per group: nd + 1 entries
    std::vector<size_t> starts_off;
    for (size_t L = 0; L < levels; L++) {
        std::vector<std::pair<const PassVariant *, size_t>> items;  // (variant, transform)
        for (size_t i = 0; i < k; i++)
            if (plans[i] && L < plans[i]->passes.size()) items.emplace_back(plans[i]->passes[L].v, i);
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<const PassVariant *, size_t> &x,
                            const std::pair<const PassVariant *, size_t> &y) { return x.first < y.first; });
        for (size_t a = 0; a < items.size();) {
            size_t b = a;
            while (b < items.size() && items[b].first == items[a].first) b++;
            Group g{items[a].first, descs.size(), b - a, 0};
            starts_off.push_back(starts.size());
            u32 blocks = 0;
            for (size_t j = a; j < b; j++) {
                PassArgs pa;
                unsigned grid = 0;
                if ((rc = cur[items[j].second].next(pa, grid))) return rc;
                descs.push_back(pa);
                starts.push_back(blocks);
                if ((u64)blocks + grid >= (1ull << 31)) return set_error(GL_E_UNSUPPORTED, "grouped launch too large");
                blocks += grid;
            }
            starts.push_back(blocks);
            g.blocks = blocks;
            groups.push_back(g);
            a = b;
        }
        level_end.push_back(groups.size());
    }
    size_t dbytes = descs.size() * sizeof(PassArgs), sbytes = starts.size() * sizeof(u32);
    size_t words = (dbytes + sbytes) / 8 + 64;
    u64 *meta = stream_scratch(s, words, 1);
    if (!meta) return set_error(GL_E_NOMEM, "descriptor allocation failed");
    std::vector<unsigned char> host(dbytes + sbytes);
    std::memcpy(host.data(), descs.data(), dbytes);
    std::memcpy(host.data() + dbytes, starts.data(), sbytes);
    GL_HIP(hipMemcpyAsync(meta, host.data(), host.size(), hipMemcpyHostToDevice, s));
    const PassArgs *ddesc = (const PassArgs *)meta;
    const u32 *dstarts = (const u32 *)((unsigned char *)meta + dbytes);
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const Group &g = groups[gi];
        if (g.blocks == 0) continue;
        g.v->launch_grouped(ddesc + g.desc_off, dstarts + starts_off[gi], (u32)g.nd, g.blocks, s);
        GL_HIP(hipGetLastError());
    }
    // The descriptor staging vector is freed on return; make sure the copy
    // has consumed it (pageable source) before leaving.
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
