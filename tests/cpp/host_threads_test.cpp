// Multi-threaded stress of the host side of the C ABI, built with a host
// sanitizer (thread or address) by tests/test_host_sanitizers.py and run on
// the CPU: plan lookups, explicit schedules installed and removed, variant
// names, divisor queries and cache clears from many threads at once.  Device
// calls are made only when a gfx950 device is present; without one they
// must fail cleanly with GL_E_DEVICE.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "goldilocks_gpu.h"

static std::atomic<int> g_fail{0};

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            g_fail.fetch_add(1);                                      \
        }                                                             \
    } while (0)

static uint64_t rng(uint64_t &s) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

int main() {
    const int device_ok = gl_init(-1) == GL_OK;
    if (!device_ok) CHECK(gl_ntt(nullptr, 16) == GL_E_DEVICE);
    std::vector<uint64_t> sizes;
    for (uint64_t n = 2; n <= (1ull << 24); n *= 2) {
        sizes.push_back(n);
        sizes.push_back(n * 3);
        sizes.push_back(n * 5);
        sizes.push_back(n * 17);
        sizes.push_back(n * 257);
        sizes.push_back(n * 15);
    }
    const size_t nvar = gl_variants(nullptr, nullptr, nullptr, 0);
    std::vector<uint32_t> r1(nvar), r2(nvar), tt(nvar);
    gl_variants(r1.data(), r2.data(), tt.data(), nvar);
    uint32_t id16x16 = UINT32_MAX, id16x1 = UINT32_MAX;
    for (uint32_t i = 0; i < nvar; i++) {
        if (r1[i] == 16 && r2[i] == 16 && tt[i] == 16) id16x16 = i;
        if (r1[i] == 16 && r2[i] == 1) id16x1 = i;
    }
    CHECK(id16x16 != UINT32_MAX && id16x1 != UINT32_MAX);

    auto worker = [&](int t) {
        uint64_t s = 0x9e3779b97f4a7c15ull * (t + 1);
        for (int it = 0; it < 3000; it++) {
            const uint64_t n = sizes[rng(s) % sizes.size()];
            const int op = (int)(rng(s) % 17);
            if (op < 8) {
                uint32_t a[8], b[8], c[8];
                size_t np = 0;
                const int rc = gl_plan_passes(n, a, b, c, 8, &np);
                CHECK(rc == GL_OK || rc == GL_E_UNSUPPORTED);
                if (rc == GL_OK) {
                    uint64_t prod = 1;
                    for (size_t i = 0; i < np && i < 8; i++) prod *= (uint64_t)a[i] * b[i];
                    CHECK(np > 8 || prod == n || (np == 1 && a[0] == 65537));
                }
            } else if (op < 10) {
                uint32_t ids[8];
                size_t np = 0;
                const int rc = (op == 8 ? gl_plan_variants : gl_plan_variants_cols)(n, ids, 8, &np);
                CHECK(rc == GL_OK || rc == GL_E_UNSUPPORTED);
            } else if (op == 10) {
                // install and remove an explicit schedule for 2^20 / 2^12
                const uint32_t ids[3] = {id16x16, id16x16, id16x1};
                CHECK(gl_plan_set(1ull << 20, ids, 3) == GL_OK);
                CHECK(gl_plan_set(1ull << 20, nullptr, 0) == GL_OK);
                const uint32_t ids2[2] = {id16x16, id16x1};
                CHECK(gl_plan_set_cols(1ull << 12, ids2, 2) == GL_OK);
                CHECK(gl_plan_set_cols(1ull << 12, nullptr, 0) == GL_OK);
            } else if (op == 11) {
                CHECK(gl_variant_name((uint32_t)(rng(s) % (nvar + 1))) != nullptr);  // ids < count + 2
            } else if (op == 12) {
                CHECK(gl_next_size(n) == n);
                CHECK(gl_is_divisor(n) == 1);
                CHECK(gl_next_layer_size(160 * n) == 32 * n);
            } else if (op == 13 && it % 64 == 0) {
                gl_clear_cache();
            } else if (op == 14) {
                CHECK(gl_plan_set(n * 7, nullptr, 0) == GL_E_SIZE);
            } else if (op == 15) {
                // row tree sizes: leaves + ceil halves up to the root
                const size_t leaves = 1 + rng(s) % 5000;
                size_t nodes = leaves, l = leaves, depth = 0;
                while (l > 1) {
                    l = (l + 1) / 2;
                    nodes += l;
                    depth++;
                }
                CHECK(gl_row_tree_size(leaves) == 32 * nodes);
                CHECK(gl_row_tree_depth(leaves) == depth);
            } else if (!device_ok) {
                uint64_t v[4] = {1, 2, 3, 4};
                CHECK(gl_ntt(v, 4) == GL_E_DEVICE);
                CHECK(gl_last_error()[0] != 0);
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < 8; t++) th.emplace_back(worker, t);
    for (auto &x : th) x.join();
    // the tune file's 2^20 schedule is in force again after the explicit
    // ones were removed and the caches cleared
    uint32_t a[8], b[8], c[8];
    size_t np = 0;
    CHECK(gl_plan_passes(1ull << 20, a, b, c, 8, &np) == GL_OK);
    CHECK(np == 2 && a[0] == 16 && b[0] == 64);
    std::printf("host_threads_test: %s (%d failures)\n", g_fail.load() ? "FAILED" : "ok", g_fail.load());
    return g_fail.load() ? 1 : 0;
}
