// gl_ntt_dev_noalloc: gl_ntt_dev for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The ping-pong buffer of multi-pass schedules, which
// gl_ntt_dev takes from the device pool, lives here in a file-scope device
// array of fixed size instead; nothing is allocated on any call.
//
// Every call on one device shares that device's array, and it keeps the last
// call's contents (no call needs it zeroed: every pass writes each scratch
// word before the next pass reads it).  Calls on one stream run in order;
// calls that overlap on one device (different streams, or several host
// threads on different streams) race on the array.
#include <mutex>

#include "gl_runtime.hpp"

namespace {

// Scratch words: one 2^24-word (128 MiB) transform, or count transforms of
// n words with n * count <= 2^24.
constexpr size_t kScratchWords = size_t(1) << 24;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

// The array's device address, per device.  In host code `scratch` names a
// host-side stand-in, never device memory: kernels and copies get the
// address from hipGetSymbolAddress, taken once per device with that device
// current.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t *g_ptr[kMaxDevices];

int scratch_ptr(uint64_t **out) {
    int dev = 0;
    GL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return gl::set_error(GL_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return gl::hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = (uint64_t *)p;
    }
    *out = g_ptr[dev];
    return GL_OK;
}

}  // namespace

extern "C" int gl_ntt_dev_noalloc(uint64_t *d, size_t n, size_t count, int inverse, void *stream) {
    using namespace gl;
    int rc = ensure_init();
    if (rc) return rc;
    if (inverse && n <= 1) return GL_OK;  // ntt/src/ntt/mod.rs:105-107
    PlanRef plan;
    if ((rc = get_plan(n, &plan))) return rc;
    if (count == 0 || plan->passes.empty()) return GL_OK;
    if (!d) return set_error(GL_E_ARG, "null device pointer");
    if (plan_is_composite(*plan))
        return set_error(GL_E_UNSUPPORTED, "lengths with a factor 65537 need pool scratch; use gl_ntt_dev");
    if (words_overflow(n, count)) return set_error(GL_E_ARG, "n * count words overflow the address space");
    const size_t words = plan_tmp_words(*plan, count, n);
    if (words > kScratchWords)
        return set_error(GL_E_ARG, "n * count exceeds the no-allocation scratch (2^24 words); use gl_ntt_dev");
    uint64_t *tmp = nullptr;
    if (words && (rc = scratch_ptr(&tmp))) return rc;
    hipStream_t s = (hipStream_t)stream;
    PassCursor c{plan.get(), d, tmp, 1, count, n, inverse != 0};
    c.create_tables = false;  // a missing table is an error, never an allocation
    c.begin();
    while (!c.done()) {
        const PassVariant *v = plan->passes[c.p].v;
        PassArgs a;
        unsigned grid = 0;
        if ((rc = c.next(a, grid))) return rc;
        v->launch(a, grid, s);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}
