// Host build of the device field arithmetic and register codelets
// (goldilocks_amd/csrc/gl_field.hpp, gl_dft.hpp), checked against closed
// forms in 128-bit integer arithmetic and against the oracle's naive DFT.
// Built with `hipcc -x hip --offload-host-only`; needs no GPU.
#include <cstdio>
#include <random>
#include <vector>

#include "../../goldilocks_amd/csrc/gl_dft.hpp"
#include "../../goldilocks_amd/csrc/gl_internal.hpp"

extern "C" int or_naive_ntt(uint64_t *, size_t);
extern "C" void or_fill_random(uint64_t *, size_t, uint64_t);

using namespace gl;
typedef unsigned __int128 u128;

static u64 mulm(u64 a, u64 b) { return (u64)(((u128)a * b) % P); }
static u64 powm(u64 a, u64 e) {
    u64 r = 1;
    for (; e; e >>= 1, a = mulm(a, a))
        if (e & 1) r = mulm(r, a);
    return r;
}
static u64 pw2(int s) { return powm(2, (u64)s); }

static long g_bad = 0;
static void fail(const char *what, u64 a, u64 b) {
    if (g_bad++ < 10) printf("FAIL %s a=%016llx b=%016llx\n", what, (unsigned long long)a, (unsigned long long)b);
}

template <int S>
static void check_shift(u64 x, u64 raw) {
    constexpr int Sm = ((S % 192) + 192) % 192;
    if (shl<S>(x) != mulm(x, pw2(Sm))) fail("shl", x, (u64)S);
    if constexpr (Sm > 0 && Sm < 96) {
        // the sub-96 form accepts any 64-bit word, not only canonical ones
        if (shl_lt96<Sm>(raw) != mulm(raw % P, pw2(Sm))) fail("shl_lt96(raw)", raw, (u64)S);
    }
}
template <int... S>
static void check_shifts(u64 x, u64 raw) { (check_shift<S>(x, raw), ...); }

template <int N>
static void check_codelet() {
    u64 x[N], y[N];
    for (int seed = 0; seed < 12; seed++) {
        or_fill_random(x, N, 977u * seed + N);
        if (seed == 0)
            for (int i = 0; i < N; i++) x[i] = P - 1 - i;
        if (seed == 1)
            for (int i = 0; i < N; i++) x[i] = P - 1;
        for (int i = 0; i < N; i++) y[i] = x[i];
        dft<N>(x);
        or_naive_ntt(y, N);
        for (int i = 0; i < N; i++)
            if (x[i] != y[i]) {
                fail("codelet", (u64)N, (u64)seed);
                break;
            }
    }
}

int main() {
    const u64 rinv = powm(pw2(64), P - 2);  // 2^-64 mod p
    std::mt19937_64 g(7);
    std::vector<u64> v = {0, 1, 2, P - 1, P - 2, EPS, EPS + 1, EPS - 1, 1ull << 32, 1ull << 63, 0xffffffff00000000ull};
    for (int i = 0; i < 64; i++) {
        v.push_back(1ull << i);
        v.push_back((1ull << i) - 1);
        v.push_back(P - (1ull << i));
    }
    for (u64 i = 0; i < 40; i++) {
        v.push_back(P - 1 - (i << 32));
        v.push_back(0xffffffff00000000ull - i);
        v.push_back(i << 32);
    }
    for (int i = 0; i < 60000; i++) v.push_back(g() % P);
    for (int i = 0; i < 6000; i++) {
        v.push_back(P - 1 - (g() & 0xffffffffffull));
        v.push_back(g() & 0xfffffffffull);
        v.push_back((g() | 0xffffffff00000000ull) % P);
    }
    for (size_t i = 0; i < v.size(); i++) {
        const u64 a = v[i];
        for (u64 b : {v[(i * 7919 + 13) % v.size()], a, v[i % 200]}) {
            if (add(a, b) != (u64)(((u128)a + b) % P)) fail("add", a, b);
            if (sub(a, b) != (u64)(((u128)a + P - b) % P)) fail("sub", a, b);
            if (mul(a, b) != mulm(a, b)) fail("mul", a, b);
            if (mont_mul(a, b) != mulm(mulm(a, b), rinv)) fail("mont_mul", a, b);
        }
        check_shifts<1, 3, 6, 12, 24, 31, 32, 33, 36, 40, 47, 48, 60, 63, 64, 65, 72, 84, 95, 96, 97, 108, 120,
                     132, 144, 156, 168, 180, 191, -12>(a, g());
        if (g_bad) break;
    }
    check_codelet<2>(); check_codelet<3>(); check_codelet<4>(); check_codelet<5>(); check_codelet<6>();
    check_codelet<8>(); check_codelet<10>(); check_codelet<15>(); check_codelet<16>(); check_codelet<17>();
    check_codelet<20>(); check_codelet<32>(); check_codelet<34>(); check_codelet<40>(); check_codelet<48>();
    check_codelet<51>(); check_codelet<60>(); check_codelet<64>(); check_codelet<85>();
    // the kernels' instance division: (x * mul) >> sh == x / d for x < 2^31
    // and 1 <= d < 2^32 (fastdiv_params), near every power of two and at random
    std::vector<u64> ds;
    for (int k = 0; k < 32; k++)
        for (int o = -2; o <= 2; o++)
            if ((1ll << k) + o >= 1) ds.push_back((1ull << k) + o);
    for (int i = 0; i < 2000; i++) ds.push_back((g() >> (g() % 33 + 31)) | 1);
    for (u64 d : ds) {
        u32 m, sh;
        fastdiv_params(d, m, sh);
        const u64 top = (u64(1) << 31) - 1;
        for (u64 x : {u64(0), d - 1, d, d + 1, 2 * d - 1, top, top - (top + 1) % d, u64(g() % (top + 1)),
                      u64(g() % (top + 1))}) {
            if (x > top) continue;
            if (fastdiv((u32)x, m, sh) != x / d) fail("fastdiv", x, d);
        }
    }
    printf("%zu field values, 19 codelet sizes, %zu divisors: %s\n", v.size(), ds.size(), g_bad ? "FAIL" : "ok");
    return g_bad != 0;
}
