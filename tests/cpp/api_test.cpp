// The reference's own tests, written against the C++ mirror (include/goldilocks.hpp):
// test_ntt vs naive (ntt/src/ntt/mod.rs:128-139), intt roundtrip, the
// convolution KAT (convolve.rs:44-60), K12 KATs (pcs/src/k12/generic.rs:113-160),
// merkle_tree layer sizes (mod.rs:75-84) and PoW KATs (pow.rs:67-99).
// Needs a gfx950 GPU.  Exit status 0 = all passed.
#include <cstdio>
#include <cstring>

#include "goldilocks.hpp"

using namespace goldilocks;
typedef unsigned __int128 u128;
static const Field P = 0xffffffff00000001ull;

static Field mulm(Field a, Field b) { return (Field)((u128)a * b % P); }
static Field powm(Field a, Field e) {
    Field r = 1;
    while (e) {
        if (e & 1) r = mulm(r, a);
        a = mulm(a, a);
        e >>= 1;
    }
    return r;
}

static std::vector<Field> naive(const std::vector<Field> &x) {
    size_t n = x.size();
    Field w = powm(2717, (P - 1) / n);
    std::vector<Field> out(n);
    for (size_t k = 0; k < n; k++) {
        Field acc = 0, wk = powm(w, k), wj = 1;
        for (size_t j = 0; j < n; j++) {
            acc = (Field)(((u128)acc + mulm(x[j], wj)) % P);
            wj = mulm(wj, wk);
        }
        out[k] = acc;
    }
    return out;
}

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);       \
            failures++;                                                   \
        }                                                                 \
    } while (0)

int main() {
    std::uint64_t s = 0x9e3779b97f4a7c15ull;
    auto rnd = [&]() {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        return s % P;
    };
    for (size_t n : {2, 3, 4, 5, 8, 15, 16, 17, 24, 64, 128, 240, 257, 1024}) {
        std::vector<Field> v(n);
        for (auto &x : v) x = rnd();
        auto expect = naive(v);
        auto w = v;
        ntt::strategy(n)->ntt(w);
        CHECK(w == expect);
        ntt::intt(w);
        CHECK(w == v);
    }
    bool threw = false;
    try {
        ntt::strategy(7);
    } catch (const Error &e) {
        threw = e.code == GL_E_SIZE;
    }
    CHECK(threw);
    {
        auto mont = [](Field x) { return (Field)(((u128)x << 64) % P); };
        std::vector<Field> a{mont(1), mont(2), mont(3)}, b{mont(4), mont(5), mont(6)};
        ntt::ntt(b);
        ntt::circular(a, b);
        CHECK((a == std::vector<Field>{mont(31), mont(31), mont(28)}));
    }
    {
        // permute/mod.rs test_copy_2_3, and a bit reversal of 8 words
        std::vector<Field> m{0, 1, 2, 3, 4, 5};
        ntt::permute::transpose(m, 2, 3);
        CHECK((m == std::vector<Field>{0, 3, 1, 4, 2, 5}));
        std::vector<Field> r{0, 1, 2, 3, 4, 5, 6, 7};
        ntt::permute::bit_reverse(r);
        CHECK((r == std::vector<Field>{0, 4, 2, 6, 1, 5, 3, 7}));
    }
    {
        auto h = pcs::k12<32>(nullptr, 0);
        const std::uint8_t e[4] = {0x1a, 0xc2, 0xd4, 0x50};
        CHECK(std::memcmp(h.data(), e, 4) == 0);
        std::uint8_t p17[17];
        for (int i = 0; i < 17; i++) p17[i] = (std::uint8_t)i;
        auto h17 = pcs::k12<32>(p17, 17);
        const std::uint8_t e17[4] = {0x6b, 0xf7, 0x5f, 0xa2};
        CHECK(std::memcmp(h17.data(), e17, 4) == 0);
    }
    {
        std::vector<std::uint8_t> in(1 << 20);
        for (size_t i = 0; i < in.size(); i++) in[i] = (std::uint8_t)(i % 0xfa);
        auto layers = pcs::merkle_tree(in.data(), in.size());
        CHECK(!layers.empty() && layers.back().size() == 32);
        for (auto &l : layers) CHECK(l.size() % 32 == 0);
    }
    {
        std::uint8_t ch[32];
        std::memset(ch, 0x55, 32);
        pcs::Pow pow(ch, 1ull << 63);
        CHECK(!pow.verify(0) && !pow.verify(1) && !pow.verify(2) && pow.verify(3));
        CHECK(pow.solve() == 3);
        pcs::Pow pow16(ch, ~0ull - (1ull << 48));
        CHECK(pow16.solve() == 18455);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
