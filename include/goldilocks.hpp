// goldilocks.hpp - C++ mirror of the reference crates' entry points over the
// C ABI in goldilocks_gpu.h (header only; link libgoldilocks_amd.so).
//
//   goldilocks::ntt::ntt / intt / strategy / clear_cache   (ntt/src/ntt/mod.rs:43-120)
//   goldilocks::pcs::merkle_tree / process_layer / k12 /
//                    next_layer_size                       (pcs/src/k12/mod.rs:33-60,
//                                                           generic.rs:62-94)
//   goldilocks::pcs::Pow                                   (pcs/src/pow.rs:9-60)
//
// Field words are the raw u64 the reference stores in `[Field]`.  Where the
// reference panics these throw goldilocks::Error (a std::runtime_error that
// carries the GL_E_* code); nothing aborts the process.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "goldilocks_gpu.h"

namespace goldilocks {

using Field = std::uint64_t;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != GL_OK) throw Error(rc, gl_last_error());
}

namespace ntt {

// trait Ntt (mod.rs:25-31): a transform of length len() applied to every
// consecutive chunk of len() words.
class Ntt {
  public:
    explicit Ntt(std::size_t n) : n_(n) {
        if (!gl_is_divisor(n))
            throw Error(GL_E_SIZE, std::to_string(n) + " is not a supported NTT size");
    }
    std::size_t len() const { return n_; }
    void ntt(Field *values, std::size_t count_words) const {
        if (count_words % n_) throw Error(GL_E_ARG, "values.len() must be a multiple of len()");
        check(gl_ntt_batch(values, n_, count_words / n_, 0));
    }
    void ntt(std::vector<Field> &v) const { ntt(v.data(), v.size()); }

  private:
    std::size_t n_;
};

// strategy(size) (mod.rs:48-78); device plans are cached by the library.
inline std::shared_ptr<Ntt> strategy(std::size_t size) { return std::make_shared<Ntt>(size); }

inline void ntt(std::vector<Field> &values) { check(gl_ntt(values.data(), values.size())); }
inline void intt(std::vector<Field> &values) { check(gl_intt(values.data(), values.size())); }
inline void clear_cache() { gl_clear_cache(); }

// convolve::circular (ntt/src/convolve.rs:13-20): b already in NTT form.
inline void circular(std::vector<Field> &a, const std::vector<Field> &b) {
    if (a.size() != b.size()) throw Error(GL_E_ARG, "lengths differ");
    check(gl_convolve(a.data(), b.data(), a.size()));
}

// permute::transpose (ntt/src/permute/mod.rs:53-58): rows x cols -> cols x rows in place.
namespace permute {
inline void transpose(std::vector<Field> &matrix, std::size_t rows, std::size_t cols) {
    if (matrix.size() != rows * cols) throw Error(GL_E_ARG, "matrix length is not rows * cols");
    check(gl_transpose(matrix.data(), rows, cols));
}
// permute::bit_reverse (mod.rs:71-74): the length must be a power of two.
inline void bit_reverse(std::vector<Field> &values) { check(gl_bit_reverse(values.data(), values.size())); }
}  // namespace permute

}  // namespace ntt

namespace pcs {

constexpr std::size_t BLOCK_SIZE = 160, HASH_SIZE = 32;

inline std::size_t next_layer_size(std::size_t size) { return gl_next_layer_size(size); }

template <std::size_t H>
std::vector<std::uint8_t> k12(const std::uint8_t *input, std::size_t len) {
    static_assert(H <= 168, "single-block K12 output");
    std::vector<std::uint8_t> out(H);
    check(gl_k12(input, len, out.data(), H));
    return out;
}

inline void process_layer(const std::uint8_t *input, std::size_t len, std::uint8_t *output) {
    check(gl_k12_layer(input, len, output));
}

// merkle_tree (mod.rs:41-60): every layer, root last.
inline std::vector<std::vector<std::uint8_t>> merkle_tree(const std::uint8_t *values, std::size_t len) {
    std::size_t total = gl_merkle_tree_size(len);
    std::vector<std::uint8_t> flat(total ? total : 1);
    std::size_t got = 0;
    check(gl_merkle_tree(values, len, flat.data(), total, &got));
    std::vector<std::vector<std::uint8_t>> layers;
    std::size_t off = 0, n = len;
    while (n > HASH_SIZE) {
        n = next_layer_size(n);
        layers.emplace_back(flat.begin() + off, flat.begin() + off + n);
        off += n;
    }
    return layers;
}

// Pow (pow.rs:9-60)
class Pow {
  public:
    Pow(const std::uint8_t (&challenge)[32], std::uint64_t threshold) : threshold_(threshold) {
        for (int i = 0; i < 32; i++) challenge_[i] = challenge[i];
    }
    bool verify(std::uint64_t nonce) const {
        std::uint8_t msg[40];
        for (int i = 0; i < 32; i++) msg[i] = challenge_[i];
        for (int i = 0; i < 8; i++) msg[32 + i] = (std::uint8_t)(nonce >> (8 * i));
        auto h = k12<8>(msg, 40);
        std::uint64_t v = 0;
        for (int i = 7; i >= 0; i--) v = (v << 8) | h[i];
        return v >= threshold_;
    }
    std::uint64_t solve(std::uint64_t max_nonce = ~0ull) const {
        std::uint64_t nonce = 0;
        check(gl_pow_solve(challenge_, threshold_, max_nonce, &nonce));
        return nonce;
    }

  private:
    std::uint8_t challenge_[32];
    std::uint64_t threshold_;
};

}  // namespace pcs
}  // namespace goldilocks
