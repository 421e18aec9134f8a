// NTT pass kernels for gfx950 (CDNA4, wave64): templates shared by the
// translation units that instantiate the variant table (gl_variants_*.hip).
//
// A pass computes one Stockham DIF step of radix R = R1*R2 for T consecutive
// instances per workgroup (see PassArgs).  The radix-R DFT of an instance is
// split as R = R1 x R2 (ntt/src/ntt/cooley_tukey.rs:80-148 restated for one
// level): sub-step a runs DFT_R1 in registers, multiplies by w_R^(ka*rb) and
// writes to LDS; sub-step b reads the transposed view, runs DFT_R2 in
// registers, applies the pass's output twiddle w_M^(m*k) and stores.
//
// Memory: lanes map to consecutive instances first, so every global load and
// store instruction touches runs of T (or more) consecutive words; each word
// is read once and written once per pass.  LDS rows are padded (odd stride
// per instance) and XOR-swizzled (power-of-two R2) so the transposed reads
// of sub-step b do not serialize on banks.
#pragma once

#include "gl_dft.hpp"
#include "gl_internal.hpp"

namespace gl {

// Byte offset k * stride (stride a uniform word count, k a compile-time
// constant) formed on the scalar unit; the empty asm keeps the product out
// of a per-element v_mad_u64_u32, so each strided access costs one 64-bit
// vector add.
__device__ __forceinline__ u64 uniform_offset(u64 stride, int k) {
    u64 o = stride * (u64)k * 8;
    asm("" : "+s"(o));
    return o;
}
template <typename W>
__device__ __forceinline__ W *at_offset(W *p, u64 bytes) {
    return (W *)((char *)p + bytes);
}

// Montgomery form of w_M^e from the two-level power table.
__device__ __forceinline__ u64 tw_pow(const PassArgs &a, u64 e) {
    return mont_mul(a.tw_lo[e & ((1ull << a.lo_bits) - 1)], a.tw_hi[e >> a.lo_bits]);
}

// v[j] -> dst[j * dstep] for j < NK.
template <int NK>
__device__ __forceinline__ void store_all(u64 *dst, u64 dstep, const u64 *v) {
#pragma unroll
    for (int j = 0; j < NK; j++) *at_offset(dst, uniform_offset(dstep, j)) = v[j];
}

// Store the NK outputs v[j] (output index k = k0 + j*kstep) of instance m,
// applying the pass's output twiddle w_M^(m*k) and scale.  With the full
// table each twiddle is one load at a compile-time offset from the
// instance's row.  Otherwise the twiddles form the geometric sequence
// A * B^j with A = w^(m*k0), B = w^(m*kstep): two table lookups per thread,
// then one product per output to advance (exact field arithmetic, so the
// result equals the per-output lookup).
template <int NK, int KSTEP>
__device__ __forceinline__ void store_outputs(const PassArgs &a, u64 *dst, u64 dstep, u32 s, u32 m, u32 k0, u32 R,
                                              u64 *v) {
    // The scale factors commute with the twiddle (exact products), so they
    // come first and each twiddle mode ends in its own store loop: no join
    // of the branches' registers before the stores.
    if (a.flags & FLAG_SCALE) {
#pragma unroll
        for (int j = 0; j < NK; j++) v[j] = mont_mul(v[j], a.scale);
    } else if (a.flags & FLAG_ROW_SCALE) {
        // output j sits at position s + S*(k0 + j*KSTEP + R*m) of its item:
        // row s/S0 + SS*(k0 + R*m) + SS*KSTEP*j
        const u64 *rs = a.row_scale + fastdiv(s, a.S0_mul, a.S0_sh) + a.SS * ((u64)k0 + (u64)R * m);
        const u64 rstep = a.SS * (u64)KSTEP;
#pragma unroll
        for (int j = 0; j < NK; j++) v[j] = mont_mul(v[j], *at_offset(rs, uniform_offset(rstep, j)));
    }
    if (a.flags & FLAG_TWIDDLE_FULL) {
        const u64 *tw = a.tw_lo + (u64)R * (u64)m + k0;
        u64 w[NK];
#pragma unroll
        for (int j = 0; j < NK; j++) w[j] = tw[j * KSTEP];
#pragma unroll
        for (int j = 0; j < NK; j++) v[j] = mont_mul(v[j], w[j]);
        store_all<NK>(dst, dstep, v);
    } else if (a.flags & FLAG_TWIDDLE) {
        u64 w = tw_pow(a, (u64)m * k0);
        if (a.flags & FLAG_SCALE_TW) w = mont_mul(w, a.scale);
        const u64 B = tw_pow(a, (u64)m * KSTEP);
#pragma unroll
        for (int j = 0; j < NK; j++) {
            v[j] = mont_mul(v[j], w);
            if (j + 1 < NK) w = mont_mul(w, B);
        }
        store_all<NK>(dst, dstep, v);
    } else {
        store_all<NK>(dst, dstep, v);
    }
}

// LDS tile of the two-sub-step pass: instance t owns TS words, row ka of it
// RS words (rb along the row).  Odd strides keep both the row-wise writes
// (lanes over t) and the column-wise reads (lanes over t, or over ka) on
// distinct banks, and every per-element offset is a compile-time constant.
template <int R2>
constexpr int lds_rs() { return R2 | 1; }
template <int R1, int R2>
constexpr int lds_ts() { return (R1 * lds_rs<R2>()) | 1; }

// Exchange through LDS in 32-bit halves (low words, then high words): half
// the LDS of the 64-bit tile, so up to twice the workgroups fit per CU, for two
// more barriers and the high words held in registers meanwhile.  Used for
// radices >= GL_HALF_EXCHANGE_MIN_R (512: the 1024-point 32x32 pass on a
// 2^28-word transform ran 1783 -> 1423 us, the 256-point passes gained
// nothing) where each thread has one task per sub-step.
#ifndef GL_HALF_EXCHANGE_MIN_R
#define GL_HALF_EXCHANGE_MIN_R 512
#endif
template <int R1, int R2, int NT, int T>
constexpr bool half_exchange() {
    return GL_HALF_EXCHANGE_MIN_R > 0 && R2 > 1 && R1 * R2 >= GL_HALF_EXCHANGE_MIN_R && T * R2 <= NT &&
           T * R1 <= NT && R1 + R2 <= 64;
}

// Worst bank multiplicity (32 banks, per 32-lane half wave) of the 32-bit
// tile's three access patterns: sub-step a writes (lanes over t, then rb),
// sub-step b reads with S == 1 (lanes over ka, then t) and with S > 1
// (lanes over t, then ka).
template <int R1, int T>
constexpr int lds32_conflicts(int RS, int TS) {
    int worst = 0;
    for (int pat = 0; pat < 3; pat++) {
        for (int h = 0; h < 2; h++) {
            int cnt[32] = {};
            for (int l = 0; l < 32; l++) {
                const int L = h * 32 + l;
                const int addr = pat == 0 ? (L % T) * TS + L / T
                               : pat == 1 ? (L / R1) * TS + (L % R1) * RS
                                          : (L % T) * TS + (L / T) * RS;
                const int c = ++cnt[addr % 32];
                worst = c > worst ? c : worst;
            }
        }
    }
    return worst;
}
// Instance stride of the 32-bit tile: the least padding with the fewest
// conflicts (rows keep the odd stride R2 | 1).
template <int R1, int R2, int T>
constexpr int lds32_ts() {
    constexpr int RS = lds_rs<R2>();
    int best = R1 * RS, fewest = 1 << 30;
    for (int ts = R1 * RS; ts < R1 * RS + 64; ts++) {
        const int c = lds32_conflicts<R1, T>(RS, ts);
        if (c < fewest) {
            fewest = c;
            best = ts;
        }
    }
    return best;
}

// Waves per SIMD the half-exchange kernels are compiled for (register cap;
// 1 = none: the 1024-point pass takes 147 VGPRs, three waves, and a cap of
// four waves made it spill and run 25-50% slower).
#ifndef GL_HALF_MIN_WAVES
#define GL_HALF_MIN_WAVES 1
#endif
template <int R1, int R2, int NT, int T>
constexpr int pass_min_waves() { return half_exchange<R1, R2, NT, T>() ? GL_HALF_MIN_WAVES : 1; }

template <int R1, int R2, int NT, int T>
constexpr int lds_words() {
    if constexpr (half_exchange<R1, R2, NT, T>()) return (T * lds32_ts<R1, R2, T>() + 1) / 2;
    return R2 > 1 ? T * lds_ts<R1, R2>() : 1;
}

// Sub-step a of the task (instance I, row rb): the R1 inputs of row rb,
// DFT_R1 and, with R2 > 1, the inner twiddles w_R^(ka*rb).  b, i, s, m
// locate the instance; lanes past the last instance load the last
// instance's words (their results are never stored), so the loads need no
// branch and no zero fill.  pass_body keeps the same steps inline: calling
// this helper there raised its register use (82 -> 110 VGPRs for 16x16).
template <int R1, int R2>
__device__ __forceinline__ void substep_a(const PassArgs &a, u32 I, int rb, u64 *x, u32 &b, u32 &i, u32 &s,
                                          u32 &m) {
    const bool valid = I < a.total_inst;
    // inner twiddles: row rb of the table, issued before the data loads so
    // one wait covers both
    u64 w[R1];
    if constexpr (R2 > 1) {
        const u64 *twr = a.tw_inner + rb * R1;
#pragma unroll
        for (int ka = 1; ka < R1; ka++) w[ka] = twr[ka];
    }
    locate(a, valid ? I : a.total_inst - 1, b, i, s, m);
    const u64 *src = a.in + (u64)b * a.bstride;
    if (a.flags & FLAG_REVERSE) {
        // as in pass_body: a descending run, pos = 0 read at 0
        const u64 step = a.SMp * (u64)R2;
        const u64 *p = src + s + a.S * (a.n_rev - (u64)m - (u64)a.Mp * (u64)rb);
        const u64 *q0 = (m == 0 && rb == 0) ? src + s : p;
        x[0] = *q0;
#pragma unroll
        for (int ra = 1; ra < R1; ra++) x[ra] = *at_offset(p, 0 - uniform_offset(step, ra));
    } else if (a.flags & FLAG_COSET) {
        // input row j = m + Mp*r of a zero-extended, coset-shifted
        // column: coef[j] * g^j below coset_rows, zero above
        const u64 *p = src + (u64)i + a.SMp * (u64)rb;
        const u64 step = a.SMp * (u64)R2;
#pragma unroll
        for (int ra = 0; ra < R1; ra++) {
            const u64 j = (u64)m + (u64)a.Mp * (u64)(ra * R2 + rb);
            x[ra] = j < a.coset_rows ? mont_mul(*at_offset(p, uniform_offset(step, ra)), a.coset[j]) : 0;
        }
    } else if (a.flags & FLAG_ZEXT) {
        // zero-extended column: rows j >= coset_rows read as zero
        const u64 *p = src + (u64)i + a.SMp * (u64)rb;
        const u64 step = a.SMp * (u64)R2;
#pragma unroll
        for (int ra = 0; ra < R1; ra++) {
            const u64 j = (u64)m + (u64)a.Mp * (u64)(ra * R2 + rb);
            x[ra] = j < a.coset_rows ? *at_offset(p, uniform_offset(step, ra)) : 0;
        }
    } else {
        const u64 *p = src + (u64)i + a.SMp * (u64)rb;
        const u64 step = a.SMp * (u64)R2;
#pragma unroll
        for (int ra = 0; ra < R1; ra++) x[ra] = *at_offset(p, uniform_offset(step, ra));
    }
    dft<R1>(x);
    if constexpr (R2 > 1) {
#pragma unroll
        for (int ka = 1; ka < R1; ka++) x[ka] = mont_mul(x[ka], w[ka]);
    }
}

// Sub-step b of the task (instance at b, s, m; column ka): DFT_R2 over the
// column and the stores of outputs k = ka + R1*kb.
template <int R1, int R2>
__device__ __forceinline__ void substep_b(const PassArgs &a, bool valid, u32 b, u32 s, u32 m, int ka, u64 *z) {
    constexpr int R = R1 * R2;
    dft<R2>(z);
    if (valid) {
        u64 *dst = a.out + (u64)b * a.bstride + s + a.S * ((u64)ka + (u64)R * (u64)m);
        store_outputs<R2, R1>(a, dst, a.S * (u64)R1, s, m, (u32)ka, (u32)R, z);
    }
}

// Sub-step b's task for thread slot tb.  S == 1: an instance's outputs are
// contiguous, so lanes run over ka first; otherwise over consecutive
// instances.
template <int R1, int T>
__device__ __forceinline__ void task_b(const PassArgs &a, int tb, int &t, int &ka) {
    if (a.S == 1) {
        ka = tb % R1;
        t = tb / R1;
    } else {
        t = tb % T;
        ka = tb / T;
    }
}

// Two-sub-step pass exchanging 32-bit halves, one task per thread and
// sub-step: sub-step a writes the low words and keeps the high words in
// registers; sub-step b reads the low words, then (after the high words
// replace them) the high words.
template <int R1, int R2, int NT, int T>
__device__ __forceinline__ void pass_body_half(const PassArgs &a, u32 blk, u32 *lds) {
    constexpr int RS = lds_rs<R2>(), TS = lds32_ts<R1, R2, T>();
    const int tid = threadIdx.x;
    const u32 I0 = blk * (u32)T;
    const bool act_a = tid < T * R2, act_b = tid < T * R1;
    u32 *wrow = lds + (tid % T) * TS + tid / T;
    u32 xh[R1];
    if (act_a) {
        u64 x[R1];
        u32 b, i, s, m;
        substep_a<R1, R2>(a, I0 + (u32)(tid % T), tid / T, x, b, i, s, m);
#pragma unroll
        for (int ka = 0; ka < R1; ka++) {
            wrow[ka * RS] = lo32(x[ka]);
            xh[ka] = hi32(x[ka]);
        }
    }
    __syncthreads();
    int t, ka;
    task_b<R1, T>(a, tid, t, ka);
    const u32 I = I0 + (u32)t;
    const bool valid = act_b && I < a.total_inst;
    u32 b = 0, i = 0, s = 0, m = 0;
    if (valid) locate(a, I, b, i, s, m);
    const u32 *rrow = lds + t * TS + ka * RS;
    u32 zl[R2];
    if (act_b) {
#pragma unroll
        for (int rb = 0; rb < R2; rb++) zl[rb] = rrow[rb];
    }
    __syncthreads();
    if (act_a) {
#pragma unroll
        for (int ka2 = 0; ka2 < R1; ka2++) wrow[ka2 * RS] = xh[ka2];
    }
    __syncthreads();
    if (act_b) {
        u64 z[R2];
#pragma unroll
        for (int rb = 0; rb < R2; rb++) z[rb] = mk64(zl[rb], rrow[rb]);
        substep_b<R1, R2>(a, valid, b, s, m, ka, z);
    }
}

// One workgroup's share of a pass: instances [blk*T, blk*T + T).
template <int R1, int R2, int NT, int T>
__device__ __forceinline__ void pass_body(const PassArgs &a, u32 blk, u64 *lds) {
    if constexpr (half_exchange<R1, R2, NT, T>()) {
        pass_body_half<R1, R2, NT, T>(a, blk, (u32 *)lds);
        return;
    }
    constexpr int R = R1 * R2;
    constexpr int RS = lds_rs<R2>(), TS = lds_ts<R1, R2>();
    const int tid = threadIdx.x;
    const u32 I0 = blk * (u32)T;

    constexpr int TA = T * R2;
    constexpr int QA = (TA + NT - 1) / NT;
#pragma unroll 1
    for (int q = 0; q < QA; q++) {
        const int ta = tid + q * NT;
        if (ta >= TA) break;
        const int t = ta % T, rb = ta / T;
        const u32 I = I0 + (u32)t;
        const bool valid = I < a.total_inst;
        if constexpr (R2 == 1 && R1 >= 32 && is_pow2(R1)) {
            if (a.flags & FLAG_ZEXT_HALF) {
                // zero extension to twice the rows: inputs r >= R1/2 are the
                // zero rows, so only the lower half is read and the first
                // codelet layer needs no additions
                u64 x[R1];
                u32 b, i, s, m;
                locate(a, valid ? I : a.total_inst - 1, b, i, s, m);
                const u64 *p = a.in + (u64)b * a.bstride + (u64)i;
#pragma unroll
                for (int ra = 0; ra < R1 / 2; ra++) x[ra] = *at_offset(p, uniform_offset(a.SMp, ra));
                dft_pow2_upper_zero<R1>(x);
                if (valid) {
                    u64 *dst = a.out + (u64)b * a.bstride + s + a.S * (u64)R * (u64)m;
                    store_outputs<R1, 1>(a, dst, a.S, s, m, 0, R, x);
                }
                continue;
            }
        }
        // inner twiddles w_R^(ka*rb): row rb of the table, issued before
        // the data loads so one wait covers both
        u64 w[R1];
        if constexpr (R2 > 1) {
            const u64 *twr = a.tw_inner + rb * R1;
#pragma unroll
            for (int ka = 1; ka < R1; ka++) w[ka] = twr[ka];
        }
        // lanes past the last instance load the last instance's words (their
        // results are never stored), so the loads need no branch and no
        // zero fill
        u64 x[R1];
        u32 b, i, s, m;
        {
            locate(a, valid ? I : a.total_inst - 1, b, i, s, m);
            const u64 *src = a.in + (u64)b * a.bstride;
            if (a.flags & FLAG_REVERSE) {
                // position pos = m + Mp*(ra*R2 + rb) is read at n - pos (0 at
                // pos = 0): a descending run with the forward path's uniform
                // step, except pos = 0 (ra = rb = m = 0)
                const u64 step = a.SMp * (u64)R2;
                const u64 *p = src + s + a.S * (a.n_rev - (u64)m - (u64)a.Mp * (u64)rb);
                const u64 *q0 = (m == 0 && rb == 0) ? src + s : p;
                x[0] = *q0;
#pragma unroll
                for (int ra = 1; ra < R1; ra++) x[ra] = *at_offset(p, 0 - uniform_offset(step, ra));
            } else if (a.flags & FLAG_COSET) {
                // input row j = m + Mp*r of a zero-extended, coset-shifted
                // column: coef[j] * g^j below coset_rows, zero above
                const u64 *p = src + (u64)i + a.SMp * (u64)rb;
                const u64 step = a.SMp * (u64)R2;
#pragma unroll
                for (int ra = 0; ra < R1; ra++) {
                    const u64 j = (u64)m + (u64)a.Mp * (u64)(ra * R2 + rb);
                    x[ra] = j < a.coset_rows ? mont_mul(*at_offset(p, uniform_offset(step, ra)), a.coset[j]) : 0;
                }
            } else if (a.flags & FLAG_ZEXT) {
                // zero-extended column: rows j >= coset_rows read as zero
                const u64 *p = src + (u64)i + a.SMp * (u64)rb;
                const u64 step = a.SMp * (u64)R2;
#pragma unroll
                for (int ra = 0; ra < R1; ra++) {
                    const u64 j = (u64)m + (u64)a.Mp * (u64)(ra * R2 + rb);
                    x[ra] = j < a.coset_rows ? *at_offset(p, uniform_offset(step, ra)) : 0;
                }
            } else {
                const u64 *p = src + (u64)i + a.SMp * (u64)rb;
                const u64 step = a.SMp * (u64)R2;
#pragma unroll
                for (int ra = 0; ra < R1; ra++) x[ra] = *at_offset(p, uniform_offset(step, ra));
            }
        }
        dft<R1>(x);
        if constexpr (R2 > 1) {
#pragma unroll
            for (int ka = 1; ka < R1; ka++) x[ka] = mont_mul(x[ka], w[ka]);
            u64 *row = lds + t * TS + rb;
#pragma unroll
            for (int ka = 0; ka < R1; ka++) row[ka * RS] = x[ka];
        } else if (valid) {
            // single sub-step pass: k = ka
            u64 *dst = a.out + (u64)b * a.bstride + s + a.S * (u64)R * (u64)m;
            store_outputs<R1, 1>(a, dst, a.S, s, m, 0, R, x);
        }
    }
    if constexpr (R2 > 1) {
        __syncthreads();
        constexpr int TB = T * R1;
        constexpr int QB = (TB + NT - 1) / NT;
#pragma unroll 1
        for (int q = 0; q < QB; q++) {
            const int tb = tid + q * NT;
            if (tb >= TB) break;
            // S == 1: an instance's outputs are contiguous, so lanes run over
            // ka first; otherwise over consecutive instances.
            int t, ka;
            if (a.S == 1) {
                ka = tb % R1;
                t = tb / R1;
            } else {
                t = tb % T;
                ka = tb / T;
            }
            const u32 I = I0 + (u32)t;
            const bool valid = I < a.total_inst;
            u32 b = 0, i = 0, s = 0, m = 0;
            if (valid) locate(a, I, b, i, s, m);
            u64 z[R2];
            const u64 *row = lds + t * TS + ka * RS;
#pragma unroll
            for (int rb = 0; rb < R2; rb++) z[rb] = row[rb];
            dft<R2>(z);
            if (valid) {
                u64 *dst = a.out + (u64)b * a.bstride + s + a.S * ((u64)ka + (u64)R * (u64)m);
                store_outputs<R2, R1>(a, dst, a.S * (u64)R1, s, m, (u32)ka, (u32)R, z);
            }
        }
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Remap so that consecutive logical blocks, which touch
// neighbouring words of the same cache lines, run on one XCD and share its L2.
__device__ __forceinline__ u32 xcd_block(u32 b, u32 grid) {
    const u32 per = grid / 8;
    if (b >= per * 8) return b;
    return (b % 8) * per + b / 8;
}

// Three-stage pass: R = R1*R2*R3, DFT_R by decimation in frequency with
// register codelets DFT_R1, DFT_R2, DFT_R3 and two LDS exchanges.
//   stage 1, task (t, r2, r3): DFT_R1 over r1, times w_R^(k1*(r2*R3 + r3))
//   stage 2, task (t, k1, r3): DFT_R2 over r2, times w_(R2*R3)^(k2*r3)
//   stage 3, task (t, k1, k2): DFT_R3 over r3 -> k = k1 + R1*k2 + R1*R2*k3
// Lanes run over consecutive instances first for loads; for stores the
// lane order follows the output layout (k1 first when S == 1, where each
// instance's outputs are contiguous).
template <int R3>
__device__ __forceinline__ int sw(int c, int key) {
    if constexpr (is_pow2(R3)) return c ^ (key & (R3 - 1));
    else return c;
}

template <int R1, int R2, int R3, int NT, int T>
__device__ __forceinline__ void pass3_body(const PassArgs &a, u32 blk, u64 *lds) {
    constexpr int R23 = R2 * R3, R = R1 * R23;
    constexpr int TS = R + 1;
    static_assert(T * R1 * R3 == NT, "stage 2 runs one task per thread");
    constexpr int TA = T * R23, QA = (TA + NT - 1) / NT;
    constexpr int TC = T * R1 * R2, QC = (TC + NT - 1) / NT;
    const int tid = threadIdx.x;
    const u32 I0 = blk * (u32)T;
    // ---- stage 1
#pragma unroll 1
    for (int q = 0; q < QA; q++) {
        const int ta = tid + q * NT;
        if (ta >= TA) break;
        const int t = ta % T, r3 = (ta / T) % R3, r2 = ta / (T * R3);
        const u32 I = I0 + (u32)t;
        const int c = r2 * R3 + r3;
        u64 w[R1];
#pragma unroll
        for (int k1 = 1; k1 < R1; k1++) w[k1] = a.tw_inner[(k1 * c) % R];
        u64 x[R1];
        {
            // lanes past the last instance load its words; never stored
            u32 b, i, s, m;
            locate(a, I < a.total_inst ? I : a.total_inst - 1, b, i, s, m);
            const u64 *src = a.in + (u64)b * a.bstride;
            if (a.flags & FLAG_REVERSE) {
#pragma unroll
                for (int r1 = 0; r1 < R1; r1++) {
                    u64 pos = (u64)m + (u64)a.Mp * (u64)(r1 * R23 + r2 * R3 + r3);
                    x[r1] = src[s + a.S * (pos ? a.n_rev - pos : 0)];
                }
            } else if (a.flags & FLAG_ZEXT) {
                // zero-extended column: rows j >= coset_rows read as zero
                const u64 *p = src + (u64)i + a.SMp * (u64)(r2 * R3 + r3);
#pragma unroll
                for (int r1 = 0; r1 < R1; r1++) {
                    const u64 j = (u64)m + (u64)a.Mp * (u64)(r1 * R23 + r2 * R3 + r3);
                    x[r1] = j < a.coset_rows ? *at_offset(p, uniform_offset(a.SMp, r1 * R23)) : 0;
                }
            } else {
                const u64 *p = src + (u64)i + a.SMp * (u64)(r2 * R3 + r3);
#pragma unroll
                for (int r1 = 0; r1 < R1; r1++) x[r1] = *at_offset(p, uniform_offset(a.SMp, r1 * R23));
            }
        }
        dft<R1>(x);
#pragma unroll
        for (int k1 = 1; k1 < R1; k1++) x[k1] = mont_mul(x[k1], w[k1]);
        u64 *row = lds + t * TS + r2 * R3 + sw<R3>(r3, r2);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) row[k1 * R23] = x[k1];
    }
    __syncthreads();
    // ---- stage 2 (reads complete before any thread overwrites)
    {
        const int t = tid % T, r3 = (tid / T) % R3, k1 = tid / (T * R3);
        u64 w[R2];
#pragma unroll
        for (int k2 = 1; k2 < R2; k2++) w[k2] = a.tw_inner[(R1 * k2 * r3) % R];
        u64 y[R2];
        const u64 *row = lds + t * TS + k1 * R23;
#pragma unroll
        for (int r2 = 0; r2 < R2; r2++) y[r2] = row[r2 * R3 + sw<R3>(r3, r2)];
        dft<R2>(y);
#pragma unroll
        for (int k2 = 1; k2 < R2; k2++) y[k2] = mont_mul(y[k2], w[k2]);
        __syncthreads();
        u64 *wrow = lds + t * TS + k1 * R23 + sw<R3>(r3, k1);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) wrow[k2 * R3] = y[k2];
    }
    __syncthreads();
    // ---- stage 3
#pragma unroll 1
    for (int q = 0; q < QC; q++) {
        const int tc = tid + q * NT;
        if (tc >= TC) break;
        int t, k1, k2;
        if (a.S == 1) {
            k1 = tc % R1;
            t = (tc / R1) % T;
            k2 = tc / (R1 * T);
        } else {
            t = tc % T;
            k1 = (tc / T) % R1;
            k2 = tc / (T * R1);
        }
        const u32 I = I0 + (u32)t;
        u64 z[R3];
        const u64 *row = lds + t * TS + k1 * R23 + k2 * R3;
#pragma unroll
        for (int r3 = 0; r3 < R3; r3++) z[r3] = row[sw<R3>(r3, k1)];
        dft<R3>(z);
        if (I < a.total_inst) {
            u32 b, i, s, m;
            locate(a, I, b, i, s, m);
            const u32 kk = (u32)(k1 + R1 * k2);
            u64 *dst = a.out + (u64)b * a.bstride + s + a.S * ((u64)kk + (u64)R * (u64)m);
            store_outputs<R3, R1 * R2>(a, dst, a.S * (u64)(R1 * R2), s, m, kk, (u32)R, z);
        }
    }
}

template <int R1, int R2, int R3, int NT, int T>
__global__ __launch_bounds__(NT) void ntt_pass3_kernel(PassArgs a) {
    __shared__ u64 lds[T * (R1 * R2 * R3 + 1)];
    pass3_body<R1, R2, R3, NT, T>(a, xcd_block(blockIdx.x, gridDim.x), lds);
}

template <int R1, int R2, int R3, int NT, int T>
__global__ __launch_bounds__(NT) void ntt_pass3_grouped_kernel(const PassArgs *__restrict__ descs,
                                                               const u32 *__restrict__ starts, u32 nd) {
    __shared__ u64 lds[T * (R1 * R2 * R3 + 1)];
    const u32 b = blockIdx.x;
    u32 lo = 0, hi = nd;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) / 2;
        if (starts[mid] <= b) lo = mid; else hi = mid;
    }
    const PassArgs a = descs[lo];
    pass3_body<R1, R2, R3, NT, T>(a, b - starts[lo], lds);
}

template <int R1, int R2, int R3, int NT, int T>
static void launch_pass3(const PassArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((ntt_pass3_kernel<R1, R2, R3, NT, T>), dim3(grid), dim3(NT), 0, s, a);
}

template <int R1, int R2, int R3, int NT, int T>
static void launch_pass3_grouped(const PassArgs *descs, const u32 *starts, u32 nd, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((ntt_pass3_grouped_kernel<R1, R2, R3, NT, T>), dim3(grid), dim3(NT), 0, s, descs, starts,
                       nd);
}

template <int R1, int R2, int NT, int T>
__global__ __launch_bounds__(NT, (pass_min_waves<R1, R2, NT, T>())) void ntt_pass_kernel(PassArgs a) {
    __shared__ u64 lds[lds_words<R1, R2, NT, T>()];
    pass_body<R1, R2, NT, T>(a, xcd_block(blockIdx.x, gridDim.x), lds);
}

// Grouped launch: descriptor j owns blocks [starts[j], starts[j+1]).  Used
// to run the same pass variant of many transforms (of different lengths)
// in one launch.
template <int R1, int R2, int NT, int T>
__global__ __launch_bounds__(NT, (pass_min_waves<R1, R2, NT, T>())) void ntt_pass_grouped_kernel(const PassArgs *__restrict__ descs,
                                                              const u32 *__restrict__ starts, u32 nd) {
    __shared__ u64 lds[lds_words<R1, R2, NT, T>()];
    const u32 b = blockIdx.x;
    u32 lo = 0, hi = nd;  // find the last j with starts[j] <= b
    while (hi - lo > 1) {
        u32 mid = (lo + hi) / 2;
        if (starts[mid] <= b) lo = mid; else hi = mid;
    }
    const PassArgs a = descs[lo];
    pass_body<R1, R2, NT, T>(a, b - starts[lo], lds);
}

template <int R1, int R2, int NT, int T>
static void launch_pass(const PassArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((ntt_pass_kernel<R1, R2, NT, T>), dim3(grid), dim3(NT), 0, s, a);
}

template <int R1, int R2, int NT, int T>
static void launch_grouped(const PassArgs *descs, const u32 *starts, u32 nd, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((ntt_pass_grouped_kernel<R1, R2, NT, T>), dim3(grid), dim3(NT), 0, s, descs, starts, nd);
}

// T: enough instances that both sub-steps have >= NT tasks, within 136 KiB LDS.
template <int R1, int R2, int NT>
constexpr int pick_T() {
    if (R2 == 1) return NT;
    constexpr int mn = R1 < R2 ? R1 : R2;
    int t = (NT + mn - 1) / mn;
    while (t > 1 && (size_t)t * lds_ts<R1, R2>() * 8 > 136 * 1024) t--;
    return t;
}

#define GL_VARIANT_T(R1, R2, NT, T)                                                          \
    {R1, R2, T, NT, (size_t)lds_words<R1, R2, NT, T>() * 8, &launch_pass<R1, R2, NT, T>,         \
     &launch_grouped<R1, R2, NT, T>, VK_TWO_STAGE, 1}
#define GL_VARIANT(R1, R2) GL_VARIANT_T(R1, R2, 256, (pick_T<R1, R2, 256>()))
// Three-stage variants: the R2 field holds R2*R3.
#define GL_VARIANT3(R1, R2, R3, NT, T)                                                   \
    {R1, (R2) * (R3), T, NT, (size_t)(T) * ((R1) * (R2) * (R3) + 1) * 8,                 \
     &launch_pass3<R1, R2, R3, NT, T>, &launch_pass3_grouped<R1, R2, R3, NT, T>, VK_THREE_STAGE, R3}

}  // namespace gl
