/*
 * Oracle: Goldilocks NTT plans (naive, Cooley-Tukey, Good-Thomas, Rader).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates ntt/src/ntt/{mod,naive,cooley_tukey,good_thomas,rader}.rs and the
 * size rules of ntt/src/divisors.rs.  The reference's in-place permutations
 * (ntt/src/permute/) are replaced by out-of-place index maps with the same
 * meaning.  Two reference defects are not reproduced, because the contract
 * its own tests assert (ntt/src/ntt/mod.rs:128-139) is the exact DFT:
 *   - cooley_tukey.rs:126 starts the on-the-fly twiddles of chunk 0 at row
 *     a_work_size/a instead of row 1 (wrong output for n > 2^15);
 *   - permute/gcd.rs:76-77 races in the parallel Gcd transpose.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static int g_threads = 1;
void or_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

/* ---- naive.rs ---- */

/* naive.rs:3-22 */
int or_naive_ntt(uint64_t *v, size_t n) {
    if (n == 0) return 0;
    uint64_t root = or_root(n);
    if (!root) return -1;
    uint64_t *copy = (uint64_t *)malloc(n * sizeof(uint64_t));
    memcpy(copy, v, n * sizeof(uint64_t));
    uint64_t wi = 1;
    for (size_t i = 0; i < n; i++) {
        uint64_t sum = 0, wij = 1;
        for (size_t j = 0; j < n; j++) {
            sum = or_add(sum, or_mul(copy[j], wij));
            wij = or_mul(wij, wi);
        }
        v[i] = sum;
        wi = or_mul(wi, root);
    }
    free(copy);
    return 0;
}

/* naive.rs:24-49 */
int or_naive_intt(uint64_t *v, size_t n) {
    if (n == 0) return 0;
    uint64_t root = or_root(n);
    if (!root) return -1;
    uint64_t scale = or_inv(n % OR_P);
    for (size_t i = 0; i < n; i++) v[i] = or_mul(v[i], scale);
    uint64_t iroot = or_pow(root, n - 1);
    uint64_t *copy = (uint64_t *)malloc(n * sizeof(uint64_t));
    memcpy(copy, v, n * sizeof(uint64_t));
    uint64_t wi = 1;
    for (size_t i = 0; i < n; i++) {
        uint64_t sum = 0, wij = 1;
        for (size_t j = 0; j < n; j++) {
            sum = or_add(sum, or_mul(copy[j], wij));
            wij = or_mul(wij, wi);
        }
        v[i] = sum;
        wi = or_mul(wi, iroot);
    }
    free(copy);
    return 0;
}

uint64_t or_eval_dft(const uint64_t *x, size_t n, uint64_t k) {
    uint64_t w = or_pow(or_root(n), k % n), wj = 1, sum = 0;
    for (size_t j = 0; j < n; j++) {
        sum = or_add(sum, or_mul(x[j], wj));
        wj = or_mul(wj, w);
    }
    return sum;
}

/* ---- plans (mod.rs:48-78) ---- */

enum { K_NOP, K_TWO, K_THREE, K_CT, K_GT, K_RADER };

typedef struct Plan Plan;
struct Plan {
    size_t n;
    int kind;
    size_t a, b;            /* CT: n = a*b; GT: n1 = a, n2 = b */
    Plan *inner_a, *inner_b;
    uint64_t root;
    uint64_t *twiddles;     /* CT, n <= 2^15: twiddles[j*a + i] = w^(i*j) */
    int shift_twiddles;     /* CT with n | 384 and n <= 128: twiddles are shifts */
    size_t k3, k4;          /* GT output map (good_thomas.rs:41-42) */
    size_t gi, gk;          /* Rader generators (rader.rs:21-29) */
    uint64_t *rader_tw;     /* Rader: transformed twiddles (rader.rs:41-48) */
};

static pthread_mutex_t g_cache_lock = PTHREAD_MUTEX_INITIALIZER;
static Plan *g_cache[4096];
static size_t g_cache_n = 0;

static size_t gcd_sz(size_t a, size_t b) {
    while (b) { size_t t = a % b; a = b; b = t; }
    return a;
}

/* utils.rs:55-65 */
static size_t modinv_sz(size_t a, size_t m) {
    long long r0 = (long long)a, r1 = (long long)m, s0 = 1, s1 = 0;
    while (r1) {
        long long q = r0 / r1, t;
        t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    long long x = s0 % (long long)m;
    if (x < 0) x += (long long)m;
    return (size_t)x;
}

static Plan *strategy(size_t n);

static Plan *build(size_t n) {
    Plan *p = (Plan *)calloc(1, sizeof(Plan));
    p->n = n;
    if (n <= 1) {
        p->kind = K_NOP;
    } else if (n == 2) {
        p->kind = K_TWO;
    } else if (n == 3) {
        p->kind = K_THREE;   /* codegen.rs naive_3 */
    } else if (n == 5 || n == 17 || n == 257 || n == 65537) {
        /* rader.rs:19-57 */
        p->kind = K_RADER;
        switch (n) {
            case 5: p->gi = 2; p->gk = 3; break;
            case 17: p->gi = 3; p->gk = 6; break;
            case 257: p->gi = 3; p->gk = 86; break;
            default: p->gi = 3; p->gk = 21846; break;
        }
        p->inner_a = strategy(n - 1);
        p->rader_tw = (uint64_t *)malloc((n - 1) * sizeof(uint64_t));
        uint64_t root = or_root(n), scale = or_inv((uint64_t)(n - 1));
        size_t pi = 1;
        for (size_t i = 0; i < n - 1; i++) {
            p->rader_tw[i] = or_mul(or_pow(root, pi), scale);
            pi = (pi * p->gi) % n;
        }
        or_ntt_batch(p->rader_tw, n - 1, 1, 0);
    } else {
        size_t a = (size_t)or_split(n), b = n / a;
        p->a = a;
        p->b = b;
        p->inner_a = strategy(a);
        p->inner_b = strategy(b);
        if (gcd_sz(a, b) == 1) {
            /* good_thomas.rs:26-96 with (k1,k2) = (n2,n1) */
            p->kind = K_GT;
            p->k3 = modinv_sz(b % a, a) * b;
            p->k4 = modinv_sz(a % b, b) * a;
        } else {
            /* cooley_tukey.rs:25-72 */
            p->kind = K_CT;
            p->root = or_root(n);
            if (n <= 128 && 384 % n == 0) {
                /* codegen.rs:21-36,167-195: w_n^e = w_384^(e*384/n) */
                p->shift_twiddles = 1;
            } else if (n <= (1u << 15)) {
                p->twiddles = (uint64_t *)malloc(n * sizeof(uint64_t));
                uint64_t wb = 1;
                for (size_t j = 0; j < b; j++) {
                    uint64_t wab = 1;
                    for (size_t i = 0; i < a; i++) {
                        p->twiddles[j * a + i] = wab;
                        wab = or_mul(wab, wb);
                    }
                    wb = or_mul(wb, p->root);
                }
            }
        }
    }
    return p;
}

/* mod.rs:48-78: cached per size.  Callers have checked or_is_divisor(n). */
static Plan *strategy(size_t n) {
    pthread_mutex_lock(&g_cache_lock);
    for (size_t i = 0; i < g_cache_n; i++)
        if (g_cache[i]->n == n) {
            Plan *p = g_cache[i];
            pthread_mutex_unlock(&g_cache_lock);
            return p;
        }
    pthread_mutex_unlock(&g_cache_lock);
    Plan *p = build(n);
    pthread_mutex_lock(&g_cache_lock);
    for (size_t i = 0; i < g_cache_n; i++)
        if (g_cache[i]->n == n) {  /* lost a race: keep the first (plans leak by design) */
            Plan *q = g_cache[i];
            pthread_mutex_unlock(&g_cache_lock);
            return q;
        }
    if (g_cache_n < 4096) g_cache[g_cache_n++] = p;
    pthread_mutex_unlock(&g_cache_lock);
    return p;
}

static void exec(const Plan *p, uint64_t *v);

static uint64_t *scratch(size_t n, uint64_t *stack, size_t cap) {
    return n <= cap ? stack : (uint64_t *)malloc(n * sizeof(uint64_t));
}
static void release(uint64_t *s, uint64_t *stack) {
    if (s != stack) free(s);
}

/* Parallel loop for the top level of one large transform: the reference's
 * ntt-par runs the inner NTTs of its Cooley-Tukey step as rayon par_chunks
 * (cooley_tukey.rs:100-112); here g_threads pthreads split the index range.
 * Nested levels stay serial. */
static __thread int tl_in_parallel;

typedef struct {
    void (*fn)(void *ctx, size_t i);
    void *ctx;
    size_t begin, end;
} Range;

static void *run_range(void *arg) {
    Range *r = (Range *)arg;
    tl_in_parallel = 1;
    for (size_t i = r->begin; i < r->end; i++) r->fn(r->ctx, i);
    return NULL;
}

static void par_for(size_t count, size_t work_per_item, void (*fn)(void *, size_t), void *ctx);

void or_par_for(size_t count, size_t work_per_item, void (*fn)(void *, size_t), void *ctx) {
    par_for(count, work_per_item, fn, ctx);
}

static void par_for(size_t count, size_t work_per_item, void (*fn)(void *, size_t), void *ctx) {
    int threads = g_threads;
    if (threads > 256) threads = 256;
    if ((size_t)threads > count) threads = (int)count;
    if (threads <= 1 || tl_in_parallel || count * work_per_item < (1u << 15)) {
        for (size_t i = 0; i < count; i++) fn(ctx, i);
        return;
    }
    pthread_t tid[256];
    Range r[256];
    for (int t = 0; t < threads; t++) {
        r[t].fn = fn;
        r[t].ctx = ctx;
        r[t].begin = count * (size_t)t / (size_t)threads;
        r[t].end = count * (size_t)(t + 1) / (size_t)threads;
        pthread_create(&tid[t], NULL, run_range, &r[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* rows x cols -> cols x rows, in 16 x 16 blocks; block rows in parallel */
typedef struct {
    const uint64_t *src;
    uint64_t *dst;
    size_t rows, cols;
} TrCtx;

static void transpose_rows(void *c, size_t blk) {
    const TrCtx *x = (const TrCtx *)c;
    const size_t B = 16, i0 = blk * B;
    size_t i1 = i0 + B < x->rows ? i0 + B : x->rows;
    for (size_t j0 = 0; j0 < x->cols; j0 += B) {
        size_t j1 = j0 + B < x->cols ? j0 + B : x->cols;
        for (size_t i = i0; i < i1; i++)
            for (size_t j = j0; j < j1; j++) x->dst[j * x->rows + i] = x->src[i * x->cols + j];
    }
}

static void transpose(const uint64_t *src, uint64_t *dst, size_t rows, size_t cols) {
    TrCtx c = {src, dst, rows, cols};
    par_for((rows + 15) / 16, 16 * cols, transpose_rows, &c);
}

typedef struct {
    const Plan *p;
    uint64_t *v, *t;
} CtCtx;

static void ct_cols(void *c, size_t j) {
    CtCtx *x = (CtCtx *)c;
    exec(x->p->inner_a, x->t + j * x->p->a);
}
static void ct_rows(void *c, size_t i) {
    CtCtx *x = (CtCtx *)c;
    exec(x->p->inner_b, x->v + i * x->p->b);
}
static void ct_twiddle_row(void *c, size_t j) {
    CtCtx *x = (CtCtx *)c;
    const Plan *p = x->p;
    size_t a = p->a, n = p->n;
    if (j == 0) return;
    if (p->shift_twiddles) {
        size_t step = 384 / n;
        for (size_t i = 1; i < a; i++) x->t[j * a + i] = or_root_384(x->t[j * a + i], (step * ((i * j) % n)) % 384);
    } else if (p->twiddles) {
        for (size_t i = 1; i < a; i++) x->t[j * a + i] = or_mul(x->t[j * a + i], p->twiddles[j * a + i]);
    } else {
        /* row j multiplies column i by w^(i*j) */
        uint64_t wcol = or_pow(p->root, j), wrow = wcol;
        for (size_t i = 1; i < a; i++) {
            x->t[j * a + i] = or_mul(x->t[j * a + i], wrow);
            wrow = or_mul(wrow, wcol);
        }
    }
}

/* cooley_tukey.rs:80-148 (six-step: T(a,b), b x DFT_a, twiddle, T(b,a),
 * a x DFT_b, T(a,b)), with twiddle rows starting at 1 for every chunk. */
static void exec_ct(const Plan *p, uint64_t *v) {
    size_t a = p->a, b = p->b, n = p->n;
    uint64_t stack[1024];
    uint64_t *t = scratch(n, stack, 1024);
    CtCtx cx = {p, v, t};
    transpose(v, t, a, b);
    par_for(b, a, ct_cols, &cx);
    par_for(b, a, ct_twiddle_row, &cx);
    transpose(t, v, b, a);
    par_for(a, b, ct_rows, &cx);
    transpose(v, t, a, b);
    memcpy(v, t, n * sizeof(uint64_t));
    release(t, stack);
}

/* good_thomas.rs:104-126: CRT gather, n1 x DFT_n2, transpose, n2 x DFT_n1,
 * CRT scatter. */
static void exec_gt(const Plan *p, uint64_t *v) {
    size_t n1 = p->a, n2 = p->b, n = p->n;
    uint64_t stack[1024];
    uint64_t *t = scratch(n, stack, 1024);
    for (size_t i1 = 0; i1 < n1; i1++)
        for (size_t i2 = 0; i2 < n2; i2++) t[i1 * n2 + i2] = v[(i1 * n2 + i2 * n1) % n];
    for (size_t i1 = 0; i1 < n1; i1++) exec(p->inner_b, t + i1 * n2);
    transpose(t, v, n1, n2);
    for (size_t k2 = 0; k2 < n2; k2++) exec(p->inner_a, v + k2 * n1);
    for (size_t k2 = 0; k2 < n2; k2++)
        for (size_t k1 = 0; k1 < n1; k1++) t[(k1 * p->k3 + k2 * p->k4) % n] = v[k2 * n1 + k1];
    memcpy(v, t, n * sizeof(uint64_t));
    release(t, stack);
}

/* rader.rs:65-93: cyclic convolution of x[gk^q] with w^(gi^q) through
 * DFT_(P-1); outputs land at X[gi^m]. */
static void exec_rader(const Plan *p, uint64_t *v) {
    size_t P = p->n, m = P - 1;
    uint64_t stack[1024];
    uint64_t *c = scratch(m, stack, 1024);
    size_t idx = 1;
    for (size_t q = 0; q < m; q++) {
        c[q] = v[idx];
        idx = (idx * p->gk) % P;
    }
    exec(p->inner_a, c);
    uint64_t x0 = v[0];
    v[0] = or_add(x0, c[0]);
    for (size_t q = 0; q < m; q++) c[q] = or_mul(c[q], p->rader_tw[q]);
    c[0] = or_add(c[0], x0);
    for (size_t i = 1, j = m - 1; i < j; i++, j--) {
        uint64_t t = c[i]; c[i] = c[j]; c[j] = t;
    }
    exec(p->inner_a, c);
    idx = 1;
    for (size_t q = 0; q < m; q++) {
        v[idx] = c[q];
        idx = (idx * p->gi) % P;
    }
    release(c, stack);
}

static void exec(const Plan *p, uint64_t *v) {
    switch (p->kind) {
        case K_NOP: return;
        case K_TWO: {
            uint64_t a = v[0], b = v[1];
            v[0] = or_add(a, b);
            v[1] = or_sub(a, b);
            return;
        }
        case K_THREE: {
            /* w_3 = 2^64 (codegen.rs:21-36) */
            uint64_t a = v[0], b = v[1], c = v[2];
            v[0] = or_add(or_add(a, b), c);
            v[1] = or_add(or_add(a, or_shift(b, 64)), or_shift(c, 128));
            v[2] = or_add(or_add(a, or_shift(b, 128)), or_shift(c, 64));
            return;
        }
        case K_CT: exec_ct(p, v); return;
        case K_GT: exec_gt(p, v); return;
        case K_RADER: exec_rader(p, v); return;
    }
}

typedef struct {
    const Plan *p;
    uint64_t *v;
    size_t n, begin, end;
} Job;

static void *run_job(void *arg) {
    Job *j = (Job *)arg;
    for (size_t t = j->begin; t < j->end; t++) exec(j->p, j->v + t * j->n);
    return NULL;
}

/* intt = scale by 1/n, reverse [1..], forward (mod.rs:104-120). */
int or_ntt_batch(uint64_t *v, size_t n, size_t count, int inverse) {
    if (inverse && n <= 1) return or_is_divisor(n) || n == 0 ? 0 : -1;
    if (!or_is_divisor(n)) return -1;
    const Plan *p = strategy(n);
    if (inverse) {
        uint64_t scale = or_inv(n % OR_P);
        for (size_t t = 0; t < count; t++) {
            uint64_t *x = v + t * n;
            for (size_t i = 0; i < n; i++) x[i] = or_mul(x[i], scale);
            for (size_t i = 1, j = n - 1; i < j; i++, j--) {
                uint64_t s = x[i]; x[i] = x[j]; x[j] = s;
            }
        }
    }
    int threads = g_threads;
    if (threads > 1 && count > 1) {
        if ((size_t)threads > count) threads = (int)count;
        pthread_t tid[256];
        Job jobs[256];
        if (threads > 256) threads = 256;
        for (int i = 0; i < threads; i++) {
            jobs[i].p = p; jobs[i].v = v; jobs[i].n = n;
            jobs[i].begin = count * (size_t)i / (size_t)threads;
            jobs[i].end = count * (size_t)(i + 1) / (size_t)threads;
            pthread_create(&tid[i], NULL, run_job, &jobs[i]);
        }
        for (int i = 0; i < threads; i++) pthread_join(tid[i], NULL);
    } else {
        for (size_t t = 0; t < count; t++) exec(p, v + t * n);
    }
    return 0;
}

int or_ntt(uint64_t *v, size_t n) { return or_ntt_batch(v, n, 1, 0); }
int or_intt(uint64_t *v, size_t n) { return or_ntt_batch(v, n, 1, 1); }

/* convolve.rs:13-20 (circular): a <- intt(ntt(a) * b_ntt), b already in NTT
 * form, the product being the Field (Montgomery) product of the raw words. */
int or_convolve(uint64_t *a, const uint64_t *b_ntt, size_t n) {
    if (or_ntt(a, n)) return -1;
    for (size_t i = 0; i < n; i++) a[i] = or_mont_mul(a[i], b_ntt[i]);
    return or_intt(a, n);
}
