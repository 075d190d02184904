/*
 * gridref.c — CPU restatement of the reference ring-arithmetic path (ORACLE).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product never calls it.
 *
 * Arithmetic follows `PrimeField<Q, u64>` (algebra/src/arith/prime.rs) with
 * Montgomery radix R = 2^64; transforms follow algebra/src/arith/ntt.rs and
 * algebra/src/poly/ntt.rs step by step (bit-reverse + DIT forward, DIF inverse
 * with n^-1 folded into the last stage, twist by psi^i).
 */
#include "gridref.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Field: PrimeField<Q, u64>                                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t q;
    uint64_t q_inv;  /* -q^{-1} mod 2^64            prime.rs:82-91 */
    uint64_t r1;     /* 2^64 mod q  (Montgomery one) prime.rs:59-66 */
    uint64_t r2;     /* 2^128 mod q                  prime.rs:74-78 */
    int wide;        /* q >= 2^63 -> montgomery_reduce_wide, prime.rs:141-146 */
} field_t;

static void field_init(field_t *f, uint64_t q) {
    f->q = q;
    /* Newton iteration for q^{-1} mod 2^64, then negate (prime.rs:82-91). */
    uint64_t inv = 1;
    for (int i = 0; i < 6; i++) inv *= 2 - q * inv;
    f->q_inv = (uint64_t)0 - inv;
    f->r1 = (uint64_t)(((u128)1 << 64) % q);
    f->r2 = (uint64_t)(((u128)f->r1 * f->r1) % q);
    f->wide = q >= ((uint64_t)1 << 63);
}

/* add_raw_words, prime.rs:110-127 */
static inline uint64_t f_add(const field_t *f, uint64_t a, uint64_t b) {
    if (!f->wide) {
        uint64_t s = a + b;
        return s >= f->q ? s - f->q : s;
    }
    uint64_t s = a + b;
    int carry = s < a;
    return (carry || s >= f->q) ? s - f->q : s;
}

/* sub_raw_words, prime.rs:130-137 */
static inline uint64_t f_sub(const field_t *f, uint64_t a, uint64_t b) {
    return a >= b ? a - b : f->q - b + a;
}

/* montgomery_reduce / montgomery_reduce_wide, prime.rs:141-183 */
static inline uint64_t f_reduce(const field_t *f, u128 t) {
    uint64_t t_lo = (uint64_t)t;
    uint64_t m = t_lo * f->q_inv;
    u128 mq = (u128)m * f->q;
    if (!f->wide) {
        uint64_t res = (uint64_t)((t + mq) >> 64);
        return res >= f->q ? res - f->q : res;
    }
    uint64_t t_hi = (uint64_t)(t >> 64);
    uint64_t p_lo = (uint64_t)mq, p_hi = (uint64_t)(mq >> 64);
    uint64_t lo_sum = t_lo + p_lo;
    int carry_lo = lo_sum < t_lo;
    uint64_t hi = t_hi + p_hi;
    int c1 = hi < t_hi;
    uint64_t hi2 = hi + (uint64_t)carry_lo;
    int c2 = hi2 < hi;
    if (c1 || c2) return hi2 - f->q; /* wrapping add of -q */
    return hi2 >= f->q ? hi2 - f->q : hi2;
}

/* mul_raw_words, prime.rs:224-226 */
static inline uint64_t f_mul(const field_t *f, uint64_t a, uint64_t b) {
    return f_reduce(f, (u128)a * b);
}

/* to_montgomery / from_montgomery, prime.rs:187-202 */
static inline uint64_t f_to(const field_t *f, uint64_t x) {
    return f_reduce(f, (u128)(x % f->q) * f->r2);
}
static inline uint64_t f_from(const field_t *f, uint64_t x) { return f_reduce(f, (u128)x); }

/* Ring::pow (square and multiply), Montgomery in / out */
static uint64_t f_pow(const field_t *f, uint64_t base, uint64_t e) {
    uint64_t acc = f->r1;
    while (e) {
        if (e & 1) acc = f_mul(f, acc, base);
        base = f_mul(f, base, base);
        e >>= 1;
    }
    return acc;
}

/* Field::inv = self^(q-2), prime.rs:556-562 */
static uint64_t f_inv(const field_t *f, uint64_t a) { return f_pow(f, a, f->q - 2); }

/* ---- canonical scalar API ------------------------------------------------ */

uint64_t gr_add(uint64_t q, uint64_t a, uint64_t b) {
    field_t f; field_init(&f, q);
    return f_from(&f, f_add(&f, f_to(&f, a), f_to(&f, b)));
}
uint64_t gr_sub(uint64_t q, uint64_t a, uint64_t b) {
    field_t f; field_init(&f, q);
    return f_from(&f, f_sub(&f, f_to(&f, a), f_to(&f, b)));
}
uint64_t gr_mul(uint64_t q, uint64_t a, uint64_t b) {
    field_t f; field_init(&f, q);
    return f_from(&f, f_mul(&f, f_to(&f, a), f_to(&f, b)));
}
uint64_t gr_pow(uint64_t q, uint64_t a, uint64_t e) {
    field_t f; field_init(&f, q);
    return f_from(&f, f_pow(&f, f_to(&f, a), e));
}
uint64_t gr_inv(uint64_t q, uint64_t a) {
    field_t f; field_init(&f, q);
    return f_from(&f, f_inv(&f, f_to(&f, a)));
}
uint64_t gr_to_montgomery(uint64_t q, uint64_t a) { field_t f; field_init(&f, q); return f_to(&f, a); }
uint64_t gr_from_montgomery(uint64_t q, uint64_t a) { field_t f; field_init(&f, q); return f_from(&f, a); }
uint64_t gr_q_inv_neg(uint64_t q) { field_t f; field_init(&f, q); return f.q_inv; }

/* ---- Ring slice hooks of PrimeField, prime.rs:433-491 -------------------- */
/* Canonical words in and out; in between they live in Montgomery form as the
 * reference stores them. op 0: add_assign_slice (montgomery.rs add_assign),
 * 1: sub_assign_slice, 2: scalar_mul_slice (dst *= scalar), 3:
 * add_assign_scaled_slice: slices longer than CHUNK = 64 go chunk by chunk,
 * the chunk's products scaled first and then added (prime.rs:449-465). */
int gr_slice_op(uint64_t q, int op, uint64_t *dst, const uint64_t *src, uint64_t scalar,
                size_t len) {
    if (op < 0 || op > 3) return GR_E_INVALID_ARGUMENT;
    field_t f; field_init(&f, q);
    const uint64_t c = f_to(&f, scalar);
    enum { CHUNK = 64 };
    for (size_t base = 0; base < len; base += CHUNK) {
        const size_t n = len - base < CHUNK ? len - base : CHUNK;
        uint64_t prod[CHUNK];
        for (size_t i = 0; i < n; i++) {
            const uint64_t d = f_to(&f, dst[base + i]);
            uint64_t r;
            if (op == 0) {
                r = f_add(&f, d, f_to(&f, src[base + i]));
            } else if (op == 1) {
                r = f_sub(&f, d, f_to(&f, src[base + i]));
            } else if (op == 2) {
                r = f_mul(&f, d, c);
            } else {
                prod[i] = f_mul(&f, f_to(&f, src[base + i]), c);
                r = f_add(&f, d, prod[i]);
            }
            dst[base + i] = f_from(&f, r);
        }
    }
    return GR_OK;
}

/* dot_product, prime.rs:467-491: chunks of 64 pointwise products, each added
 * into one running sum in order. */
uint64_t gr_dot_product(uint64_t q, const uint64_t *a, const uint64_t *b, size_t len) {
    field_t f; field_init(&f, q);
    uint64_t acc = 0; /* Montgomery zero */
    for (size_t i = 0; i < len; i++)
        acc = f_add(&f, acc, f_mul(&f, f_to(&f, a[i]), f_to(&f, b[i])));
    return f_from(&f, acc);
}

/* ------------------------------------------------------------------------ */
/* Generator and roots of unity                                              */
/* ------------------------------------------------------------------------ */

#define GOLDILOCKS 0xffffffff00000001ULL

/* find_generator, prime.rs:645-687: Goldilocks is pinned to 7; otherwise the
 * smallest g >= 2 with g^((q-1)/p) != 1 for every prime p | q-1, where the
 * factors come from trial division. Returns the generator canonically. */
int gr_find_generator(uint64_t q, uint64_t *g_out) {
    if (q < 3 || (q & 1) == 0) return GR_E_INVALID_ARGUMENT;
    if (q == GOLDILOCKS) { *g_out = 7; return GR_OK; }
    field_t f; field_init(&f, q);
    uint64_t factors[64];
    int nf = 0;
    uint64_t n = q - 1;
    for (uint64_t d = 2; d <= n / d; d++) {
        if (n % d == 0) {
            factors[nf++] = d;
            while (n % d == 0) n /= d;
        }
    }
    if (n > 1) factors[nf++] = n;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t gm = f_to(&f, g);
        int ok = 1;
        for (int i = 0; i < nf; i++) {
            if (f_pow(&f, gm, (q - 1) / factors[i]) == f.r1) { ok = 0; break; }
        }
        if (ok) { *g_out = g; return GR_OK; }
    }
    return GR_E_UNSUPPORTED_SIZE;
}

/* root_of_unity(n), prime.rs:567-585; result in Montgomery form. */
static int f_root_of_unity(const field_t *f, uint64_t n, uint64_t *root) {
    if (n == 0 || (f->q - 1) % n != 0) return GR_E_UNSUPPORTED_SIZE;
    uint64_t g;
    if (gr_find_generator(f->q, &g) != GR_OK) return GR_E_UNSUPPORTED_SIZE;
    uint64_t r = f_pow(f, f_to(f, g), (f->q - 1) / n);
    if (f_pow(f, r, n) != f->r1) return GR_E_UNSUPPORTED_SIZE;
    if (n > 1 && f_pow(f, r, n / 2) == f->r1) return GR_E_UNSUPPORTED_SIZE;
    *root = r;
    return GR_OK;
}

int gr_root_of_unity(uint64_t q, uint64_t n, uint64_t *root_out) {
    field_t f; field_init(&f, q);
    uint64_t r;
    int st = f_root_of_unity(&f, n, &r);
    if (st == GR_OK) *root_out = f_from(&f, r);
    return st;
}

/* max_ntt_size, prime.rs:597-606 */
uint64_t gr_max_ntt_size(uint64_t q) {
    uint64_t v = q - 1, m = 1;
    while ((v & 1) == 0 && v) { v >>= 1; m <<= 1; }
    return m;
}

/* ------------------------------------------------------------------------ */
/* NttPlan (arith/ntt.rs:52-153) and TwistedNttPlan (poly/ntt.rs:26-153)     */
/* ------------------------------------------------------------------------ */

typedef struct {
    field_t f;
    uint64_t len;
    int stages;
    uint64_t *fwd;   /* stage s (len 2^(s+1)) twiddles at fwd + (2^s - 1) */
    uint64_t *inv;   /* inverse stage s (len n>>s) twiddles at inv + off[s] */
    uint64_t inv_scale;
    uint64_t *twists, *untwists; /* psi^i, psi^-i (Montgomery) */
} plan_t;

static int is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

static void plan_free(plan_t *p) {
    free(p->fwd); free(p->inv); free(p->twists); free(p->untwists);
    memset(p, 0, sizeof *p);
}

/* NttPlan::build, arith/ntt.rs:124-153; twisted part poly/ntt.rs:140-153. */
static int plan_build(plan_t *p, uint64_t q, uint64_t len, int twisted) {
    memset(p, 0, sizeof *p);
    field_init(&p->f, q);
    if (!is_pow2(len)) return GR_E_LENGTH_NOT_POW2;
    const field_t *f = &p->f;
    uint64_t root;
    if (f_root_of_unity(f, len, &root) != GR_OK) return GR_E_UNSUPPORTED_SIZE;
    uint64_t inv_root = f_inv(f, root);
    p->len = len;
    int stages = 0;
    while (((uint64_t)1 << stages) < len) stages++;
    p->stages = stages;
    p->fwd = (uint64_t *)malloc(sizeof(uint64_t) * (len ? len : 1));
    p->inv = (uint64_t *)malloc(sizeof(uint64_t) * (len ? len : 1));
    /* forward: stage_len = 2, 4, ..., len; w_len = root^(len/stage_len) */
    uint64_t off = 0;
    for (int s = 0; s < stages; s++) {
        uint64_t stage_len = (uint64_t)2 << s, half = stage_len / 2;
        uint64_t w = f_pow(f, root, len / stage_len), cur = f->r1;
        for (uint64_t j = 0; j < half; j++) { p->fwd[off + j] = cur; cur = f_mul(f, cur, w); }
        off += half;
    }
    /* inverse: stage_len = len, len/2, ..., 2; w_len = inv_root^(len/stage_len) */
    off = 0;
    for (int s = 0; s < stages; s++) {
        uint64_t stage_len = len >> s, half = stage_len / 2;
        uint64_t w = f_pow(f, inv_root, len / stage_len), cur = f->r1;
        for (uint64_t j = 0; j < half; j++) { p->inv[off + j] = cur; cur = f_mul(f, cur, w); }
        off += half;
    }
    /* inverse_ntt_scale = len^-1, prime.rs:592-596 */
    p->inv_scale = f_inv(f, f_to(f, len));
    if (twisted) {
        uint64_t psi;
        if (f_root_of_unity(f, 2 * len, &psi) != GR_OK) { plan_free(p); return GR_E_UNSUPPORTED_TWIST; }
        uint64_t ipsi = f_inv(f, psi);
        p->twists = (uint64_t *)malloc(sizeof(uint64_t) * len);
        p->untwists = (uint64_t *)malloc(sizeof(uint64_t) * len);
        uint64_t a = f->r1, b = f->r1;
        for (uint64_t i = 0; i < len; i++) {
            p->twists[i] = a; p->untwists[i] = b;
            a = f_mul(f, a, psi); b = f_mul(f, b, ipsi);
        }
    }
    return GR_OK;
}

/* bit_reverse_permute, arith/ntt.rs:180-194 */
static void bit_reverse(uint64_t *v, uint64_t n) {
    uint64_t j = 0;
    for (uint64_t i = 1; i < n; i++) {
        uint64_t bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) { uint64_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}

/* scalar_ntt_forward_with_plan, arith/ntt.rs:292-312 (butterfly :205-235) */
static void plan_forward(const plan_t *p, uint64_t *v) {
    const field_t *f = &p->f;
    uint64_t n = p->len;
    bit_reverse(v, n);
    uint64_t off = 0;
    for (int s = 0; s < p->stages; s++) {
        uint64_t len = (uint64_t)2 << s, half = len / 2;
        const uint64_t *tw = p->fwd + off;
        for (uint64_t start = 0; start < n; start += len) {
            uint64_t *e = v + start, *o = v + start + half;
            for (uint64_t i = 0; i < half; i++) {
                uint64_t t = f_mul(f, o[i], tw[i]);
                o[i] = f_sub(f, e[i], t);
                e[i] = f_add(f, e[i], t);
            }
        }
        off += half;
    }
}

/* scalar_ntt_inverse_with_plan, arith/ntt.rs:314-335 (butterfly :237-283) */
static void plan_inverse(const plan_t *p, uint64_t *v) {
    const field_t *f = &p->f;
    uint64_t n = p->len;
    uint64_t off = 0;
    for (int s = 0; s < p->stages; s++) {
        uint64_t len = n >> s, half = len / 2;
        int final_stage = (len == 2);
        const uint64_t *tw = p->inv + off;
        for (uint64_t start = 0; start < n; start += len) {
            uint64_t *e = v + start, *o = v + start + half;
            for (uint64_t i = 0; i < half; i++) {
                uint64_t diff = f_sub(f, e[i], o[i]);
                uint64_t sum = f_add(f, e[i], o[i]);
                uint64_t od = f_mul(f, diff, tw[i]);
                if (final_stage) { sum = f_mul(f, sum, p->inv_scale); od = f_mul(f, od, p->inv_scale); }
                e[i] = sum; o[i] = od;
            }
        }
        off += half;
    }
    bit_reverse(v, n);
}

/* TwistedNttPlan::forward_in_place_trusted, poly/ntt.rs:75-87 */
static void plan_twisted_forward(const plan_t *p, uint64_t *v) {
    for (uint64_t i = 0; i < p->len; i++) v[i] = f_mul(&p->f, v[i], p->twists[i]);
    plan_forward(p, v);
}

/* TwistedNttPlan::inverse_in_place_trusted, poly/ntt.rs:90-104 */
static void plan_twisted_inverse(const plan_t *p, uint64_t *v) {
    plan_inverse(p, v);
    for (uint64_t i = 0; i < p->len; i++) v[i] = f_mul(&p->f, v[i], p->untwists[i]);
}

static void to_mont_vec(const field_t *f, uint64_t *dst, const uint64_t *src, size_t n) {
    for (size_t i = 0; i < n; i++) dst[i] = f_to(f, src[i]);
}
static void from_mont_vec(const field_t *f, uint64_t *v, size_t n) {
    for (size_t i = 0; i < n; i++) v[i] = f_from(f, v[i]);
}

typedef void (*xform_fn)(const plan_t *, uint64_t *);

static int run_xform(uint64_t q, uint64_t len, uint64_t *values, size_t count, int twisted,
                     xform_fn fn) {
    if (!values && count) return GR_E_INVALID_ARGUMENT;
    plan_t p;
    int st = plan_build(&p, q, len, twisted);
    if (st != GR_OK) return st;
    for (size_t c = 0; c < count; c++) {
        uint64_t *v = values + c * len;
        to_mont_vec(&p.f, v, v, len);
        fn(&p, v);
        from_mont_vec(&p.f, v, len);
    }
    plan_free(&p);
    return GR_OK;
}

int gr_ntt_forward(uint64_t q, uint64_t len, uint64_t *v, size_t c) { return run_xform(q, len, v, c, 0, plan_forward); }
int gr_ntt_inverse(uint64_t q, uint64_t len, uint64_t *v, size_t c) { return run_xform(q, len, v, c, 0, plan_inverse); }
int gr_twisted_forward(uint64_t q, uint64_t len, uint64_t *v, size_t c) { return run_xform(q, len, v, c, 1, plan_twisted_forward); }
int gr_twisted_inverse(uint64_t q, uint64_t len, uint64_t *v, size_t c) { return run_xform(q, len, v, c, 1, plan_twisted_inverse); }

/* TwistedNttPlan::multiply_in_place, poly/ntt.rs:110-126 */
static void plan_multiply(const plan_t *p, uint64_t *lhs, uint64_t *rhs) {
    plan_twisted_forward(p, lhs);
    plan_twisted_forward(p, rhs);
    for (uint64_t i = 0; i < p->len; i++) lhs[i] = f_mul(&p->f, lhs[i], rhs[i]);
    plan_twisted_inverse(p, lhs);
}

int gr_negacyclic_mul(uint64_t q, uint64_t len, const uint64_t *a, const uint64_t *b,
                      uint64_t *out, size_t count) {
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * 2 * len);
    for (size_t c = 0; c < count; c++) {
        to_mont_vec(&p.f, tmp, a + c * len, len);
        to_mont_vec(&p.f, tmp + len, b + c * len, len);
        plan_multiply(&p, tmp, tmp + len);
        from_mont_vec(&p.f, tmp, len);
        memcpy(out + c * len, tmp, sizeof(uint64_t) * len);
    }
    free(tmp);
    plan_free(&p);
    return GR_OK;
}

/* naive_negacyclic_mul_coeffs, poly/ring.rs:60-77 */
int gr_negacyclic_mul_naive(uint64_t q, uint64_t len, const uint64_t *a, const uint64_t *b,
                            uint64_t *out, size_t count) {
    field_t f; field_init(&f, q);
    uint64_t *acc = (uint64_t *)malloc(sizeof(uint64_t) * (len ? len : 1));
    for (size_t c = 0; c < count; c++) {
        const uint64_t *x = a + c * len, *y = b + c * len;
        for (uint64_t i = 0; i < len; i++) acc[i] = 0;
        for (uint64_t i = 0; i < len; i++) {
            uint64_t xi = f_to(&f, x[i]);
            for (uint64_t j = 0; j < len; j++) {
                uint64_t prod = f_mul(&f, xi, f_to(&f, y[j]));
                uint64_t idx = i + j;
                if (idx < len) acc[idx] = f_add(&f, acc[idx], prod);
                else acc[idx - len] = f_sub(&f, acc[idx - len], prod);
            }
        }
        for (uint64_t i = 0; i < len; i++) out[c * len + i] = f_from(&f, acc[i]);
    }
    free(acc);
    return GR_OK;
}

/* ------------------------------------------------------------------------ */
/* Linear commitment                                                          */
/* ------------------------------------------------------------------------ */

/* TwistedNttPoly::dot_product, twisted_ntt.rs:378-408: acc[k] += a_j[k]*b_j[k] */
static void ntt_dot(const plan_t *p, const uint64_t *a, const uint64_t *b, uint32_t cols,
                    uint64_t *acc) {
    const field_t *f = &p->f;
    uint64_t len = p->len;
    for (uint64_t k = 0; k < len; k++) acc[k] = 0;
    for (uint32_t j = 0; j < cols; j++)
        for (uint64_t k = 0; k < len; k++)
            acc[k] = f_add(f, acc[k], f_mul(f, a[(size_t)j * len + k], b[(size_t)j * len + k]));
}

/* prepare_poly_vector / prepare_poly_matrix (linear.rs:329-359): copy into
 * Montgomery form and, for coefficient-domain input, twisted-forward each poly. */
static uint64_t *prepare(const plan_t *p, const uint64_t *src, size_t polys, int already_ntt) {
    uint64_t *dst = (uint64_t *)malloc(sizeof(uint64_t) * (polys * p->len + 1));
    to_mont_vec(&p->f, dst, src, polys * p->len);
    if (!already_ntt)
        for (size_t i = 0; i < polys; i++) plan_twisted_forward(p, dst + i * p->len);
    return dst;
}

int gr_commit_prepared(uint64_t q, uint64_t len, uint32_t rows, const uint64_t *a_msg,
                       uint32_t msg_len, const uint64_t *a_open, uint32_t open_len,
                       int key_domain, const uint64_t *message, const uint64_t *randomness,
                       uint64_t *t_out) {
    if (rows == 0 || msg_len == 0 || open_len == 0) return GR_E_INVALID_PARAMETERS;
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    const field_t *f = &p.f;
    uint64_t *am = prepare(&p, a_msg, (size_t)rows * msg_len, key_domain);
    uint64_t *ao = prepare(&p, a_open, (size_t)rows * open_len, key_domain);
    uint64_t *m = prepare(&p, message, msg_len, 0);
    uint64_t *r = prepare(&p, randomness, open_len, 0);
    uint64_t *acc1 = (uint64_t *)malloc(sizeof(uint64_t) * len * 2), *acc2 = acc1 + len;
    /* recompute_linear_commitment_parts_trusted, linear.rs:476-506 */
    for (uint32_t row = 0; row < rows; row++) {
        ntt_dot(&p, am + (size_t)row * msg_len * len, m, msg_len, acc1);
        ntt_dot(&p, ao + (size_t)row * open_len * len, r, open_len, acc2);
        for (uint64_t k = 0; k < len; k++) acc1[k] = f_add(f, acc1[k], acc2[k]);
        /* finish_prepared_vector, linear.rs:361-375 */
        plan_twisted_inverse(&p, acc1);
        for (uint64_t k = 0; k < len; k++) t_out[(size_t)row * len + k] = f_from(f, acc1[k]);
    }
    free(acc1); free(am); free(ao); free(m); free(r);
    plan_free(&p);
    return GR_OK;
}

/* Coefficient-domain Ring::dot_product on CyclotomicPolyRing (arith/ring.rs:113-121):
 * acc += mul_ref(a_j, b_j), each product one twisted multiply (poly/ring.rs:83-101). */
static void coeff_dot(const plan_t *p, const uint64_t *a, const uint64_t *b, uint32_t cols,
                      uint64_t *acc, uint64_t *scratch) {
    uint64_t len = p->len;
    uint64_t *x = scratch, *y = scratch + len;
    for (uint64_t k = 0; k < len; k++) acc[k] = 0;
    for (uint32_t j = 0; j < cols; j++) {
        memcpy(x, a + (size_t)j * len, sizeof(uint64_t) * len);
        memcpy(y, b + (size_t)j * len, sizeof(uint64_t) * len);
        plan_multiply(p, x, y);
        for (uint64_t k = 0; k < len; k++) acc[k] = f_add(&p->f, acc[k], x[k]);
    }
}

int gr_commit_coeff(uint64_t q, uint64_t len, uint32_t rows, const uint64_t *a_msg,
                    uint32_t msg_len, const uint64_t *a_open, uint32_t open_len,
                    const uint64_t *message, const uint64_t *randomness, uint64_t *t_out) {
    if (rows == 0 || msg_len == 0 || open_len == 0) return GR_E_INVALID_PARAMETERS;
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    const field_t *f = &p.f;
    uint64_t *am = prepare(&p, a_msg, (size_t)rows * msg_len, 1);
    uint64_t *ao = prepare(&p, a_open, (size_t)rows * open_len, 1);
    uint64_t *m = prepare(&p, message, msg_len, 1);
    uint64_t *r = prepare(&p, randomness, open_len, 1);
    uint64_t *buf = (uint64_t *)malloc(sizeof(uint64_t) * len * 4);
    uint64_t *acc1 = buf, *acc2 = buf + len, *scratch = buf + 2 * len;
    for (uint32_t row = 0; row < rows; row++) {
        coeff_dot(&p, am + (size_t)row * msg_len * len, m, msg_len, acc1, scratch);
        coeff_dot(&p, ao + (size_t)row * open_len * len, r, open_len, acc2, scratch);
        for (uint64_t k = 0; k < len; k++)
            t_out[(size_t)row * len + k] = f_from(f, f_add(f, acc1[k], acc2[k]));
    }
    free(buf); free(am); free(ao); free(m); free(r);
    plan_free(&p);
    return GR_OK;
}

/* ------------------------------------------------------------------------ */
/* Gram (garbage.rs:76-93)                                                    */
/* ------------------------------------------------------------------------ */

int gr_gram(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *s,
            uint64_t *g_out, int ntt_once) {
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    const field_t *f = &p.f;
    size_t part = (size_t)n * len;
    uint64_t *sm = prepare(&p, s, (size_t)r * n, ntt_once ? 0 : 1);
    uint64_t *buf = (uint64_t *)malloc(sizeof(uint64_t) * len * 3);
    size_t idx = 0;
    for (uint32_t i = 0; i < r; i++) {
        for (uint32_t j = i; j < r; j++) {
            uint64_t *acc = buf;
            if (ntt_once) {
                ntt_dot(&p, sm + i * part, sm + j * part, n, acc);
                plan_twisted_inverse(&p, acc);
            } else {
                coeff_dot(&p, sm + i * part, sm + j * part, n, acc, buf + len);
            }
            for (uint64_t k = 0; k < len; k++) g_out[idx * len + k] = f_from(f, acc[k]);
            idx++;
        }
    }
    free(buf); free(sm);
    plan_free(&p);
    return GR_OK;
}

/* ------------------------------------------------------------------------ */
/* LaBRADOR step around the commitment                                        */
/* ------------------------------------------------------------------------ */

/* decompose_coeff_balanced / centered_mod / decompose_poly_balanced
 * (labrador/src/main_protocol/mod.rs:258-320), flattened as decompose_polys
 * (mod.rs:731-745): out[(poly * limbs + l) * len + k]. */
int gr_decompose_balanced(uint64_t q, uint64_t len, size_t npolys, const uint64_t *in,
                          uint64_t base, uint32_t limbs, uint64_t *out) {
    if (base < 2 || limbs == 0) return GR_E_INVALID_PARAMETERS;
    const __int128 half_q = (__int128)(q / 2), b = (__int128)base, half = b / 2;
    for (size_t p = 0; p < npolys; p++) {
        for (uint64_t k = 0; k < len; k++) {
            __int128 carry = (__int128)in[p * len + k];
            if (carry > half_q) carry -= (__int128)q;
            for (uint32_t l = 0; l < limbs; l++) {
                __int128 r = carry % b;  /* truncating, like i128 % in Rust */
                if (r > half) r -= b;
                else if (r < -half) r += b;
                out[(p * limbs + l) * len + k] = r >= 0 ? (uint64_t)r : q - (uint64_t)(-r);
                carry = (carry - r) / b;
            }
        }
    }
    return GR_OK;
}

/* compute_garbage_h (labrador/src/main_protocol/garbage.rs:101-133):
 * h_ij = (<phi_i, s_j> + <phi_j, s_i>) * 2^-1 for 0 <= i <= j < r, row-major
 * upper triangle; phi and s: r parts x n polys x len, coefficient domain. */
int gr_garbage_h(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *phi,
                 const uint64_t *s, uint64_t *h_out) {
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    const field_t *f = &p.f;
    const size_t part = (size_t)n * len;
    uint64_t *pm = prepare(&p, phi, (size_t)r * n, 1);
    uint64_t *sm = prepare(&p, s, (size_t)r * n, 1);
    uint64_t *buf = (uint64_t *)malloc(sizeof(uint64_t) * len * 4);
    uint64_t *t1 = buf, *t2 = buf + len, *scratch = buf + 2 * len;
    const uint64_t two_inv = f_to(f, q / 2 + 1);
    size_t idx = 0;
    for (uint32_t i = 0; i < r; i++) {
        for (uint32_t j = i; j < r; j++) {
            coeff_dot(&p, pm + i * part, sm + j * part, n, t1, scratch);
            coeff_dot(&p, pm + j * part, sm + i * part, n, t2, scratch);
            for (uint64_t k = 0; k < len; k++)
                h_out[idx * len + k] = f_from(f, f_mul(f, f_add(f, t1[k], t2[k]), two_inv));
            idx++;
        }
    }
    free(buf); free(pm); free(sm);
    plan_free(&p);
    return GR_OK;
}

/* compute_amortized_witness (labrador/src/main_protocol/amortization.rs:61-84):
 * z[k] = sum_i c_i * s_i[k]; c: r polys, s: r parts x n polys, z: n polys. */
int gr_amortize(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *c,
                const uint64_t *s, uint64_t *z_out) {
    plan_t p;
    int st = plan_build(&p, q, len, 1);
    if (st != GR_OK) return st;
    const field_t *f = &p.f;
    uint64_t *cm = prepare(&p, c, r, 1);
    uint64_t *sm = prepare(&p, s, (size_t)r * n, 1);
    uint64_t *buf = (uint64_t *)malloc(sizeof(uint64_t) * len * 3);
    uint64_t *acc = buf, *x = buf + len, *y = buf + 2 * len;
    for (uint32_t k = 0; k < n; k++) {
        for (uint64_t e = 0; e < len; e++) acc[e] = 0;
        for (uint32_t i = 0; i < r; i++) {
            memcpy(x, cm + (size_t)i * len, sizeof(uint64_t) * len);
            memcpy(y, sm + ((size_t)i * n + k) * len, sizeof(uint64_t) * len);
            plan_multiply(&p, x, y);
            for (uint64_t e = 0; e < len; e++) acc[e] = f_add(f, acc[e], x[e]);
        }
        for (uint64_t e = 0; e < len; e++) z_out[(size_t)k * len + e] = f_from(f, acc[e]);
    }
    free(buf); free(cm); free(sm);
    plan_free(&p);
    return GR_OK;
}

/* ------------------------------------------------------------------------ */
/* ChaCha20 (RFC 8439 section 2.3) and the JL projection                      */
/* ------------------------------------------------------------------------ */

static inline uint32_t rotl32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
static inline uint32_t ld32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
#define QR(a, b, c, d)                                   \
    a += b; d ^= a; d = rotl32(d, 16);                   \
    c += d; b ^= c; b = rotl32(b, 12);                   \
    a += b; d ^= a; d = rotl32(d, 8);                    \
    c += d; b ^= c; b = rotl32(b, 7);

void gr_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                       uint8_t out[64]) {
    uint32_t s[16], x[16];
    s[0] = 0x61707865; s[1] = 0x3320646e; s[2] = 0x79622d32; s[3] = 0x6b206574;
    for (int i = 0; i < 8; i++) s[4 + i] = ld32(key + 4 * i);
    s[12] = counter;
    for (int i = 0; i < 3; i++) s[13 + i] = ld32(nonce + 4 * i);
    memcpy(x, s, sizeof s);
    for (int r = 0; r < 10; r++) {
        QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) {
        uint32_t v = x[i] + s[i];
        out[4 * i] = (uint8_t)v; out[4 * i + 1] = (uint8_t)(v >> 8);
        out[4 * i + 2] = (uint8_t)(v >> 16); out[4 * i + 3] = (uint8_t)(v >> 24);
    }
}

/* project_and_extract_seq, jl.rs:388-508. The keystream of part `part` is
 * ChaCha20(key = seed, nonce = LE u32(part) || 0^8) from block 0; column col
 * uses bytes [col*packed_len, (col+1)*packed_len) with packed_len = ceil(rows/4);
 * row m reads bits 2*(m%4) of byte m/4: 00 -> +1, 01 -> -1, else 0. */
int gr_jl_project(uint64_t q, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                  uint64_t cols, uint64_t len, const uint64_t *parts, uint64_t *p_modq,
                  int64_t *p_exact, int8_t *flat) {
    if (rows == 0 || len == 0 || cols % len != 0) return GR_E_INVALID_ARGUMENT;
    field_t f; field_init(&f, q);
    uint64_t n = cols / len;
    uint64_t packed = (rows + 3) / 4;
    uint64_t *acc = (uint64_t *)calloc(rows, sizeof(uint64_t));
    for (uint32_t m = 0; m < rows; m++) acc[m] = 0; /* Montgomery zero */
    if (p_exact) for (uint32_t m = 0; m < rows; m++) p_exact[m] = 0;
    uint8_t block[64];
    uint8_t nonce[12];
    for (uint32_t part = 0; part < num_parts; part++) {
        memset(nonce, 0, sizeof nonce);
        nonce[0] = (uint8_t)part; nonce[1] = (uint8_t)(part >> 8);
        nonce[2] = (uint8_t)(part >> 16); nonce[3] = (uint8_t)(part >> 24);
        const uint64_t *sp = parts + (size_t)part * cols;
        uint64_t cur_block = UINT64_MAX;
        for (uint64_t col = 0; col < cols; col++) {
            uint64_t cm = f_to(&f, sp[col]);
            uint64_t c = sp[col] % q;
            int64_t centred = c > q / 2 ? (int64_t)c - (int64_t)q : (int64_t)c;
            uint64_t poly = col / len, coeff = col % len;
            for (uint64_t b = 0; b < packed; b++) {
                uint64_t off = col * packed + b;
                if (off / 64 != cur_block) {
                    cur_block = off / 64;
                    gr_chacha20_block(seed, (uint32_t)cur_block, nonce, block);
                }
                uint8_t byte = block[off % 64];
                for (int nib = 0; nib < 4; nib++) {
                    uint64_t row = b * 4 + nib;
                    if (row >= rows) break;
                    int v = (byte >> (2 * nib)) & 3;
                    int8_t e = v == 0 ? 1 : (v == 1 ? -1 : 0);
                    if (e == 1) acc[row] = f_add(&f, acc[row], cm);
                    else if (e == -1) acc[row] = f_sub(&f, acc[row], cm);
                    if (p_exact) p_exact[row] += e * centred;
                    if (flat)
                        flat[row * num_parts * cols + ((uint64_t)part * len + coeff) * n + poly] = e;
                }
            }
        }
    }
    for (uint32_t m = 0; m < rows; m++) p_modq[m] = f_from(&f, acc[m]);
    free(acc);
    return GR_OK;
}
