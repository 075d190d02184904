/*
 * gridref_avx2.c — restatement of the reference's AVX2 CPU path for the
 * prepared Ajtai commitment over a prime q < 2^32 (ORACLE / CPU BASELINE).
 *
 * TEST INFRASTRUCTURE ONLY: timed by bench.py's cpu_baseline leg and checked
 * against gridref.c by the tests. Not part of the product.
 *
 * The commit mirrors CyclotomicPolyRing::recompute_linear_commitment_runtime
 * (commit/src/linear.rs:717-770) with a prepared (NTT-domain) key:
 *   1. twisted forward NTT of every witness polynomial (poly/ntt.rs:75-87):
 *      pointwise twist (simd/montgomery.rs:512-633), bit_reverse_permute
 *      (arith/ntt.rs:180-194), AVX2 DIT stages (simd/montgomery.rs:688-742);
 *   2. per row, TwistedNttPoly::dot_product (poly/twisted_ntt.rs:378-408). For
 *      N = 64 <= CHUNK that function takes its scalar branch: acc += a * b per
 *      slot with the scalar Montgomery multiply — restated as such;
 *   3. inverse (simd/montgomery.rs:802-890: AVX2 DIF stages, n^-1 on the last
 *      stage, bit reverse) and untwist.
 * Two limb widths, as PrimeField<Q, L> offers:
 *   L = u64 (default, prime.rs:26): R = 2^64; for q < 2^32 products through
 *       _mm256_mul_epu32 then a per-lane scalar Montgomery reduce
 *       (u64_arith.rs:360-539); for q >= 2^32 (Goldilocks) the wide 4-lane
 *       Montgomery multiply built from four _mm256_mul_epu32 per product
 *       (mul_epu64_wide_avx2 / montgomery_mul_u64_avx2, u64_arith.rs:193-245),
 *       the carry-aware add/sub (:161-178) and the wide butterflies (:287-357),
 *       with the scalar montgomery_reduce_wide (prime.rs:158-183) for the dot
 *       product and the tails;
 *   L = u32: R = 2^32, 8-lane Montgomery with the wide carry path for
 *       q >= 2^31 (u32_arith.rs:28-315).
 * Also restated: the negacyclic product of two polynomials
 * (NegacyclicMulRing::negacyclic_mul_coeffs, poly/ring.rs:83-101 ->
 * TwistedNttPlan::multiply_in_place, poly/ntt.rs:110-126), and commitments
 * of a prefix of the key's columns (a witness shorter than the key).
 */
#include "gridref.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

__attribute__((target("avx2"))) static int cpu_has_avx2(void) {
    return __builtin_cpu_supports("avx2");
}
int gr_avx2_available(void) { return cpu_has_avx2(); }

/* ---------------- scalar field words for both limb widths ---------------- */

typedef struct {
    uint64_t q;
    int limb; /* 32 or 64 */
    int wide; /* limb 64 and q >= 2^63: montgomery_reduce_wide */
    uint64_t qinv64;  /* -q^-1 mod 2^64 */
    uint32_t qinv32;  /* -q^-1 mod 2^32 */
    uint64_t one;     /* R mod q */
    uint64_t r2;      /* R^2 mod q */
} fw_t;

static void fw_init(fw_t *f, uint64_t q, int limb) {
    f->q = q;
    f->limb = limb;
    f->wide = limb == 64 && q >= ((uint64_t)1 << 63);
    uint64_t inv = 1;
    for (int i = 0; i < 6; i++) inv *= 2 - q * inv;
    f->qinv64 = (uint64_t)0 - inv;
    f->qinv32 = (uint32_t)f->qinv64;
    if (limb == 64) {
        f->one = (uint64_t)(((u128)1 << 64) % q);
    } else {
        f->one = ((uint64_t)1 << 32) % q;
    }
    f->r2 = (uint64_t)(((u128)f->one * f->one) % q);
}

/* L = u64, q >= 2^63: montgomery_reduce_wide (prime.rs:158-183), the sum
 * t + m q split into 64-bit halves with explicit carries */
static inline uint64_t red64_wide(const fw_t *f, u128 t) {
    uint64_t t_lo = (uint64_t)t, t_hi = (uint64_t)(t >> 64);
    uint64_t m = t_lo * f->qinv64;
    u128 prod = (u128)m * f->q;
    uint64_t p_lo = (uint64_t)prod, p_hi = (uint64_t)(prod >> 64);
    uint64_t s_lo = t_lo + p_lo;
    int c_lo = s_lo < t_lo;
    uint64_t s_hi = t_hi + p_hi;
    int c_a = s_hi < t_hi;
    uint64_t s_hi2 = s_hi + (uint64_t)c_lo;
    int c_b = s_hi2 < s_hi;
    if (c_a || c_b) return s_hi2 + (0 - f->q);
    return s_hi2 >= f->q ? s_hi2 - f->q : s_hi2;
}

/* L = u64: montgomery_reduce (prime.rs:141-155), wide for q >= 2^63 */
static inline uint64_t red64(const fw_t *f, u128 t) {
    if (f->wide) return red64_wide(f, t);
    uint64_t m = (uint64_t)t * f->qinv64;
    uint64_t r = (uint64_t)((t + (u128)m * f->q) >> 64);
    return r >= f->q ? r - f->q : r;
}

/* L = u32, q >= 2^31: montgomery_reduce_wide (prime.rs:158-183) */
static inline uint32_t red32(const fw_t *f, uint64_t t) {
    uint32_t q = (uint32_t)f->q;
    uint32_t t_lo = (uint32_t)t, t_hi = (uint32_t)(t >> 32);
    uint32_t m = t_lo * f->qinv32;
    uint64_t prod = (uint64_t)m * q;
    uint32_t p_lo = (uint32_t)prod, p_hi = (uint32_t)(prod >> 32);
    uint32_t lo = t_lo + p_lo;
    int c_lo = lo < t_lo;
    uint32_t hi = t_hi + p_hi;
    int c1 = hi < t_hi;
    uint32_t hi2 = hi + (uint32_t)c_lo;
    int c2 = hi2 < hi;
    if (c1 || c2) return hi2 - q;
    return hi2 >= q ? hi2 - q : hi2;
}

static inline uint64_t fw_mul(const fw_t *f, uint64_t a, uint64_t b) {
    return f->limb == 64 ? red64(f, (u128)a * b) : red32(f, a * b);
}
/* add_raw_words (prime.rs:110-127) for either limb */
static inline uint64_t fw_add(const fw_t *f, uint64_t a, uint64_t b) {
    if (f->limb == 64) {
        uint64_t s = a + b;
        return (s < a || s >= f->q) ? s - f->q : s;
    }
    uint32_t x = (uint32_t)a, y = (uint32_t)b, q = (uint32_t)f->q;
    uint32_t s = x + y;
    return (s < x || s >= q) ? (uint32_t)(s - q) : s;
}
static inline uint64_t fw_sub(const fw_t *f, uint64_t a, uint64_t b) {
    return a >= b ? a - b : f->q - b + a;
}
static inline uint64_t fw_to(const fw_t *f, uint64_t x) { return fw_mul(f, x % f->q, f->r2); }
static inline uint64_t fw_from(const fw_t *f, uint64_t x) {
    return f->limb == 64 ? red64(f, (u128)x) : red32(f, x);
}
static uint64_t fw_pow(const fw_t *f, uint64_t b, uint64_t e) {
    uint64_t acc = f->one;
    while (e) { if (e & 1) acc = fw_mul(f, acc, b); b = fw_mul(f, b, b); e >>= 1; }
    return acc;
}

/* ------------- AVX2 wide Montgomery, limb = u64, q >= 2^32 --------------- */

/* cmpgt_epu64_avx2 (u64_arith.rs:151-158) */
__attribute__((target("avx2"))) static inline __m256i cmpgt_epu64(__m256i a, __m256i b) {
    __m256i sign = _mm256_set1_epi64x((long long)((uint64_t)1 << 63));
    return _mm256_cmpgt_epi64(_mm256_xor_si256(a, sign), _mm256_xor_si256(b, sign));
}
/* add_mod_u64_avx2 / sub_mod_u64_avx2 (u64_arith.rs:160-178) */
__attribute__((target("avx2"))) static inline __m256i addm_u64x4(__m256i a, __m256i b,
                                                                 __m256i qv, __m256i thr) {
    __m256i s = _mm256_add_epi64(a, b);
    __m256i red = _mm256_or_si256(cmpgt_epu64(a, s), cmpgt_epu64(s, thr));
    return _mm256_sub_epi64(s, _mm256_and_si256(red, qv));
}
__attribute__((target("avx2"))) static inline __m256i subm_u64x4(__m256i a, __m256i b,
                                                                 __m256i qv) {
    __m256i d = _mm256_sub_epi64(a, b);
    return _mm256_add_epi64(d, _mm256_and_si256(cmpgt_epu64(b, a), qv));
}
/* mul_lo_epu64_avx2 (u64_arith.rs:180-189): low 64 bits of a 64 x 64 product */
__attribute__((target("avx2"))) static inline __m256i mullo_u64x4(__m256i a, __m256i b) {
    __m256i ah = _mm256_srli_epi64(a, 32), bh = _mm256_srli_epi64(b, 32);
    __m256i cross = _mm256_add_epi64(_mm256_mul_epu32(a, bh), _mm256_mul_epu32(ah, b));
    return _mm256_add_epi64(_mm256_mul_epu32(a, b), _mm256_slli_epi64(cross, 32));
}
/* mul_epu64_wide_avx2 (u64_arith.rs:191-218): 128-bit products from four
 * 32 x 32 partial products with the two carries */
__attribute__((target("avx2"))) static inline void mulwide_u64x4(__m256i a, __m256i b,
                                                                 __m256i *lo, __m256i *hi) {
    __m256i ah = _mm256_srli_epi64(a, 32), bh = _mm256_srli_epi64(b, 32);
    __m256i p0 = _mm256_mul_epu32(a, b), p1 = _mm256_mul_epu32(a, bh);
    __m256i p2 = _mm256_mul_epu32(ah, b), p3 = _mm256_mul_epu32(ah, bh);
    __m256i cross = _mm256_add_epi64(p1, p2);
    __m256i cross_carry = cmpgt_epu64(p1, cross);
    __m256i l = _mm256_add_epi64(p0, _mm256_slli_epi64(cross, 32));
    __m256i lo_carry = cmpgt_epu64(p0, l);
    __m256i h = _mm256_add_epi64(
        _mm256_add_epi64(p3, _mm256_add_epi64(_mm256_srli_epi64(cross, 32),
                                              _mm256_and_si256(cross_carry,
                                                               _mm256_set1_epi64x((long long)1 << 32)))),
        _mm256_and_si256(lo_carry, _mm256_set1_epi64x(1)));
    *lo = l;
    *hi = h;
}
/* montgomery_mul_u64_avx2 (u64_arith.rs:220-245) */
__attribute__((target("avx2"))) static inline __m256i montw_u64x4(const fw_t *f, __m256i a,
                                                                  __m256i b) {
    __m256i qv = _mm256_set1_epi64x((long long)f->q);
    __m256i thr = _mm256_set1_epi64x((long long)(f->q - 1));
    __m256i qinv = _mm256_set1_epi64x((long long)f->qinv64);
    __m256i one = _mm256_set1_epi64x(1);
    __m256i t_lo, t_hi, mq_lo, mq_hi;
    mulwide_u64x4(a, b, &t_lo, &t_hi);
    __m256i m = mullo_u64x4(t_lo, qinv);
    mulwide_u64x4(m, qv, &mq_lo, &mq_hi);
    __m256i sum_lo = _mm256_add_epi64(t_lo, mq_lo);
    __m256i carry_lo = cmpgt_epu64(t_lo, sum_lo);
    __m256i sum_hi_a = _mm256_add_epi64(t_hi, mq_hi);
    __m256i carry_a = cmpgt_epu64(t_hi, sum_hi_a);
    __m256i sum_hi = _mm256_add_epi64(sum_hi_a, _mm256_and_si256(carry_lo, one));
    __m256i carry_b = cmpgt_epu64(sum_hi_a, sum_hi);
    __m256i red = _mm256_or_si256(_mm256_or_si256(carry_a, carry_b), cmpgt_epu64(sum_hi, thr));
    return _mm256_sub_epi64(sum_hi, _mm256_and_si256(red, qv));
}

/* mul_assign / scalar_mul / butterflies, prime_wide_u64 (u64_arith.rs:247-357) */
__attribute__((target("avx2"))) static void mul_assign_w64(const fw_t *f, uint64_t *dst,
                                                           const uint64_t *src, size_t len) {
    size_t i = 0;
    for (; i + 4 <= len; i += 4) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        __m256i b = _mm256_loadu_si256((const __m256i *)(src + i));
        _mm256_storeu_si256((__m256i *)(dst + i), montw_u64x4(f, a, b));
    }
    for (; i < len; i++) dst[i] = red64(f, (u128)dst[i] * src[i]);
}
__attribute__((target("avx2"))) static void scalar_mul_w64(const fw_t *f, uint64_t *dst,
                                                           size_t len, uint64_t s) {
    size_t i = 0;
    __m256i sv = _mm256_set1_epi64x((long long)s);
    for (; i + 4 <= len; i += 4) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), montw_u64x4(f, a, sv));
    }
    for (; i < len; i++) dst[i] = red64(f, (u128)dst[i] * s);
}
__attribute__((target("avx2"))) static void bfly_fwd_w64(const fw_t *f, uint64_t *even,
                                                         uint64_t *odd, const uint64_t *tw,
                                                         size_t len) {
    size_t i = 0;
    __m256i qv = _mm256_set1_epi64x((long long)f->q);
    __m256i thr = _mm256_set1_epi64x((long long)(f->q - 1));
    for (; i + 4 <= len; i += 4) {
        __m256i u = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i o = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        __m256i v = montw_u64x4(f, o, w);
        _mm256_storeu_si256((__m256i *)(even + i), addm_u64x4(u, v, qv, thr));
        _mm256_storeu_si256((__m256i *)(odd + i), subm_u64x4(u, v, qv));
    }
    for (; i < len; i++) {
        uint64_t u = even[i], v = red64(f, (u128)odd[i] * tw[i]);
        even[i] = fw_add(f, u, v);
        odd[i] = fw_sub(f, u, v);
    }
}
__attribute__((target("avx2"))) static void bfly_inv_w64(const fw_t *f, uint64_t *even,
                                                         uint64_t *odd, const uint64_t *tw,
                                                         size_t len) {
    size_t i = 0;
    __m256i qv = _mm256_set1_epi64x((long long)f->q);
    __m256i thr = _mm256_set1_epi64x((long long)(f->q - 1));
    for (; i + 4 <= len; i += 4) {
        __m256i u = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i v = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        _mm256_storeu_si256((__m256i *)(even + i), addm_u64x4(u, v, qv, thr));
        _mm256_storeu_si256((__m256i *)(odd + i), montw_u64x4(f, subm_u64x4(u, v, qv), w));
    }
    for (; i < len; i++) {
        uint64_t u = even[i], v = odd[i];
        even[i] = fw_add(f, u, v);
        odd[i] = red64(f, (u128)fw_sub(f, u, v) * tw[i]);
    }
}

/* --------------------- AVX2 kernels, limb = u64 ------------------------- */

/* mul_assign_prime_montgomery_u64_avx2, Q < 2^32 branch (u64_arith.rs:360-391) */
__attribute__((target("avx2"))) static void mul_assign_u64(const fw_t *f, uint64_t *dst,
                                                           const uint64_t *src, size_t len) {
    if (f->q >= ((uint64_t)1 << 32)) { mul_assign_w64(f, dst, src, len); return; }
    size_t i = 0;
    for (; i + 4 <= len; i += 4) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        __m256i b = _mm256_loadu_si256((const __m256i *)(src + i));
        uint64_t lanes[4];
        _mm256_storeu_si256((__m256i *)lanes, _mm256_mul_epu32(a, b));
        for (int l = 0; l < 4; l++) lanes[l] = red64(f, lanes[l]);
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_loadu_si256((const __m256i *)lanes));
    }
    for (; i < len; i++) dst[i] = red64(f, dst[i] * src[i]);
}

/* scalar_mul_prime_montgomery_u64_avx2, Q < 2^32 branch (u64_arith.rs:394-425) */
__attribute__((target("avx2"))) static void scalar_mul_u64(const fw_t *f, uint64_t *dst,
                                                           size_t len, uint64_t s) {
    if (f->q >= ((uint64_t)1 << 32)) { scalar_mul_w64(f, dst, len, s); return; }
    size_t i = 0;
    __m256i sv = _mm256_set1_epi64x((long long)s);
    for (; i + 4 <= len; i += 4) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        uint64_t lanes[4];
        _mm256_storeu_si256((__m256i *)lanes, _mm256_mul_epu32(a, sv));
        for (int l = 0; l < 4; l++) lanes[l] = red64(f, lanes[l]);
        _mm256_storeu_si256((__m256i *)(dst + i), _mm256_loadu_si256((const __m256i *)lanes));
    }
    for (; i < len; i++) dst[i] = red64(f, dst[i] * s);
}

/* butterfly_forward_prime_montgomery_u64_avx2, Q < 2^32 (u64_arith.rs:428-481) */
__attribute__((target("avx2"))) static void bfly_fwd_u64(const fw_t *f, uint64_t *even,
                                                         uint64_t *odd, const uint64_t *tw,
                                                         size_t len) {
    if (f->q >= ((uint64_t)1 << 32)) { bfly_fwd_w64(f, even, odd, tw, len); return; }
    size_t i = 0;
    __m256i qv = _mm256_set1_epi64x((long long)f->q);
    __m256i thr = _mm256_set1_epi64x((long long)(f->q - 1));
    for (; i + 4 <= len; i += 4) {
        __m256i u = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i o = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        uint64_t lanes[4];
        _mm256_storeu_si256((__m256i *)lanes, _mm256_mul_epu32(o, w));
        for (int l = 0; l < 4; l++) lanes[l] = red64(f, lanes[l]);
        __m256i v = _mm256_loadu_si256((const __m256i *)lanes);
        __m256i sum = _mm256_add_epi64(u, v);
        __m256i e = _mm256_sub_epi64(sum, _mm256_and_si256(_mm256_cmpgt_epi64(sum, thr), qv));
        __m256i d = _mm256_add_epi64(_mm256_sub_epi64(u, v),
                                     _mm256_and_si256(_mm256_cmpgt_epi64(v, u), qv));
        _mm256_storeu_si256((__m256i *)(even + i), e);
        _mm256_storeu_si256((__m256i *)(odd + i), d);
    }
    for (; i < len; i++) {
        uint64_t u = even[i], v = red64(f, (u128)odd[i] * tw[i]);
        even[i] = u + v >= f->q ? u + v - f->q : u + v;
        odd[i] = u >= v ? u - v : u + f->q - v;
    }
}

/* butterfly_inverse_prime_montgomery_u64_avx2, Q < 2^32 (u64_arith.rs:484-539) */
__attribute__((target("avx2"))) static void bfly_inv_u64(const fw_t *f, uint64_t *even,
                                                         uint64_t *odd, const uint64_t *tw,
                                                         size_t len) {
    if (f->q >= ((uint64_t)1 << 32)) { bfly_inv_w64(f, even, odd, tw, len); return; }
    size_t i = 0;
    __m256i qv = _mm256_set1_epi64x((long long)f->q);
    __m256i thr = _mm256_set1_epi64x((long long)(f->q - 1));
    for (; i + 4 <= len; i += 4) {
        __m256i u = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i v = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i sum = _mm256_add_epi64(u, v);
        __m256i e = _mm256_sub_epi64(sum, _mm256_and_si256(_mm256_cmpgt_epi64(sum, thr), qv));
        __m256i d = _mm256_add_epi64(_mm256_sub_epi64(u, v),
                                     _mm256_and_si256(_mm256_cmpgt_epi64(v, u), qv));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        uint64_t lanes[4];
        _mm256_storeu_si256((__m256i *)lanes, _mm256_mul_epu32(d, w));
        for (int l = 0; l < 4; l++) lanes[l] = red64(f, lanes[l]);
        _mm256_storeu_si256((__m256i *)(even + i), e);
        _mm256_storeu_si256((__m256i *)(odd + i), _mm256_loadu_si256((const __m256i *)lanes));
    }
    for (; i < len; i++) {
        uint64_t u = even[i], v = odd[i];
        uint64_t s = u + v >= f->q ? u + v - f->q : u + v;
        uint64_t d = u >= v ? u - v : u + f->q - v;
        even[i] = s;
        odd[i] = red64(f, (u128)d * tw[i]);
    }
}

/* --------------------- AVX2 kernels, limb = u32 ------------------------- */

__attribute__((target("avx2"))) static inline __m256i cmpgt_epu32(__m256i a, __m256i b) {
    __m256i sign = _mm256_set1_epi32((int)0x80000000u);
    return _mm256_cmpgt_epi32(_mm256_xor_si256(a, sign), _mm256_xor_si256(b, sign));
}

/* widen_mul_u32_full (u32_arith.rs:28-51): 8-lane 32x32 -> (lo, hi) */
__attribute__((target("avx2"))) static inline void widen_mul(__m256i a, __m256i b, __m256i *lo,
                                                             __m256i *hi) {
    *lo = _mm256_mullo_epi32(a, b);
    __m256i ev = _mm256_srli_epi64(_mm256_mul_epu32(a, b), 32);
    __m256i od = _mm256_srli_epi64(_mm256_mul_epu32(_mm256_srli_epi64(a, 32),
                                                    _mm256_srli_epi64(b, 32)), 32);
    *hi = _mm256_or_si256(ev, _mm256_slli_epi64(od, 32));
}

/* montgomery_mul_u32_avx2, wide-carry branch for q >= 2^31 (u32_arith.rs:58-92) */
__attribute__((target("avx2"))) static inline __m256i mont_mul_u32x8(const fw_t *f, __m256i a,
                                                                     __m256i b) {
    __m256i qv = _mm256_set1_epi32((int)(uint32_t)f->q);
    __m256i qinv = _mm256_set1_epi32((int)f->qinv32);
    __m256i thr = _mm256_set1_epi32((int)(uint32_t)(f->q - 1));
    __m256i ones = _mm256_set1_epi32(1);
    __m256i t_lo, t_hi, mq_lo, mq_hi;
    widen_mul(a, b, &t_lo, &t_hi);
    __m256i m = _mm256_mullo_epi32(t_lo, qinv);
    widen_mul(m, qv, &mq_lo, &mq_hi);
    __m256i sum_lo = _mm256_add_epi32(t_lo, mq_lo);
    __m256i c_lo = cmpgt_epu32(t_lo, sum_lo);
    __m256i sum_hi_a = _mm256_add_epi32(t_hi, mq_hi);
    __m256i sum_hi = _mm256_add_epi32(sum_hi_a, _mm256_and_si256(c_lo, ones));
    if (f->q < ((uint64_t)1 << 31)) {
        __m256i red = cmpgt_epu32(sum_hi, thr);
        return _mm256_sub_epi32(sum_hi, _mm256_and_si256(red, qv));
    }
    __m256i ov = _mm256_or_si256(cmpgt_epu32(t_hi, sum_hi_a), cmpgt_epu32(sum_hi_a, sum_hi));
    __m256i red = _mm256_or_si256(ov, cmpgt_epu32(sum_hi, thr));
    return _mm256_sub_epi32(sum_hi, _mm256_and_si256(red, qv));
}

/* modular add of 8 lanes with the carry test (u32_arith.rs:143-182) */
__attribute__((target("avx2"))) static inline __m256i add_u32x8(const fw_t *f, __m256i a,
                                                                __m256i b) {
    __m256i qv = _mm256_set1_epi32((int)(uint32_t)f->q);
    __m256i thr = _mm256_set1_epi32((int)(uint32_t)(f->q - 1));
    __m256i s = _mm256_add_epi32(a, b);
    __m256i red = _mm256_or_si256(cmpgt_epu32(a, s), cmpgt_epu32(s, thr));
    return _mm256_sub_epi32(s, _mm256_and_si256(red, qv));
}
__attribute__((target("avx2"))) static inline __m256i sub_u32x8(const fw_t *f, __m256i a,
                                                                __m256i b) {
    __m256i qv = _mm256_set1_epi32((int)(uint32_t)f->q);
    __m256i borrow = cmpgt_epu32(b, a);
    return _mm256_add_epi32(_mm256_sub_epi32(a, b), _mm256_and_si256(borrow, qv));
}

__attribute__((target("avx2"))) static void mul_assign_u32(const fw_t *f, uint32_t *dst,
                                                           const uint32_t *src, size_t len) {
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        __m256i b = _mm256_loadu_si256((const __m256i *)(src + i));
        _mm256_storeu_si256((__m256i *)(dst + i), mont_mul_u32x8(f, a, b));
    }
    for (; i < len; i++) dst[i] = red32(f, (uint64_t)dst[i] * src[i]);
}
__attribute__((target("avx2"))) static void scalar_mul_u32(const fw_t *f, uint32_t *dst,
                                                           size_t len, uint32_t s) {
    size_t i = 0;
    __m256i sv = _mm256_set1_epi32((int)s);
    for (; i + 8 <= len; i += 8) {
        __m256i a = _mm256_loadu_si256((const __m256i *)(dst + i));
        _mm256_storeu_si256((__m256i *)(dst + i), mont_mul_u32x8(f, a, sv));
    }
    for (; i < len; i++) dst[i] = red32(f, (uint64_t)dst[i] * s);
}

/* butterfly_forward_prime_montgomery_u32_avx2 (u32_arith.rs:215-263) */
__attribute__((target("avx2"))) static void bfly_fwd_u32(const fw_t *f, uint32_t *even,
                                                         uint32_t *odd, const uint32_t *tw,
                                                         size_t len) {
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
        __m256i e = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i o = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        __m256i t = mont_mul_u32x8(f, o, w);
        _mm256_storeu_si256((__m256i *)(even + i), add_u32x8(f, e, t));
        _mm256_storeu_si256((__m256i *)(odd + i), sub_u32x8(f, e, t));
    }
    for (; i < len; i++) {
        uint32_t e = even[i], t = red32(f, (uint64_t)odd[i] * tw[i]);
        even[i] = (uint32_t)fw_add(f, e, t);
        odd[i] = (uint32_t)fw_sub(f, e, t);
    }
}

/* butterfly_inverse_prime_montgomery_u32_avx2 (u32_arith.rs:266-315) */
__attribute__((target("avx2"))) static void bfly_inv_u32(const fw_t *f, uint32_t *even,
                                                         uint32_t *odd, const uint32_t *tw,
                                                         size_t len) {
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
        __m256i e = _mm256_loadu_si256((const __m256i *)(even + i));
        __m256i o = _mm256_loadu_si256((const __m256i *)(odd + i));
        __m256i w = _mm256_loadu_si256((const __m256i *)(tw + i));
        _mm256_storeu_si256((__m256i *)(even + i), add_u32x8(f, e, o));
        _mm256_storeu_si256((__m256i *)(odd + i), mont_mul_u32x8(f, sub_u32x8(f, e, o), w));
    }
    for (; i < len; i++) {
        uint32_t e = even[i], o = odd[i];
        even[i] = (uint32_t)fw_add(f, e, o);
        odd[i] = red32(f, (uint64_t)(uint32_t)fw_sub(f, e, o) * tw[i]);
    }
}

/* --------------------------- plan + commit -------------------------------- */

typedef struct {
    fw_t f;
    uint32_t rows, n, d;
    uint32_t n_use;  /* witness polynomials of the next commit (a prefix of the key's columns) */
    /* negacyclic products: count pairs in Montgomery form, out in Montgomery form */
    void *pa, *pb, *pout;
    size_t pairs;
    /* coefficient-domain sums of products (Gram, inner commitments):
     * x: nx vectors of len polys, y: ny vectors of len polys; task t sums
     * negacyclic products x[tx[t]][k] * y[ty[t]][k] over k into out[t] */
    void *vx, *vy, *vout;
    uint32_t *tx, *ty;
    uint32_t vlen, ntasks;
    /* independent commitments: item i's witness is pool polys
     * [ioff[i], ioff[i] + ilen[i]) against the key's first ilen[i] columns */
    void *pool, *iout;
    uint64_t *ioff;
    uint32_t *ilen;
    uint32_t nitems;
    void *a;        /* rows*n*d limbs, Montgomery */
    void *s;        /* n*d limbs, Montgomery, coefficient domain */
    void *s_work;   /* NTT scratch */
    void *t;        /* rows*d limbs */
    void *fwd, *inv, *twist, *untwist; /* plan tables (limb words) */
    uint64_t inv_scale;
} avx_ctx_t;

static void bit_reverse_u64(uint64_t *v, uint32_t n) {
    uint32_t j = 0;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) { uint64_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}
static void bit_reverse_u32(uint32_t *v, uint32_t n) {
    uint32_t j = 0;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) { uint32_t t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}

/* twisted forward / inverse on one polynomial with the AVX2 kernels */
static void twisted_fwd(avx_ctx_t *c, void *poly) {
    uint32_t d = c->d;
    if (c->f.limb == 64) {
        uint64_t *v = (uint64_t *)poly;
        mul_assign_u64(&c->f, v, (const uint64_t *)c->twist, d);
        bit_reverse_u64(v, d);
        const uint64_t *tw = (const uint64_t *)c->fwd;
        for (uint32_t len = 2; len <= d; len <<= 1) {
            uint32_t half = len / 2;
            for (uint32_t st = 0; st < d; st += len) bfly_fwd_u64(&c->f, v + st, v + st + half, tw, half);
            tw += half;
        }
    } else {
        uint32_t *v = (uint32_t *)poly;
        mul_assign_u32(&c->f, v, (const uint32_t *)c->twist, d);
        bit_reverse_u32(v, d);
        const uint32_t *tw = (const uint32_t *)c->fwd;
        for (uint32_t len = 2; len <= d; len <<= 1) {
            uint32_t half = len / 2;
            for (uint32_t st = 0; st < d; st += len) bfly_fwd_u32(&c->f, v + st, v + st + half, tw, half);
            tw += half;
        }
    }
}

static void twisted_inv(avx_ctx_t *c, void *poly) {
    uint32_t d = c->d;
    if (c->f.limb == 64) {
        uint64_t *v = (uint64_t *)poly;
        const uint64_t *tw = (const uint64_t *)c->inv;
        for (uint32_t len = d; len >= 2; len >>= 1) {
            uint32_t half = len / 2;
            for (uint32_t st = 0; st < d; st += len) {
                bfly_inv_u64(&c->f, v + st, v + st + half, tw, half);
                if (len == 2) scalar_mul_u64(&c->f, v + st, len, c->inv_scale);
            }
            tw += half;
        }
        bit_reverse_u64(v, d);
        mul_assign_u64(&c->f, v, (const uint64_t *)c->untwist, d);
    } else {
        uint32_t *v = (uint32_t *)poly;
        const uint32_t *tw = (const uint32_t *)c->inv;
        for (uint32_t len = d; len >= 2; len >>= 1) {
            uint32_t half = len / 2;
            for (uint32_t st = 0; st < d; st += len) {
                bfly_inv_u32(&c->f, v + st, v + st + half, tw, half);
                if (len == 2) scalar_mul_u32(&c->f, v + st, len, (uint32_t)c->inv_scale);
            }
            tw += half;
        }
        bit_reverse_u32(v, d);
        mul_assign_u32(&c->f, v, (const uint32_t *)c->untwist, d);
    }
}

static size_t limb_bytes(const avx_ctx_t *c) { return c->f.limb == 64 ? 8 : 4; }
static uint64_t get_w(const avx_ctx_t *c, const void *base, size_t i) {
    return c->f.limb == 64 ? ((const uint64_t *)base)[i] : ((const uint32_t *)base)[i];
}
static void set_w(const avx_ctx_t *c, void *base, size_t i, uint64_t v) {
    if (c->f.limb == 64) ((uint64_t *)base)[i] = v; else ((uint32_t *)base)[i] = (uint32_t)v;
}

/* TwistedNttPoly::dot_product, N <= CHUNK branch (twisted_ntt.rs:382-394):
 * acc[k] = sum over col < len of a[row][col][k] * s[col][k] with the scalar
 * mul_raw_words / add_raw_words of the limb type; the limb case is decided
 * once per row, as the reference's monomorphised code does at compile time. */
static void dot_row(const avx_ctx_t *c, uint32_t row, const void *s, uint32_t len,
                    uint64_t *acc) {
    const fw_t *f = &c->f;
    const size_t d = c->d;
    for (size_t k = 0; k < d; k++) acc[k] = 0;
    if (f->limb == 64 && !f->wide) {
        const uint64_t *a = (const uint64_t *)c->a + (size_t)row * c->n * d;
        const uint64_t *sw = (const uint64_t *)s, q = f->q, qinv = f->qinv64;
        for (uint32_t col = 0; col < len; col++)
            for (size_t k = 0; k < d; k++) {
                u128 t = (u128)a[col * d + k] * sw[col * d + k];
                uint64_t m = (uint64_t)t * qinv;
                uint64_t r = (uint64_t)((t + (u128)m * q) >> 64);
                r = r >= q ? r - q : r;
                uint64_t x = acc[k] + r; /* q < 2^63: no carry (prime.rs:113-119) */
                acc[k] = x >= q ? x - q : x;
            }
    } else if (f->limb == 64) {
        const uint64_t *a = (const uint64_t *)c->a + (size_t)row * c->n * d;
        const uint64_t *sw = (const uint64_t *)s;
        for (uint32_t col = 0; col < len; col++)
            for (size_t k = 0; k < d; k++)
                acc[k] = fw_add(f, acc[k], red64_wide(f, (u128)a[col * d + k] * sw[col * d + k]));
    } else {
        const uint32_t *a = (const uint32_t *)c->a + (size_t)row * c->n * d;
        const uint32_t *sw = (const uint32_t *)s;
        for (uint32_t col = 0; col < len; col++)
            for (size_t k = 0; k < d; k++)
                acc[k] = fw_add(f, acc[k], red32(f, (uint64_t)a[col * d + k] * sw[col * d + k]));
    }
}

void *gr_avx2_prepare(uint64_t q, int limb_bits, uint32_t rows, uint32_t n, uint32_t d,
                      const uint64_t *a_ntt) {
    if ((limb_bits != 32 && limb_bits != 64) || d < 2 || d > 256) return NULL;
    if (limb_bits == 32 && q >= ((uint64_t)1 << 32)) return NULL;
    uint64_t psi_c, root_c;
    if (gr_root_of_unity(q, 2 * (uint64_t)d, &psi_c) != GR_OK) return NULL;
    if (gr_root_of_unity(q, d, &root_c) != GR_OK) return NULL;
    avx_ctx_t *c = (avx_ctx_t *)calloc(1, sizeof *c);
    fw_init(&c->f, q, limb_bits);
    c->rows = rows; c->n = n; c->d = d; c->n_use = n;
    size_t lb = limb_bytes(c);
    c->a = malloc(lb * ((size_t)rows * n * d + 1));
    c->s = malloc(lb * ((size_t)n * d + 1));
    c->s_work = malloc(lb * ((size_t)n * d + 1));
    c->t = malloc(lb * ((size_t)rows * d + 1));
    c->fwd = malloc(lb * d); c->inv = malloc(lb * d);
    c->twist = malloc(lb * d); c->untwist = malloc(lb * d);
    for (size_t i = 0; i < (size_t)rows * n * d; i++) set_w(c, c->a, i, fw_to(&c->f, a_ntt[i]));
    /* NttPlan / TwistedNttPlan tables (arith/ntt.rs:124-153, poly/ntt.rs:140-153) */
    const fw_t *f = &c->f;
    uint64_t root = fw_to(f, root_c), iroot = fw_pow(f, root, q - 2);
    uint64_t psi = fw_to(f, psi_c), ipsi = fw_pow(f, psi, q - 2);
    size_t off = 0;
    for (uint32_t len = 2; len <= d; len <<= 1) {
        uint64_t w = fw_pow(f, root, d / len), cur = f->one;
        for (uint32_t j = 0; j < len / 2; j++) { set_w(c, c->fwd, off + j, cur); cur = fw_mul(f, cur, w); }
        off += len / 2;
    }
    off = 0;
    for (uint32_t len = d; len >= 2; len >>= 1) {
        uint64_t w = fw_pow(f, iroot, d / len), cur = f->one;
        for (uint32_t j = 0; j < len / 2; j++) { set_w(c, c->inv, off + j, cur); cur = fw_mul(f, cur, w); }
        off += len / 2;
    }
    uint64_t a = f->one, b = f->one;
    for (uint32_t i = 0; i < d; i++) {
        set_w(c, c->twist, i, a); set_w(c, c->untwist, i, b);
        a = fw_mul(f, a, psi); b = fw_mul(f, b, ipsi);
    }
    c->inv_scale = fw_pow(f, fw_to(f, d), q - 2);
    return c;
}

void gr_avx2_free(void *h) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    if (!c) return;
    free(c->a); free(c->s); free(c->s_work); free(c->t);
    free(c->fwd); free(c->inv); free(c->twist); free(c->untwist);
    free(c->pa); free(c->pb); free(c->pout);
    free(c->vx); free(c->vy); free(c->vout); free(c->tx); free(c->ty);
    free(c->pool); free(c->iout); free(c->ioff); free(c->ilen);
    free(c);
}

/* The next commits use the key's first n_use columns (n_use <= n): the
 * witness is n_use polynomials. */
int gr_avx2_set_len(void *h, uint32_t n_use) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    if (n_use > c->n) return GR_E_DIMENSION_MISMATCH;
    c->n_use = n_use;
    return GR_OK;
}
void gr_avx2_load_witness_len(void *h, const uint64_t *s, uint32_t n_use) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < (size_t)n_use * c->d; i++) set_w(c, c->s, i, fw_to(&c->f, s[i]));
}

void gr_avx2_load_witness(void *h, const uint64_t *s) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < (size_t)c->n * c->d; i++) set_w(c, c->s, i, fw_to(&c->f, s[i]));
}

typedef struct {
    avx_ctx_t *c;
    uint32_t lo, hi; /* polynomial or row range */
    int phase;
} job_t;

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    avx_ctx_t *c = j->c;
    size_t lb = limb_bytes(c), d = c->d;
    if (j->phase == 4) {
        /* whole commitments per thread: the prepared commit of worker phases
         * 0 and 1 on the thread's own scratch */
        size_t maxlen = 0;
        for (uint32_t i = j->lo; i < j->hi; i++) if (c->ilen[i] > maxlen) maxlen = c->ilen[i];
        char *sw = (char *)malloc(lb * d * (maxlen + 1));
        uint64_t acc[256];
        for (uint32_t i = j->lo; i < j->hi; i++) {
            const uint32_t len = c->ilen[i];
            memcpy(sw, (char *)c->pool + lb * d * c->ioff[i], lb * d * len);
            for (uint32_t p = 0; p < len; p++) twisted_fwd(c, sw + lb * d * p);
            for (uint32_t row = 0; row < c->rows; row++) {
                dot_row(c, row, sw, len, acc);
                char *out = (char *)c->iout + lb * d * ((size_t)i * c->rows + row);
                for (size_t k = 0; k < d; k++) set_w(c, out, k, acc[k]);
                twisted_inv(c, out);
            }
        }
        free(sw);
    } else if (j->phase == 3) {
        /* Ring::dot_product over coefficient-domain polynomials: every
         * product a full negacyclic_mul_coeffs, summed with add_assign_slice
         * (the ring's add, restated per coefficient) */
        void *lhs = malloc(lb * d), *rhs = malloc(lb * d), *acc = malloc(lb * d);
        for (uint32_t t = j->lo; t < j->hi; t++) {
            memset(acc, 0, lb * d);
            for (uint32_t k = 0; k < c->vlen; k++) {
                memcpy(lhs, (char *)c->vx + lb * d * ((size_t)c->tx[t] * c->vlen + k), lb * d);
                memcpy(rhs, (char *)c->vy + lb * d * ((size_t)c->ty[t] * c->vlen + k), lb * d);
                twisted_fwd(c, lhs);
                twisted_fwd(c, rhs);
                if (c->f.limb == 64) mul_assign_u64(&c->f, (uint64_t *)lhs, (const uint64_t *)rhs, d);
                else mul_assign_u32(&c->f, (uint32_t *)lhs, (const uint32_t *)rhs, d);
                twisted_inv(c, lhs);
                for (size_t i = 0; i < d; i++)
                    set_w(c, acc, i, fw_add(&c->f, get_w(c, acc, i), get_w(c, lhs, i)));
            }
            memcpy((char *)c->vout + lb * d * t, acc, lb * d);
        }
        free(lhs);
        free(rhs);
        free(acc);
    } else if (j->phase == 2) {
        /* negacyclic_mul_coeffs (poly/ring.rs:83-101): clone both operands,
         * twisted forward transforms, pointwise product, inverse
         * (TwistedNttPlan::multiply_in_place, poly/ntt.rs:110-126) */
        void *lhs = malloc(lb * d), *rhs = malloc(lb * d);
        for (uint32_t p = j->lo; p < j->hi; p++) {
            memcpy(lhs, (char *)c->pa + lb * d * p, lb * d);
            memcpy(rhs, (char *)c->pb + lb * d * p, lb * d);
            twisted_fwd(c, lhs);
            twisted_fwd(c, rhs);
            if (c->f.limb == 64) mul_assign_u64(&c->f, (uint64_t *)lhs, (const uint64_t *)rhs, d);
            else mul_assign_u32(&c->f, (uint32_t *)lhs, (const uint32_t *)rhs, d);
            twisted_inv(c, lhs);
            memcpy((char *)c->pout + lb * d * p, lhs, lb * d);
        }
        free(lhs);
        free(rhs);
    } else if (j->phase == 5 || j->phase == 6) {
        /* TwistedNttPlan::forward_in_place_trusted / inverse_in_place_trusted
         * (poly/ntt.rs:75-104) in place on the first operand */
        for (uint32_t p = j->lo; p < j->hi; p++) {
            char *poly = (char *)c->pa + lb * d * p;
            if (j->phase == 5) twisted_fwd(c, poly);
            else twisted_inv(c, poly);
        }
    } else if (j->phase == 0) {
        for (uint32_t p = j->lo; p < j->hi; p++) {
            char *dst = (char *)c->s_work + lb * d * p;
            memcpy(dst, (char *)c->s + lb * d * p, lb * d);
            twisted_fwd(c, dst);
        }
    } else {
        /* TwistedNttPoly::dot_product, N <= CHUNK branch (twisted_ntt.rs:382-394) */
        uint64_t acc[256];
        for (uint32_t row = j->lo; row < j->hi; row++) {
            dot_row(c, row, c->s_work, c->n_use, acc);
            char *out = (char *)c->t + lb * d * row;
            for (size_t k = 0; k < d; k++) set_w(c, out, k, acc[k]);
            twisted_inv(c, out);
        }
    }
    return NULL;
}

static void run_phase(avx_ctx_t *c, int phase, uint32_t total, int threads) {
    if (total == 0) return;
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > total) threads = (int)total;
    if (threads <= 1) {
        job_t j = {c, 0, total, phase};
        worker(&j);
        return;
    }
    pthread_t tid[256];
    job_t jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++) {
        jobs[t].c = c; jobs[t].phase = phase;
        jobs[t].lo = (uint32_t)((uint64_t)total * t / threads);
        jobs[t].hi = (uint32_t)((uint64_t)total * (t + 1) / threads);
        pthread_create(&tid[t], NULL, worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

/* The timed region: NTT(s), MAC, INTT — everything a prepared commit does. */
void gr_avx2_commit_run(void *h, int threads) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    run_phase(c, 0, c->n_use, threads);
    run_phase(c, 1, c->rows, threads);
}

/* Negacyclic products of `count` pairs (canonical inputs, converted once
 * outside the timed run). */
int gr_avx2_negacyclic_load(void *h, const uint64_t *a, const uint64_t *b, size_t count) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    size_t lb = limb_bytes(c), words = count * c->d;
    free(c->pa); free(c->pb); free(c->pout);
    c->pa = malloc(lb * words + 1); c->pb = malloc(lb * words + 1); c->pout = malloc(lb * words + 1);
    if (!c->pa || !c->pb || !c->pout) return GR_E_INVALID_ARGUMENT;
    for (size_t i = 0; i < words; i++) {
        set_w(c, c->pa, i, fw_to(&c->f, a[i]));
        set_w(c, c->pb, i, fw_to(&c->f, b[i]));
    }
    c->pairs = count;
    return GR_OK;
}
void gr_avx2_negacyclic_run(void *h, int threads) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    run_phase(c, 2, (uint32_t)c->pairs, threads);
}
/* Sums of coefficient-domain products: x (nx x len polys), y (ny x len
 * polys), task t = (tx[t], ty[t]). The Gram matrix is x = y = s with the
 * upper-triangle pairs (garbage.rs:76-93); inner commitments t_i = A s_i
 * (crs.rs:211-216) are x = A's rows, y = the parts, tasks (row, part). */
int gr_avx2_dots_load(void *h, const uint64_t *x, uint32_t nx, const uint64_t *y, uint32_t ny,
                      uint32_t len, const uint32_t *tx, const uint32_t *ty, uint32_t ntasks) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    size_t lb = limb_bytes(c), d = c->d;
    free(c->vx); free(c->vy); free(c->vout); free(c->tx); free(c->ty);
    c->vx = malloc(lb * d * ((size_t)nx * len + 1));
    c->vy = malloc(lb * d * ((size_t)ny * len + 1));
    c->vout = malloc(lb * d * ((size_t)ntasks + 1));
    c->tx = (uint32_t *)malloc(4 * ((size_t)ntasks + 1));
    c->ty = (uint32_t *)malloc(4 * ((size_t)ntasks + 1));
    if (!c->vx || !c->vy || !c->vout || !c->tx || !c->ty) return GR_E_INVALID_ARGUMENT;
    for (size_t i = 0; i < (size_t)nx * len * d; i++) set_w(c, c->vx, i, fw_to(&c->f, x[i]));
    for (size_t i = 0; i < (size_t)ny * len * d; i++) set_w(c, c->vy, i, fw_to(&c->f, y[i]));
    for (uint32_t t = 0; t < ntasks; t++) {
        if (tx[t] >= nx || ty[t] >= ny) return GR_E_INVALID_ARGUMENT;
        c->tx[t] = tx[t];
        c->ty[t] = ty[t];
    }
    c->vlen = len;
    c->ntasks = ntasks;
    return GR_OK;
}
void gr_avx2_dots_run(void *h, int threads) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    run_phase(c, 3, c->ntasks, threads);
}
void gr_avx2_dots_result(void *h, uint64_t *out) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < (size_t)c->ntasks * c->d; i++) out[i] = fw_from(&c->f, get_w(c, c->vout, i));
}

/* Independent commitments of witnesses drawn from a pool of polynomials
 * (converted once, outside the timed run). */
int gr_avx2_items_load(void *h, const uint64_t *pool, size_t pool_polys, const uint64_t *offs,
                       const uint32_t *lens, uint32_t count) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    size_t lb = limb_bytes(c), d = c->d;
    for (uint32_t i = 0; i < count; i++)
        if (lens[i] > c->n || offs[i] + lens[i] > pool_polys) return GR_E_DIMENSION_MISMATCH;
    free(c->pool); free(c->iout); free(c->ioff); free(c->ilen);
    c->pool = malloc(lb * d * (pool_polys + 1));
    c->iout = malloc(lb * d * ((size_t)count * c->rows + 1));
    c->ioff = (uint64_t *)malloc(8 * ((size_t)count + 1));
    c->ilen = (uint32_t *)malloc(4 * ((size_t)count + 1));
    if (!c->pool || !c->iout || !c->ioff || !c->ilen) return GR_E_INVALID_ARGUMENT;
    for (size_t i = 0; i < pool_polys * d; i++) set_w(c, c->pool, i, fw_to(&c->f, pool[i]));
    memcpy(c->ioff, offs, 8 * (size_t)count);
    memcpy(c->ilen, lens, 4 * (size_t)count);
    c->nitems = count;
    return GR_OK;
}
void gr_avx2_items_run(void *h, int threads) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    run_phase(c, 4, c->nitems, threads);
}
void gr_avx2_items_result(void *h, uint64_t *out) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < (size_t)c->nitems * c->rows * c->d; i++)
        out[i] = fw_from(&c->f, get_w(c, c->iout, i));
}

/* Twisted forward (inverse = 0) or inverse transforms, in place, of the first
 * operand loaded by gr_avx2_negacyclic_load. */
void gr_avx2_xform_run(void *h, int inverse, int threads) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    run_phase(c, inverse ? 6 : 5, (uint32_t)c->pairs, threads);
}
void gr_avx2_xform_result(void *h, uint64_t *out) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < c->pairs * c->d; i++) out[i] = fw_from(&c->f, get_w(c, c->pa, i));
}

void gr_avx2_negacyclic_result(void *h, uint64_t *out) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < c->pairs * c->d; i++) out[i] = fw_from(&c->f, get_w(c, c->pout, i));
}

void gr_avx2_result(void *h, uint64_t *t_out) {
    avx_ctx_t *c = (avx_ctx_t *)h;
    for (size_t i = 0; i < (size_t)c->rows * c->d; i++) t_out[i] = fw_from(&c->f, get_w(c, c->t, i));
}

int gr_avx2_commit(uint64_t q, int limb_bits, uint32_t rows, uint32_t n, const uint64_t *a_ntt,
                   const uint64_t *s, uint64_t *t_out, int threads) {
    void *h = gr_avx2_prepare(q, limb_bits, rows, n, 64, a_ntt);
    if (!h) return GR_E_UNSUPPORTED_SIZE;
    gr_avx2_load_witness(h, s);
    gr_avx2_commit_run(h, threads);
    gr_avx2_result(h, t_out);
    gr_avx2_free(h);
    return GR_OK;
}
