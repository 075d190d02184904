/*
 * gridref — CPU restatement of the Gridland ring-arithmetic path (the ORACLE).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against, and the timed CPU baseline in bench.py. Nothing in the
 * product (lattica_amd/) links, loads or calls it.
 *
 * Every function restates the reference algorithm named in its comment
 * (paths relative to the reference checkout). Values crossing this header are
 * canonical residues in [0, q) stored as uint64_t, exactly what the reference's
 * `IntegerRing::to_canonical()` returns (algebra/src/arith/prime.rs:529-531).
 * Internally the field works in Montgomery form with R = 2^64, the default
 * `PrimeField<Q, L = u64>` (algebra/src/arith/prime.rs:24-29).
 *
 * Status codes are shared with include/gridhip.h (GH_OK = 0, negatives on error).
 */
#ifndef GRIDREF_H
#define GRIDREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_OK 0
#define GR_E_LENGTH_MISMATCH (-1)
#define GR_E_LENGTH_NOT_POW2 (-2)
#define GR_E_UNSUPPORTED_SIZE (-3)
#define GR_E_UNSUPPORTED_TWIST (-4)
#define GR_E_INVALID_PARAMETERS (-10)
#define GR_E_DIMENSION_MISMATCH (-11)
#define GR_E_INVALID_ARGUMENT (-20)

/* ---- Z_q scalar arithmetic (prime.rs:61-226) on canonical values ---------- */
uint64_t gr_add(uint64_t q, uint64_t a, uint64_t b);
uint64_t gr_sub(uint64_t q, uint64_t a, uint64_t b);
uint64_t gr_mul(uint64_t q, uint64_t a, uint64_t b);
/* Ring slice hooks (prime.rs:433-491): op 0 add, 1 sub, 2 scalar mul,
 * 3 add scaled; canonical words */
int gr_slice_op(uint64_t q, int op, uint64_t *dst, const uint64_t *src, uint64_t scalar,
                size_t len);
uint64_t gr_dot_product(uint64_t q, const uint64_t *a, const uint64_t *b, size_t len);
uint64_t gr_pow(uint64_t q, uint64_t a, uint64_t e);
uint64_t gr_inv(uint64_t q, uint64_t a);
/* raw Montgomery hooks, so tests can pin the R = 2^64 words themselves */
uint64_t gr_to_montgomery(uint64_t q, uint64_t a);
uint64_t gr_from_montgomery(uint64_t q, uint64_t a_mont);
uint64_t gr_q_inv_neg(uint64_t q); /* -q^{-1} mod 2^64, prime.rs:82-91 */

/* ---- roots (prime.rs:566-607, :645-687) ----------------------------------- */
int gr_find_generator(uint64_t q, uint64_t *g_out);
int gr_root_of_unity(uint64_t q, uint64_t n, uint64_t *root_out);
uint64_t gr_max_ntt_size(uint64_t q);

/* ---- transforms (arith/ntt.rs, poly/ntt.rs) -------------------------------
 * Each acts in place on `count` consecutive polynomials of length `len`.     */
int gr_ntt_forward(uint64_t q, uint64_t len, uint64_t *values, size_t count);
int gr_ntt_inverse(uint64_t q, uint64_t len, uint64_t *values, size_t count);
int gr_twisted_forward(uint64_t q, uint64_t len, uint64_t *values, size_t count);
int gr_twisted_inverse(uint64_t q, uint64_t len, uint64_t *values, size_t count);
/* out[i] = a[i] * b[i] mod (X^len + 1), one twisted multiply per pair
 * (poly/ntt.rs:110-126, poly/ring.rs:83-101). out may alias a. */
int gr_negacyclic_mul(uint64_t q, uint64_t len, const uint64_t *a, const uint64_t *b,
                      uint64_t *out, size_t count);
/* schoolbook product, the reference's naive oracle (poly/ring.rs:60-77) */
int gr_negacyclic_mul_naive(uint64_t q, uint64_t len, const uint64_t *a, const uint64_t *b,
                            uint64_t *out, size_t count);

/* ---- linear (Ajtai) commitment --------------------------------------------
 * a_msg: rows x msg_len polys, a_open: rows x open_len polys, row-major
 * (lattice/types.rs:230-234), each poly `len` coefficients.
 * `key_domain` 0: a_* are coefficient-domain polys; 1: twisted-NTT evals.
 * t_out: rows polys, coefficient domain.
 * Prepared path: linear.rs:717-770 -> :509-529 -> :476-506, dot product
 * twisted_ntt.rs:378-408, finish linear.rs:361-375.                          */
int gr_commit_prepared(uint64_t q, uint64_t len, uint32_t rows, const uint64_t *a_msg,
                       uint32_t msg_len, const uint64_t *a_open, uint32_t open_len,
                       int key_domain, const uint64_t *message, const uint64_t *randomness,
                       uint64_t *t_out);
/* Coefficient path: per-row Ring::dot_product over CyclotomicPolyRing, one
 * negacyclic multiply per term (linear.rs:476-506, arith/ring.rs:113-121). */
int gr_commit_coeff(uint64_t q, uint64_t len, uint32_t rows, const uint64_t *a_msg,
                    uint32_t msg_len, const uint64_t *a_open, uint32_t open_len,
                    const uint64_t *message, const uint64_t *randomness, uint64_t *t_out);

/* ---- LaBRADOR Gram g_ij = <s_i, s_j>, i <= j, row-major upper triangle ----
 * s: r parts x n polys x len coeffs. g_out: r(r+1)/2 polys.
 * garbage.rs:76-93, index main_protocol/mod.rs:664-670. `ntt_once` = 0 runs the
 * reference-faithful coefficient-domain dot product (one NTT multiply per
 * term); 1 transforms each s_i once and accumulates in the NTT domain (same
 * values, labelled variant for the CPU baseline).                           */
int gr_gram(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *s,
            uint64_t *g_out, int ntt_once);

/* ---- LaBRADOR step around the commitment ---------------------------------
 * gr_decompose_balanced: decompose_polys / decompose_poly_balanced
 *   (main_protocol/mod.rs:258-320, :731-745): out[(poly*limbs + l)*len + k],
 *   limbs centred in [-base/2, base/2] and stored mod q.
 * gr_garbage_h: compute_garbage_h (main_protocol/garbage.rs:101-133),
 *   h_ij = (<phi_i, s_j> + <phi_j, s_i>) / 2, i <= j, row-major upper triangle.
 * gr_amortize: compute_amortized_witness (main_protocol/amortization.rs:61-84),
 *   z[k] = sum_i c_i * s_i[k].                                                 */
int gr_decompose_balanced(uint64_t q, uint64_t len, size_t npolys, const uint64_t *in,
                          uint64_t base, uint32_t limbs, uint64_t *out);
int gr_garbage_h(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *phi,
                 const uint64_t *s, uint64_t *h_out);
int gr_amortize(uint64_t q, uint64_t len, uint32_t r, uint32_t n, const uint64_t *c,
                const uint64_t *s, uint64_t *z_out);

/* ---- JL projection (labrador/src/jl.rs:388-508) ---------------------------
 * parts: num_parts x cols canonical residues, col = poly*len + coeff.
 * p_modq[rows]: the reference output (sum mod q).
 * p_exact[rows] (nullable): the same sum over centred lifts (x > q/2 -> x-q).
 * flat (nullable): rows*num_parts*cols int8 entries, layout
 *   [m][part][coeff][poly] (jl.rs:262-264).                                  */
int gr_jl_project(uint64_t q, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                  uint64_t cols, uint64_t len, const uint64_t *parts, uint64_t *p_modq,
                  int64_t *p_exact, int8_t *flat);
/* RFC 8439 ChaCha20 block function (IETF: 32-bit counter, 96-bit nonce). */
void gr_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                       uint8_t out[64]);

/* ---- AVX2 restatement of the reference's SIMD CPU path (CPU baseline) -----
 * Prepared Ajtai commit t = A_ntt * s with s in the coefficient domain, and
 * negacyclic products, using the reference's AVX2 kernel structure:
 *   limb 64: PrimeField<Q, u64> (the default type) — u64_arith.rs:360-539 for
 *            q < 2^32, the wide Montgomery path u64_arith.rs:193-357 for
 *            larger q (Goldilocks)
 *   limb 32: PrimeField<Q, u32>, q < 2^32          — u32_arith.rs:28-315
 * a_ntt: rows x n polys (NTT domain, canonical), s: n polys (coefficients).
 * threads > 1 splits the rows over OpenMP-free pthreads (the build's own
 * parallelisation; the reference is single-threaded).                       */
int gr_avx2_available(void);
int gr_avx2_commit(uint64_t q, int limb_bits, uint32_t rows, uint32_t n, const uint64_t *a_ntt,
                   const uint64_t *s, uint64_t *t_out, int threads);
/* Split form for timing: prepare converts the key once (setup, untimed);
 * load_witness converts s into limb words (the caller's own ring elements);
 * commit_run is the prepared commit itself (NTT(s), MAC, INTT). */
void *gr_avx2_prepare(uint64_t q, int limb_bits, uint32_t rows, uint32_t n, uint32_t d,
                      const uint64_t *a_ntt);
void gr_avx2_load_witness(void *h, const uint64_t *s);
void gr_avx2_commit_run(void *h, int threads);
void gr_avx2_result(void *h, uint64_t *t_out);
void gr_avx2_free(void *h);
/* commits of a witness of n_use <= n polynomials against the key's first
 * n_use columns (rows = 0 and n = 0 make a handle for products only) */
int gr_avx2_set_len(void *h, uint32_t n_use);
void gr_avx2_load_witness_len(void *h, const uint64_t *s, uint32_t n_use);
/* negacyclic products of count pairs of d-coefficient polynomials (canonical
 * in and out; load converts once, run is the timed part) */
int gr_avx2_negacyclic_load(void *h, const uint64_t *a, const uint64_t *b, size_t count);
void gr_avx2_xform_run(void *h, int inverse, int threads);
void gr_avx2_xform_result(void *h, uint64_t *out);
void gr_avx2_negacyclic_run(void *h, int threads);
void gr_avx2_negacyclic_result(void *h, uint64_t *out);
/* sums over k of coefficient-domain negacyclic products x[tx[t]][k] * y[ty[t]][k]
 * (Ring::dot_product / RingMat::mul_slice of CyclotomicPolyRing) */
int gr_avx2_dots_load(void *h, const uint64_t *x, uint32_t nx, const uint64_t *y, uint32_t ny,
                      uint32_t len, const uint32_t *tx, const uint32_t *ty, uint32_t ntasks);
void gr_avx2_dots_run(void *h, int threads);
void gr_avx2_dots_result(void *h, uint64_t *out);
/* independent commitments, one per item: witness = pool polys [offs[i],
 * offs[i] + lens[i]) against the key's first lens[i] columns; run splits the
 * items over threads (the build's own parallelisation) */
int gr_avx2_items_load(void *h, const uint64_t *pool, size_t pool_polys, const uint64_t *offs,
                       const uint32_t *lens, uint32_t count);
void gr_avx2_items_run(void *h, int threads);
void gr_avx2_items_result(void *h, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
