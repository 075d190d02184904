/*
 * gridhip.h — C ABI of the MI355X ring-arithmetic library (libgridhip.so).
 *
 * This is the drop-in boundary for Gridland's word-size ring path. Each entry
 * point names the reference interface it replaces (paths relative to the
 * reference checkout). A Rust caller binds these through a thin `extern "C"`
 * crate; the backend arms sit inside the reference's existing seams:
 *   NTTRing / Ring slice hooks (algebra/src/arith/ntt.rs:341-384,
 *     algebra/src/arith/ring.rs:75-135), dispatched today by GRID_SIMD
 *     (algebra/src/simd/dispatch.rs:5-66);
 *   PreparedLinearOps (commit/src/linear.rs:576-631, impl :690-939), whose
 *     type-erased cache (linear.rs:31) holds a gh_key;
 *   the LaBRADOR hot calls inner_commit_slice (labrador/src/crs.rs:211-216),
 *     compute_garbage_g (labrador/src/main_protocol/garbage.rs:76-93) and the
 *     JL projection (labrador/src/jl.rs:388-508).
 *
 * Conventions
 *  - Values are canonical residues in [0, q). A ring's word is uint32_t when
 *    q < 2^32 and uint64_t otherwise (gh_ring_query().word_bytes); arrays are
 *    packed words, polynomial-major, `d` words per polynomial.
 *  - Polynomial matrices are row-major of polynomials, as RingMat
 *    (algebra/src/lattice/types.rs:230-234).
 *  - NTT-domain values use the reference's slot order: slot k holds
 *    a(psi^(2k+1)) (TwistedNttPoly, algebra/src/poly/twisted_ntt.rs:22-27).
 *  - `mem` says where caller buffers live: GH_MEM_HOST (copied in/out, the
 *    call returns when results are in host memory) or GH_MEM_DEVICE (device
 *    pointers from gh_alloc or any HIP allocation on the context's device;
 *    work is enqueued on the context stream and the call returns without
 *    waiting — gh_ctx_synchronize() waits). Device results are complete once
 *    gh_ctx_synchronize() returns, and visible to every later call on the
 *    context; a device-memory gh_commit (or single-item gh_commit_batch) may
 *    hold its last step back until that point, so hipDeviceSynchronize()
 *    alone does not finish it (tuning key "defer_finalize" = 0 turns this
 *    off). A later commitment whose witness is that held-back output (a
 *    commitment of a commitment) completes it first.
 *  - Return value: GH_OK (0) or a negative gh_status. Errors that the
 *    reference reports map 1:1 onto NttError (arith/ntt.rs:40-49) and
 *    CommitmentError (commit/src/error.rs:7-18).
 *  - Device objects are library-owned; a context is used from one host thread
 *    at a time.
 */
#ifndef GRIDHIP_H
#define GRIDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GH_API __attribute__((visibility("default")))

typedef enum gh_status {
  GH_OK = 0,
  /* NttError (algebra/src/arith/ntt.rs:40-49) */
  GH_E_LENGTH_MISMATCH = -1,
  GH_E_LENGTH_NOT_POW2 = -2,
  GH_E_UNSUPPORTED_SIZE = -3,
  GH_E_UNSUPPORTED_TWIST = -4,
  /* CommitmentError (commit/src/error.rs:7-18) */
  GH_E_INVALID_PARAMETERS = -10,
  GH_E_DIMENSION_MISMATCH = -11,
  GH_E_OPENING_NORM_EXCEEDED = -12,
  GH_E_INVALID_MESSAGE_ENCODING = -13,
  GH_E_INVALID_OPENING = -14,
  /* boundary / device */
  GH_E_INVALID_ARGUMENT = -20, /* null handle or pointer, bad flag */
  GH_E_NOT_CANONICAL = -21,    /* an input word >= q */
  GH_E_UNSUPPORTED_MODULUS = -22, /* q has a root of the right order but no GPU field path */
  GH_E_DEVICE = -23,           /* HIP runtime error */
  GH_E_OOM = -24,              /* device allocation failed */
  GH_E_BUSY = -25,             /* server queue full */
  GH_E_NO_DEVICE = -26,
  /* SerializationError (serialize/src/lib.rs:28-34) */
  GH_E_UNEXPECTED_END = -30, /* encoded data ends before what its headers announce */
  GH_E_INVALID_DATA = -31    /* element >= q, shape overflow, trailing bytes, inconsistent key */
} gh_status;

#define GH_MEM_HOST 0
#define GH_MEM_DEVICE 1

#define GH_DOMAIN_COEFF 0 /* coefficient domain (CyclotomicPolyRing) */
#define GH_DOMAIN_NTT 1   /* twisted-NTT evaluations (TwistedNttPoly) */

#define GH_RING_Q32 1
#define GH_RING_GOLDILOCKS 2

typedef struct gh_ctx gh_ctx;
typedef struct gh_ring gh_ring;
typedef struct gh_key gh_key;

/* ---- lifecycle ---------------------------------------------------------- */
GH_API const char *gh_status_string(int status);
GH_API int gh_version(void);
GH_API int gh_device_count(int *count);
GH_API int gh_ctx_create(int device, gh_ctx **out);
GH_API int gh_ctx_destroy(gh_ctx *ctx);
GH_API int gh_ctx_synchronize(gh_ctx *ctx);
/* number of compute units of the context's device (for batch sizing) */
GH_API int gh_ctx_num_cus(gh_ctx *ctx, int *out);

/* Kernel timing with HIP events recorded on the context stream around the
 * phases of each commitment (observability; the reference has none). While
 * enabled, every gh_commit records events around its witness transform
 * (phase 0), the streaming MAC (phase 1) and the finalize/inverse (phase 2).
 * gh_ctx_profile_read waits for the stream and returns the summed elapsed
 * milliseconds and launch count of a phase since the last read. */
GH_API int gh_ctx_profile_enable(gh_ctx *ctx, int on);
GH_API int gh_ctx_profile_read(gh_ctx *ctx, int phase, double *total_ms, uint64_t *count);

/* Launch-shape tuning of a context (written by the autotuner, read back for
 * reports). Keys: "mac_variant" (0 dword / 1 16-byte loads, q32 MAC),
 * "mac_waves" (4, 8, 16), "mac_blocks_per_cu" (1..8), "quad_max" (largest
 * witness batch transformed four lanes per polynomial), "fused" (2, the
 * default: q32 commitments transform the witness inside the MAC kernel, whose
 * waves each own four rows of a 16-row block — row counts not divisible by 16
 * take 1; 1: the same kernel with the waves splitting the block's columns; 0:
 * separate transform and MAC kernels), "fused_depth" (1..4: key columns in
 * flight per wave, 8 * depth for fused 2 and max(1, depth * 8 / mac_waves)
 * for fused 1),
 * "batch_chunk_cols" and "batch_chunk_cols_gold" (16..4096, key columns per
 * block of gh_commit_batch and of the server's waves, for q32 and Goldilocks
 * keys; halved for batches too small to fill the device), "batch_group" (4, 8,
 * 16: witnesses that share each loaded key column), "defer_finalize" (1,
 * the default: a device-memory gh_commit or single-item gh_commit_batch
 * leaves its last step to the next call on the context, so consecutive
 * commitments cost one launch each; 0: that step is enqueued before the call
 * returns, so the result is complete in stream order and a plain
 * hipStreamSynchronize or hipDeviceSynchronize finishes it; setting 0 also
 * enqueues a step held back until then), "stage1_tasks_per_block" (-1, the
 * default: the first half of a held-back step runs in about num_cus / 2
 * leading blocks of the next launch; 0: in its MAC blocks; 1..8: in leading
 * blocks of that many tasks), "batch_fused" (0, the default: a Goldilocks
 * gh_commit_batch or batch plan runs a transform, a grouped MAC and a
 * finalize kernel; 1: one launch that transforms, multiplies and finalizes,
 * slower on the batches measured), "batch_fused_cols" (32..4096, a
 * multiple of 32: key columns per block of that launch) and "batch_mac_ring"
 * (4, the default, down to 2: the grouped MAC of gh_commit_batch and the
 * server stages that many key columns per wave in an LDS ring filled by
 * global-to-LDS loads, for groups of 4 and row counts divisible by 16; 0: the
 * kernels with one column of loads in flight in registers), "q32_ring_group"
 * (4, the default, or 8: witnesses sharing each key column in that ring MAC
 * for q32 keys; Goldilocks keys use 4) and
 * "negacyclic_quad" (-1, the default: gh_negacyclic_mul runs one pair per
 * quad of lanes, or for q < 2^32 one pair per thread when "quad_lds" is 0;
 * 0 and 1: per thread or per quad for both) and "xform_quad" (1, the default: the
 * twisted transforms run four lanes per polynomial at every size; 0: the
 * forward does up to "quad_max" polynomials that way and the inverse a
 * polynomial per thread; -1: 1 for Goldilocks, 0 for q < 2^32) and "quad_lds"
 * (1, the default: the four-lane transforms finish their 4-point DFT over the
 * lanes through LDS; 0: with lane exchanges). Values are checked;
 * an unknown key or value gives GH_E_INVALID_ARGUMENT. Every setting yields
 * identical results; only the speed changes. */
GH_API int gh_ctx_set_tuning(gh_ctx *ctx, const char *key, int64_t value);
GH_API int gh_ctx_get_tuning(gh_ctx *ctx, const char *key, int64_t *value);

/* Achievable HBM bandwidth of the context's device: a streaming read and a
 * streaming copy over `bytes` (rounded down to 16 B) of freshly allocated
 * memory, best of `reps` timed kernels each, in GB/s (1e9 B/s; a copy counts
 * its read and its write). Up to 512 MiB, three source buffers are read in
 * turn so that none is served from the Infinity Cache. Measurement support for
 * the roofline report; no reference counterpart. */
GH_API int gh_measure_bandwidth(gh_ctx *ctx, size_t bytes, int reps, double *read_gbps,
                                double *copy_gbps);

/* Issue rates of the context's device: the best of three timed launches of
 * four probe kernels, in microseconds, on num_cus * waves_per_simd blocks of
 * four waves, each wave running 4096 iterations of
 *   out_us[0]: four independent v_mfma_i32_32x32x32_i8,
 *   out_us[1]: eight independent v_mfma_i32_16x16x64_i8 (the same int8
 *              multiply-accumulates as [0]),
 *   out_us[2]: four independent ChaCha20 quarter rounds (48 VALU
 *              instructions),
 *   out_us[3]: the work of [0] and [2] interleaved in one loop, which shows
 *              how much of the matrix-core time a wave's VALU work hides.
 * waves_per_simd 1..8. Measurement support for the JL projection's roofline;
 * no reference counterpart. */
GH_API int gh_measure_issue(gh_ctx *ctx, int waves_per_simd, double out_us[4]);

/* Issue rate of single VALU instruction kinds on the context's device, in G
 * wave instructions per second (best of three timed launches on num_cus *
 * waves_per_simd blocks of four waves, each lane running eight independent
 * chains of one instruction): out_gips[0] v_add_u32, [1] v_mul_hi_u32,
 * [2] v_mul_lo_u32, [3] v_mad_u64_u32, [4] v_lshl_add_u64 (64-bit add),
 * [5] v_fma_f64, [6] the modular correction v_sub_co_u32 + v_cndmask_b32 +
 * v_add_u32 (its three instructions counted), [7] v_cndmask_b32, [8] v_xor_b32,
 * [9] v_alignbit_b32. waves_per_simd 1..8. Measurement support for the VALU
 * rooflines of the transforms and multiply-accumulates; no reference
 * counterpart. */
GH_API int gh_measure_valu(gh_ctx *ctx, int waves_per_simd, double out_gips[10]);

/* device memory, so callers without a HIP runtime of their own can keep
 * inputs resident in HBM */
GH_API int gh_alloc(gh_ctx *ctx, size_t bytes, void **dev_ptr);
GH_API int gh_free(gh_ctx *ctx, void *dev_ptr);
GH_API int gh_upload(gh_ctx *ctx, void *dev_dst, const void *host_src, size_t bytes);
GH_API int gh_download(gh_ctx *ctx, void *host_dst, const void *dev_src, size_t bytes);

/* Checks that `words` ring words are canonical residues (< q). Device-memory
 * calls trust their buffers, like the reference's *_trusted paths
 * (commit/src/linear.rs:476-529); host-memory calls check before upload.
 * Counterpart of PrimeField's Valid impl (algebra/src/arith/prime.rs:723-728).
 * Returns GH_OK or GH_E_NOT_CANONICAL. */
GH_API int gh_check_canonical(gh_ring *ring, const void *values, size_t words, int mem);

/* ---- ring plan ----------------------------------------------------------
 * Replaces NttPlan::build / TwistedNttPlan::build and their plan caches
 * (algebra/src/arith/ntt.rs:124-171, algebra/src/poly/ntt.rs:140-178) with the
 * root selection of PrimeField (algebra/src/arith/prime.rs:566-607, :645-687).
 * Errors: LENGTH_NOT_POW2, UNSUPPORTED_SIZE (no d-th root, or d != 64 on the
 * GPU), UNSUPPORTED_TWIST (no 2d-th root), UNSUPPORTED_MODULUS (q >= 2^32 and
 * not Goldilocks). */
typedef struct gh_ring_info {
  uint64_t q;
  uint32_t d;
  uint32_t word_bytes; /* 4 or 8 */
  uint32_t kind;       /* GH_RING_Q32 / GH_RING_GOLDILOCKS */
  uint32_t wire_bytes; /* bytes of one encoded element: ceil(bits(q - 1) / 8) */
  uint64_t generator;  /* find_generator */
  uint64_t psi;        /* root_of_unity(2d): the twist */
  uint64_t omega;      /* root_of_unity(d) = psi^2 */
  uint64_t d_inv;      /* inverse_ntt_scale(d) */
} gh_ring_info;

GH_API int gh_ring_create(gh_ctx *ctx, uint64_t q, uint32_t d, gh_ring **out);
GH_API int gh_ring_destroy(gh_ring *ring);
GH_API int gh_ring_query(const gh_ring *ring, gh_ring_info *out);

/* ---- transforms: NTTRing seam and TwistedNttPlan -------------------------
 * In place on `count` polynomials of d words.
 * gh_ntt_forward / gh_ntt_inverse: NTTRing::ntt_forward_with_plan /
 *   ntt_inverse_with_plan, cyclic length-d NTT (arith/ntt.rs:292-335;
 *   simd/montgomery.rs:636-681).
 * gh_twisted_forward / gh_twisted_inverse: TwistedNttPlan::forward_in_place /
 *   inverse_in_place, negacyclic (poly/ntt.rs:55-104). */
GH_API int gh_ntt_forward(gh_ring *ring, void *values, size_t count, int mem);
GH_API int gh_ntt_inverse(gh_ring *ring, void *values, size_t count, int mem);
GH_API int gh_twisted_forward(gh_ring *ring, void *values, size_t count, int mem);
GH_API int gh_twisted_inverse(gh_ring *ring, void *values, size_t count, int mem);

/* out[i] = a[i] * b[i] mod (X^d + 1) for i < count: poly_mul_ntt /
 * NegacyclicMulRing::negacyclic_mul_coeffs (poly/ntt.rs:305-319,
 * poly/ring.rs:83-101), one fused forward/forward/pointwise/inverse pass.
 * `out` may alias `a` or `b`. */
GH_API int gh_negacyclic_mul(gh_ring *ring, const void *a, const void *b, void *out, size_t count,
                             int mem);
/* Pointwise product of NTT-domain polynomials, `Mul for TwistedNttPoly`
 * (poly/twisted_ntt.rs:570-581) / Ring::pointwise_mul_assign_slice. */
GH_API int gh_pointwise_mul(gh_ring *ring, const void *a, const void *b, void *out, size_t count,
                            int mem);

/* ---- Ring slice hooks ------------------------------------------------------
 * The remaining `Ring` slice hooks (arith/ring.rs:75-135), overridden for
 * PrimeField at arith/prime.rs:433-491 (simd/montgomery.rs add_assign :125,
 * sub_assign :257, scalar_mul :389). They act on `len` field elements (any
 * length, not a number of polynomials), canonical words of the ring. `src`
 * may equal `dst`; other overlaps are undefined. `scalar` is a canonical
 * residue (GH_E_NOT_CANONICAL otherwise).
 *   gh_add_assign:        dst[i] = dst[i] + src[i]           (add_assign_slice)
 *   gh_sub_assign:        dst[i] = dst[i] - src[i]           (sub_assign_slice)
 *   gh_scalar_mul:        dst[i] = dst[i] * scalar           (scalar_mul_slice)
 *   gh_add_assign_scaled: dst[i] = dst[i] + src[i] * scalar  (add_assign_scaled_slice)
 *   gh_dot_product:       out[0] = sum_i a[i] * b[i]         (dot_product), one
 *     word in `mem` (a device pointer for GH_MEM_DEVICE); 0 for len 0. */
GH_API int gh_add_assign(gh_ring *ring, void *dst, const void *src, size_t len, int mem);
GH_API int gh_sub_assign(gh_ring *ring, void *dst, const void *src, size_t len, int mem);
GH_API int gh_scalar_mul(gh_ring *ring, void *dst, uint64_t scalar, size_t len, int mem);
GH_API int gh_add_assign_scaled(gh_ring *ring, void *dst, const void *src, uint64_t scalar,
                                size_t len, int mem);
GH_API int gh_dot_product(gh_ring *ring, const void *a, const void *b, size_t len, void *out,
                          int mem);

/* ---- resident commitment key ---------------------------------------------
 * PreparedLinearOps::build_linear_key_cache (commit/src/linear.rs:701-708,
 * prepare_linear_key :377-395): A = [A_msg | A_open] kept NTT-domain in HBM in
 * the library's streaming layout. a_msg is rows x msg_len polynomials and
 * a_open rows x open_len (open_len may be 0), both row-major, in `domain`.
 * The key is what the Rust cache box holds; its Drop calls gh_key_destroy. */
GH_API int gh_key_create(gh_ring *ring, uint32_t rows, const void *a_msg, uint32_t msg_len,
                         const void *a_open, uint32_t open_len, int domain, int mem,
                         gh_key **out);
GH_API int gh_key_destroy(gh_key *key);
GH_API int gh_key_dims(const gh_key *key, uint32_t *rows, uint32_t *msg_len, uint32_t *open_len);

/* ---- plain commitment (the drop-in) --------------------------------------
 * PreparedLinearOps::recompute_linear_commitment_runtime with a prepared key
 * (commit/src/linear.rs:717-770 -> :509-529 -> :476-506, finish :361-375):
 *   t = A_msg * m + A_open * r over R_q, t coefficient-domain, rows polys.
 * message: msg_len polys, randomness: open_len polys (may be NULL when the
 * key has open_len 0), both in `domain` (GH_DOMAIN_NTT is the
 * commit_with_opening_ntt path, commit/src/ajtai/mod.rs:256-269).
 * DIMENSION_MISMATCH when lengths differ from the key (linear.rs:270-327).
 * With GH_MEM_DEVICE and open_len 0 the finalize (sum of the MAC partials and
 * the inverse NTT) is held back and runs inside the next such commitment's
 * launch over the same field (q32 or Goldilocks), or before any other call on
 * the context: back-to-back commitments cost one launch each. */
GH_API int gh_commit(gh_key *key, const void *message, uint32_t msg_len, const void *randomness,
                     uint32_t open_len, int domain, void *t_out, int mem);

/* ---- commitment that allocates nothing --------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Same result as gh_commit(key, message, msg_len, NULL, 0,
 * GH_DOMAIN_COEFF, t_out, GH_MEM_DEVICE): message and t_out are device
 * pointers, the message in the coefficient domain, and the key has open_len 0
 * and rows % 4 == 0. The partial sums go to one 4 MiB array per device
 * (static device memory the library holds until the process ends) instead of
 * the context workspace; nothing is allocated and the result is complete in
 * stream order (the finalize is never held back). A held-back finalize of an
 * earlier gh_commit is issued first. The array holds blocks x rows x 64 words
 * of partial sums, so the grid shrinks to fit it; the fewest blocks a launch
 * can have are ceil(msg_len / 128) for q32 and ceil(msg_len / 64) for
 * Goldilocks, and a key for which even those exceed 4 MiB gets
 * GH_E_INVALID_ARGUMENT (rows 16: msg_len above 131072 for q32, above 32768
 * for Goldilocks).
 * Every call on one device shares the array and each call overwrites the part
 * it uses: calls that overlap on one device (two streams, two contexts, two
 * threads without ordering) race; calls in one stream do not.
 * DIMENSION_MISMATCH when msg_len differs from the key. */
GH_API int gh_commit_noalloc(gh_key *key, const void *message, uint32_t msg_len, void *t_out);

/* ---- BDLOP commitment ------------------------------------------------------
 * BdlopCommitmentScheme::commit / verify recompute (commit/src/bdlop/mod.rs:34-54)
 * through PreparedLinearOps::recompute_bdlop_parts_runtime
 * (commit/src/linear.rs:800-938):
 *   u = A_mask * r (mask->rows polys), v = A_msg * m + A_open * r (key->rows polys),
 * coefficient domain. `mask` holds A_mask as a key of open_len message columns
 * (gh_key_create with msg_len = open_len, open_len = 0); `key` holds
 * [A_msg | A_open]. message: msg_len polys, randomness: open_len polys, both in
 * `domain`. DIMENSION_MISMATCH when the lengths disagree with the keys
 * (validate_bdlop_key, bdlop/mod.rs:56-64). */
GH_API int gh_bdlop_commit(gh_key *mask, gh_key *key, const void *message, uint32_t msg_len,
                           const void *randomness, uint32_t open_len, int domain, void *u_out,
                           void *v_out, int mem);

/* ---- batched commitments ---------------------------------------------------
 * Many independent commitments against one shared key in a single launch.
 * Item i commits witness_i (len_i polys, coefficient domain) against the first
 * len_i columns of the key: t_i = A[:, :len_i] * w_i. All pointers in one call
 * share `mem`. The per-item result equals gh_commit on the same inputs with a
 * key of len_i columns. A single device-memory item takes gh_commit's path
 * (its finalize may be held back the same way). No reference counterpart: the
 * reference commits one message per call (commit/src/ajtai/mod.rs:443-452). */
typedef struct gh_commit_item {
  const void *witness;
  uint32_t len;
  uint32_t reserved;
  void *t_out;
} gh_commit_item;
GH_API int gh_commit_batch(gh_key *key, const gh_commit_item *items, size_t count, int mem);

/* Batch plans: a fixed batch of device-memory items whose kernels (segmented
 * witness transform, grouped multiply-accumulate, batched finalize) are
 * recorded once into a HIP graph over plan-owned device memory. Every
 * gh_batch_plan_launch replays the graph on the context stream: one launch
 * per batch, recomputing each t_i from the current contents of its witness
 * buffer (callers refill witnesses in place between launches). Results equal
 * gh_commit_batch on the same items. The key must have rows % 4 == 0; the
 * launch-shape tuning in force at creation is the one recorded. No reference
 * counterpart (see gh_commit_batch). */
typedef struct gh_batch_plan gh_batch_plan;
GH_API int gh_batch_plan_create(gh_key *key, const gh_commit_item *items, size_t count,
                                gh_batch_plan **out);
GH_API int gh_batch_plan_launch(gh_batch_plan *plan);
GH_API int gh_batch_plan_destroy(gh_batch_plan *plan);

/* ---- LaBRADOR Gram matrix --------------------------------------------------
 * compute_garbage_g (labrador/src/main_protocol/garbage.rs:76-93):
 *   g_ij = <s_i, s_j> = sum_k s_i[k] * s_j[k] over R_q, 0 <= i <= j < r,
 * written row-major over the upper triangle (garbage_index,
 * main_protocol/mod.rs:664-670): r(r+1)/2 polynomials, coefficient domain.
 * s: r parts x n polys, in `domain`. */
GH_API int gh_gram(gh_ring *ring, const void *s, uint32_t r, uint32_t n, int domain, void *g_out,
                   int mem);

/* ---- LaBRADOR inner commitments --------------------------------------------
 * CommitKey::inner_commit_slice (labrador/src/crs.rs:211-216) for r parts:
 * t_i = A * s_i, s: r x n polys (key msg_len == n, open_len == 0),
 * t_out: r x rows polys, coefficient domain. */
GH_API int gh_inner_commit(gh_key *key, const void *s, uint32_t r, uint32_t n, int domain,
                           void *t_out, int mem);

/* ---- LaBRADOR step around the commitment -------------------------------------
 * gh_decompose_balanced: decompose_polys / decompose_poly_balanced
 *   (labrador/src/main_protocol/mod.rs:258-320, :731-745): each coefficient,
 *   centred (x > q/2 -> x - q), is split into `limbs` base-`base` digits in
 *   [-base/2, base/2] (centered_mod, truncating division), stored mod q;
 *   out: count x limbs x d words, [poly][limb][coeff]. Digits beyond `limbs`
 *   are dropped, as in the reference. base in [2, 2^32], limbs >= 1, else
 *   GH_E_INVALID_PARAMETERS.
 * gh_garbage_h: compute_garbage_h (labrador/src/main_protocol/garbage.rs:101-133):
 *   h_ij = (<phi_i, s_j> + <phi_j, s_i>) * 2^-1, 0 <= i <= j < r, row-major
 *   upper triangle as gh_gram; phi, s: r x n polys in `domain`.
 * gh_amortize: compute_amortized_witness (main_protocol/amortization.rs:61-84):
 *   z[k] = sum_i c_i * s_i[k]; c: r polys, s: r x n polys, both in `domain`;
 *   z: n polys, coefficient domain.
 * The outer commitments u1 = B*t + C*g and u2 = D*h (labrador/src/crs.rs:273-304)
 * are gh_commit with a key [B | C] (message t, randomness g) and a key D. */
GH_API int gh_decompose_balanced(gh_ring *ring, const void *values, size_t count, uint64_t base,
                                 uint32_t limbs, void *out, int mem);
GH_API int gh_garbage_h(gh_ring *ring, const void *phi, const void *s, uint32_t r, uint32_t n,
                        int domain, void *h_out, int mem);
GH_API int gh_amortize(gh_ring *ring, const void *c, const void *s, uint32_t r, uint32_t n,
                       int domain, void *z_out, int mem);

/* ---- JL projection -----------------------------------------------------------
 * JLMatrix::project_and_extract_jl_rows (labrador/src/jl.rs:267-303 ->
 * project_and_extract_seq :388-508). Part p's matrix is the ChaCha20 keystream
 * (RFC 8439, key = seed, nonce = LE u32(p) || 0^8, counter from 0), ceil(rows/4)
 * bytes per column, 2 bits per row: 00 -> +1, 01 -> -1, else 0.
 * parts: num_parts x cols residues (col = poly*d + coeff).
 * p_modq: rows ring words, the reference's p = sum mod q.
 * p_exact (nullable): rows int64, the same sum over centred lifts
 *   (x > q/2 -> x - q), exact in 64-bit integers; q < 2^32 only, and fewer
 *   than 2^26 polynomials (num_parts * cols / d), so that
 *   |p| <= total * d * (q - 1) / 2 < 2^63 (GH_E_INVALID_ARGUMENT otherwise).
 * flat (nullable): rows*num_parts*cols int8 entries in the reference's flat
 *   layout [m][part][coeff][poly] (jl.rs:262-264). */
GH_API int gh_jl_project(gh_ring *ring, const uint8_t seed[32], uint32_t rows, const void *parts,
                         uint32_t num_parts, uint64_t cols, void *p_modq, int64_t *p_exact,
                         int8_t *flat, int mem);

/* Launch shape of the int8 matrix-core JL kernels for `total_polys` witness
 * polynomials on a device of `num_cus` compute units (host arithmetic only;
 * no device or context needed). Their int32 sums are exact: every product is
 * a JL entry in {-1, 0, 1} times a signed base-256 digit in [-128, 127], and
 * each int32 sum takes at most `terms` of them, so |sum| <= bound = terms *
 * 128 <= 2^31 - 1. kernel: GH_JL_KERNEL_PROJECT (gh_jl_project at 253..256
 * rows; per_unit = polynomials per wave, four waves per block) or
 * GH_JL_KERNEL_EVALS (gh_jl_evals; per_unit = polynomials per block; the
 * witness runs in `chunks` launches of at most chunk_polys polynomials,
 * `blocks` blocks in the first). total_polys <= 2^36. No reference
 * counterpart: it states the bound the matrix-core path relies on. */
enum { GH_JL_KERNEL_PROJECT = 0, GH_JL_KERNEL_EVALS = 1 };
typedef struct gh_jl_shape {
  uint64_t chunk_polys, per_unit, blocks, chunks, terms, bound;
} gh_jl_shape;
GH_API int gh_jl_launch_shape(int kernel, uint64_t total_polys, int num_cus, gh_jl_shape *out);

/* ---- JL rows for aggregation --------------------------------------------------
 * gh_jl_extract_flat: JLMatrix::extract_jl_rows_flat (labrador/src/jl.rs:311-383):
 *   the entries alone, rows * num_parts * cols int8 in the layout
 *   [m][part][coeff][poly] (what gh_jl_project's `flat` holds); rows <= 256.
 * gh_jl_conjugate: jl_rows_flat_to_conjugated_polys
 *   (labrador/src/main_protocol/mod.rs:885-923): rows * num_parts * n polynomials
 *   (n = cols / d) laid out [m][part][poly]; coefficient k is e(0) for k = 0
 *   and -e(d - k) otherwise, e(k) = Pi[m][part][poly * d + k] as a residue:
 *   sigma_{-1} of the row's polynomial (labrador/src/relation.rs:636-655).
 *   `flat` (layout above, in `mem`) is the input; flat NULL draws the
 *   entries from `seed` instead, with the same result and no int8 array.
 * gh_jl_evals: the prover's JL evaluations (labrador/src/main_protocol/step_prover.rs:196-205,
 *   last_level/prover.rs:187-196): jl_eval[m] = sum_i sum_j
 *   sigma_{-1}(pi_i^(m))_j * s_ij over R_q, for rows = 256, the count
 *   JLProfile::validate requires (labrador/src/params/jl_profile.rs:47-62);
 *   parts: num_parts x cols words (cols = n * d), coefficient domain; evals:
 *   256 polynomials, coefficient domain. Their constant terms are the p of
 *   gh_jl_project.
 * gh_jl_aggregate: the JL term of the aggregated function, prover and verifier
 *   (labrador/src/main_protocol/aggregation.rs:237-245, called from
 *   step_prover.rs:187 and step_verifier.rs:359 through aggregate_step2_flat):
 *   phi[i][j] += sum_m scale_m * sigma_{-1}(pi_i^(m))_j over R_q, for rows =
 *   256. scale: 256 polynomials, the loop's jl_scale_m = sum_k beta_k *
 *   omega_k[m] formed by the caller; phi: num_parts x n polynomials (n = cols
 *   / d), read and written in place (the caller's agg_phi). Both coefficient
 *   domain. The conjugated polynomials are never materialised.
 * All four need d = 64 and cols % d == 0. */
GH_API int gh_jl_extract_flat(gh_ring *ring, const uint8_t seed[32], uint32_t rows,
                              uint32_t num_parts, uint64_t cols, int8_t *flat, int mem);
GH_API int gh_jl_conjugate(gh_ring *ring, const uint8_t seed[32], const int8_t *flat,
                           uint32_t rows, uint32_t num_parts, uint64_t cols, void *polys,
                           int mem);
GH_API int gh_jl_evals(gh_ring *ring, const uint8_t seed[32], uint32_t rows, const void *parts,
                       uint32_t num_parts, uint64_t cols, void *evals, int mem);
GH_API int gh_jl_aggregate(gh_ring *ring, const uint8_t seed[32], uint32_t rows,
                           const void *scale, uint32_t num_parts, uint64_t cols, void *phi,
                           int mem);

/* ---- canonical wire formats ----------------------------------------------------
 * The reference's canonical encodings (serialize/src/lib.rs:1-7), byte for byte:
 *   element     PrimeField: gh_ring_info.wire_bytes = ceil(bits(q - 1) / 8) bytes,
 *               little-endian, value < q (algebra/src/arith/prime.rs:54-58, :690-721);
 *   polynomial  d elements, no header: CyclotomicPolyRing coefficients
 *               (poly/ring.rs:469-505) or TwistedNttPoly evaluations in the
 *               library's slot order (poly/twisted_ntt.rs:647-685).
 * Containers, by `kind`; an item is a polynomial except for GH_WIRE_ELEMS:
 *   GH_WIRE_POLYS    `rows` polynomials back to back ([T; N], struct fields);
 *   GH_WIRE_RINGVEC  RingVec: u64 count, then the items (lattice/types.rs:183-219);
 *   GH_WIRE_RINGMAT  RingMat: u64 rows, u64 cols, then rows x cols items row-major
 *                    (lattice/types.rs:440-490);
 *   GH_WIRE_VEC      Vec<polynomial>: u32 count, then the items (serialize/src/lib.rs:210-245);
 *   GH_WIRE_ELEMS    Vec<element>: u32 count, then bare field elements.
 * Shapes are rows x cols items; cols is 1 for every kind but GH_WIRE_RINGMAT.
 * Items in memory are ring words (rows*cols*d of them, or rows for ELEMS) in
 * `mem`; encoded bytes are always host memory. Elements are packed and unpacked
 * by device kernels; headers are written and read on the host.
 *
 * gh_wire_size:   encoded bytes of a shape; INVALID_DATA when it overflows or a
 *                 Vec length exceeds u32.
 * gh_wire_encode: items at `src` -> `out` (at most out_cap bytes); *written is
 *                 the encoded size (out NULL: size only). Device items are
 *                 trusted canonical; host items are checked (NOT_CANONICAL).
 * gh_wire_peek:   the header of `bytes`: rows, cols and the total encoded size
 *                 (GH_WIRE_POLYS has no header: *rows is read as the count).
 * gh_wire_decode: `bytes` -> items at `dst`, for the shape in the header, which
 *                 must be rows x cols (else DIMENSION_MISMATCH; for
 *                 GH_WIRE_POLYS rows is the count). Errors in the reference's
 *                 order: INVALID_DATA for an element >= q or a shape that
 *                 overflows; UNEXPECTED_END when the bytes end early, after the
 *                 elements present were checked (every whole element in the
 *                 coefficient domain, as CyclotomicPolyRing decodes element by
 *                 element; whole polynomials only in the NTT domain, as
 *                 TwistedNttPoly checks its length first). `domain` names what
 *                 the polynomials hold (ignored for ELEMS). consumed NULL asks
 *                 for exact consumption (deserialize_exact, serialize/src/lib.rs:79-83):
 *                 trailing bytes give INVALID_DATA; else *consumed is the bytes
 *                 used. On error a device `dst` may be partly written; a host
 *                 `dst` is untouched. */
#define GH_WIRE_POLYS 0
#define GH_WIRE_RINGVEC 1
#define GH_WIRE_RINGMAT 2
#define GH_WIRE_VEC 3
#define GH_WIRE_ELEMS 4
GH_API int gh_wire_size(const gh_ring *ring, int kind, uint64_t rows, uint64_t cols,
                        uint64_t *bytes);
GH_API int gh_wire_encode(gh_ring *ring, int kind, uint64_t rows, uint64_t cols, const void *src,
                          int mem, void *out, uint64_t out_cap, uint64_t *written);
GH_API int gh_wire_peek(const gh_ring *ring, int kind, const void *bytes, uint64_t len,
                        uint64_t *rows, uint64_t *cols, uint64_t *total);
GH_API int gh_wire_decode(gh_ring *ring, int kind, int domain, const void *bytes, uint64_t len,
                          uint64_t rows, uint64_t cols, void *dst, int mem, uint64_t *consumed);

/* ---- prepared-key persistence ---------------------------------------------------
 * The encoding of LinearCommitmentKey (commit/src/linear.rs:169-186): RingMat
 * a_msg (rows x msg_len) then RingMat a_open (rows x open_len).
 * gh_key_save writes the resident key in `domain`: GH_DOMAIN_NTT gives the
 * TwistedNttPoly matrices a prepared key holds (PreparedLinearCommitmentKey,
 * linear.rs:162-167, :377-395), GH_DOMAIN_COEFF the coefficient-domain key
 * (finish_twisted_ring_mat, poly/twisted_ntt.rs:318-330). out NULL: size only.
 * gh_key_load rebuilds a resident key from either form, as gh_key_create
 * would from the matrices (a coefficient-domain key is transformed once).
 * Decoding errors as gh_wire_decode; then row counts that differ give
 * INVALID_DATA (deserialize_and_validate, linear.rs:188-192, serialize/src/lib.rs:106-110),
 * and an empty message block or row count INVALID_PARAMETERS
 * (CommitmentDimensions::validate, linear.rs:45-59). */
GH_API int gh_key_save(gh_key *key, int domain, void *out, uint64_t out_cap, uint64_t *written);
GH_API int gh_key_load(gh_ring *ring, const void *bytes, uint64_t len, int domain,
                       uint64_t *consumed, gh_key **out);

/* ---- serving loop -------------------------------------------------------------
 * Continuous batching of commit and Gram requests on one device. No reference
 * counterpart: the reference computes one commitment per call on the calling
 * thread (commit/src/ajtai/mod.rs:443-452, labrador/src/crs.rs:211-216,
 * labrador/src/main_protocol/garbage.rs:76-93).
 *
 * Requests of any size, ring and kind are submitted from any thread as they
 * arrive. A worker thread admits them in arrival order while the server's HBM
 * arena (mem_bytes) has room for their footprint (NTT-domain witness, partial
 * sums, device copies of host inputs and outputs) and fewer than max_inflight
 * are admitted. The requests admitted together form a wave: one witness
 * transform per ring, one grouped multiply-accumulate and finalize per key,
 * one Gram sweep and finalize per ring. Up to `slots` waves are on the device
 * at once, each on its own stream, so a wave starts while the previous one
 * drains. A request's arena block is released when its wave completes.
 *
 * Each result is bit-identical to the plain call on the same inputs, whatever
 * else shares its wave: t_i = A[:, :n] * s_i as gh_commit with a key of the
 * first n columns, g as gh_gram.
 *
 * Keys and rings used by a request must outlive it. Host buffers of a request
 * (mem == GH_MEM_HOST) must stay valid until gh_server_wait / gh_server_poll
 * reports it done; host witnesses are checked canonical at submit, device
 * witnesses are trusted. */
typedef struct gh_server gh_server;

typedef struct gh_server_config {
  uint64_t mem_bytes;    /* HBM arena for admitted requests (0: 4 GiB) */
  uint32_t max_inflight; /* admitted requests not yet completed (0: 4096) */
  uint32_t max_wave;     /* requests launched together (0: 1024) */
  uint32_t slots;        /* waves on the device at once, one stream each (0: 2, at most 4) */
  uint32_t max_pending;  /* submitted and not yet admitted; more gives GH_E_BUSY (0: 65536) */
} gh_server_config;

#define GH_OP_COMMIT 1 /* t_i = A[:, :n] * s_i for every part i (inner_commit_slice) */
#define GH_OP_GRAM 2   /* upper triangle of <s_i, s_j> (compute_garbage_g) */

typedef struct gh_request {
  uint32_t ops;        /* GH_OP_COMMIT, GH_OP_GRAM, or both (one witness transform) */
  uint32_t r;          /* witness parts, >= 1 */
  uint32_t n;          /* polynomials per part, >= 1 */
  int32_t mem;         /* where witness, t_out and g_out live */
  gh_ring *ring;       /* ring of the context the server runs on */
  gh_key *key;         /* GH_OP_COMMIT: key of `ring`, rows % 4 == 0, >= n columns */
  const void *witness; /* r x n polynomials, coefficient domain */
  void *t_out;         /* GH_OP_COMMIT: r x rows polynomials */
  void *g_out;         /* GH_OP_GRAM: r(r+1)/2 polynomials */
  uint32_t wire;       /* 1: witness, t_out and g_out hold canonical encodings */
  uint32_t reserved;
  uint64_t witness_len; /* wire: bytes at witness */
  uint64_t t_cap;       /* wire: bytes available at t_out */
  uint64_t g_cap;       /* wire: bytes available at g_out */
} gh_request;

/* Wire requests (wire = 1, mem = GH_MEM_HOST) carry the canonical encodings
 * of the LaBRADOR step's objects, so a server front end can pass network bytes
 * through unchanged:
 *   witness  LabradorWitness::parts, a Vec<Vec<poly>> (labrador/src/relation.rs:113-115):
 *            u32 r, then per part u32 n and n coefficient-domain polynomials;
 *            every part has the same n; r and n of the request are 0 (read
 *            from the witness) or must match it;
 *   t_out    the inner commitments t_vecs, a Vec<RingVec<poly>>
 *            (labrador/src/main_protocol/mod.rs:702-726): u32 r, then per part
 *            u64 rows and rows polynomials: 4 + r * (8 + rows * d * wire_bytes) bytes;
 *   g_out    the Gram polynomials, a Vec<poly> (garbage.rs:76-93): u32 r(r+1)/2,
 *            then the polynomials: 4 + r(r+1)/2 * d * wire_bytes bytes.
 * The ring's element width must equal its word width (the 32-bit prime
 * 4294966657 and Goldilocks do), else GH_E_INVALID_PARAMETERS. Submit checks
 * the structure: UNEXPECTED_END, INVALID_DATA for trailing bytes,
 * DIMENSION_MISMATCH for parts of different lengths, INVALID_ARGUMENT for
 * outputs too small. Elements are unpacked and checked on the device: one >= q
 * completes the request with GH_E_INVALID_DATA and leaves its outputs
 * unwritten. */

typedef struct gh_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t waves;           /* waves launched */
  uint64_t commits;         /* commitments completed (r per commit request) */
  uint64_t grams;           /* Gram matrices completed */
  uint64_t mem_in_use;      /* arena bytes held by admitted requests */
  uint64_t mem_peak;
  uint32_t pending;         /* submitted, not admitted */
  uint32_t admitted;        /* admitted, not completed */
  uint32_t max_wave_seen;   /* most requests in one wave */
  uint32_t reserved;
  double device_ms;         /* summed device time of completed waves */
} gh_server_stats;

GH_API int gh_server_create(gh_ctx *ctx, const gh_server_config *cfg, gh_server **out);
/* completes every submitted request, then frees the server */
GH_API int gh_server_destroy(gh_server *srv);
/* validates and queues a request; GH_E_BUSY when max_pending are queued,
 * GH_E_OOM when its footprint exceeds the whole arena */
GH_API int gh_server_submit(gh_server *srv, const gh_request *req, uint64_t *ticket);
/* blocks until the request is done; status is its result, latency_us the time
 * from submit to completion; the ticket is released */
GH_API int gh_server_wait(gh_server *srv, uint64_t ticket, int *status, double *latency_us);
/* non-blocking; when *done is 1 the outputs are written and the ticket is released */
GH_API int gh_server_poll(gh_server *srv, uint64_t ticket, int *done, int *status,
                          double *latency_us);
GH_API int gh_server_query(gh_server *srv, gh_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif /* GRIDHIP_H */
