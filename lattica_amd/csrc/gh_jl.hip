// JL rows for aggregation: the entries of the JL matrix in the reference's
// flat layout, its conjugated polynomials, the polynomial JL evaluations the
// prover folds into b'' and the JL term of the aggregated phi
// (labrador/src/jl.rs:311-383, labrador/src/main_protocol/mod.rs:835-923,
// step_prover.rs:196-205, main_protocol/aggregation.rs:237-245).
//
// Column col = poly * d + coeff of part p is keystream bytes
// [col * pk, col * pk + pk) of ChaCha20(seed, nonce = p), pk = ceil(rows / 4),
// two bits per row: 00 -> +1, 01 -> -1, else 0. With 256 rows a column is one
// keystream block, block counter = col.
//
// Kernels:
//   k_jl_flat        one thread per column, 256 consecutive polynomials of one
//                    coefficient per block: each row's bytes are written as one
//                    contiguous run of the flat layout [m][part][coeff][poly];
//   k_jl_conj_seed   one wave per polynomial (lane = column within it): row m's
//                    conjugated polynomial sigma_{-1}(pi) lands as 64 words, lane
//                    k' writing coefficient (64 - k') mod 64;
//   k_jl_conj_flat   the same from a flat int8 array, a 64 x 64 (coeff x poly)
//                    tile transposed through LDS;
//   k_jl_evals       jl_eval[m][t] = sum over columns c = (part, poly j, k) of
//                    Pi[m][c] * W[c][t], W[(j, k)][t] = s_j[k + t] for k + t < 64
//                    and -s_j[k + t - 64] otherwise (the negacyclic product
//                    sigma_{-1}(pi) * s written out), an exact integer GEMM on
//                    the int8 matrix cores: Pi is ternary, and each W entry is
//                    split into signed base-256 digits (4 for the 32-bit prime,
//                    8 for Goldilocks) whose products sum exactly in int32.
//                    The accumulator tile of a block (256 rows x 256 digit
//                    columns of int32) fills half a CU's register file, so
//                    Goldilocks runs two passes of four digits each; a block
//                    runs only the digits its witness words need (one for
//                    words in [-127, 127]);
//   k_jl_evals_fin   sums each digit plane over the blocks that computed it and
//                    recombines the digits mod q;
//   k_jl_agg         the aggregation's JL term phi_ij += sum_m c_m sigma_{-1}(pi_ij^(m))
//                    (main_protocol/aggregation.rs:237-245) as the same kind of
//                    GEMM with the polynomials as rows: the entries go from the
//                    keystream into the matrix-core A operand in registers, the
//                    digits of the rotated c_m sit in LDS, and the 2 GiB of
//                    conjugated polynomials of r 32 x 1024 are never written.
#include <algorithm>

#include "gh_internal.h"
#include "gh_chacha.h"

namespace gh {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// entry of a 2-bit code (labrador/src/jl.rs:356-362)
__device__ __forceinline__ int jl_entry(uint32_t code) { return code == 0 ? 1 : (code == 1 ? -1 : 0); }

// The keystream bytes of column `col` (pk of them) into 16 little-endian words
// of `ks` (LDS, stride 17 words per thread): two blocks when the column
// straddles a block boundary. Only used when pk < 64.
__device__ void jl_column_lds(const ChaKey& key, uint64_t col, uint32_t pk, uint32_t* ks) {
  const uint64_t off = col * pk;
  const uint32_t b0 = (uint32_t)(off >> 6), o0 = (uint32_t)(off & 63);
  uint32_t blk[16];
  chacha_block(key, b0, blk);
  uint8_t* out = reinterpret_cast<uint8_t*>(ks);
  const uint8_t* bytes = reinterpret_cast<const uint8_t*>(blk);
  const uint32_t first = 64 - o0 < pk ? 64 - o0 : pk;
  for (uint32_t i = 0; i < first; i++) out[i] = bytes[o0 + i];
  if (first < pk) {
    chacha_block(key, b0 + 1, blk);
    for (uint32_t i = first; i < pk; i++) out[i] = bytes[i - first];
  }
}

constexpr int kEvPolys = 8;       // witness polynomials per LDS round

// Exactness of the int32 matrix-core sums. Every A operand entry is a JL
// entry in {-1, 0, 1} and every B operand entry a signed base-256 digit in
// [-128, 127], so an int32 sum of K such products, and every partial sum on
// the way, lies within K * 128 in magnitude. The launch shapes keep
// K * 128 <= 2^31 - 1:
//   k_jl_proj   a block's sum for one (row, digit) adds 64 products per
//               polynomial over the polynomials of its 4 waves (each wave's
//               MFMA accumulator, then the int32 sum of the four):
//               K = 4 * 64 * ppw; ppw <= kProjMaxPpw = 65535 gives
//               K * 128 = 2147450880;
//   k_jl_evals  a block's sum for one (row, digit column) adds 64 products per
//               polynomial of the block, in one wave's accumulator:
//               K = 64 * ppb; ppb <= kEvMaxPpb = 262136 (a multiple of 8)
//               gives K * 128 = 2147418112. Witnesses longer than
//               (blocks) * kEvMaxPpb polynomials run in chunks, each reduced
//               mod q and added to the output;
//   k_jl_agg    each output sums 64 coefficients x 256 rows of one polynomial
//               (K = 2^14, K * 128 = 2^21), independent of the size.
// The reductions sum block partials in int64: at most 2^36 polynomials are
// accepted, so a digit sum stays below 2^49.
// jl_proj_shape and jl_evals_shape compute these shapes (gh_jl_launch_shape).
constexpr uint64_t kProjMaxPpw = 65535;
constexpr uint64_t kEvMaxPpb = 262136;
static_assert(4 * 64 * kProjMaxPpw * 128 <= 0x7fffffffull, "projection int32 bound");
static_assert(64 * kEvMaxPpb * 128 <= 0x7fffffffull, "evaluation int32 bound");
static_assert(kEvMaxPpb % kEvPolys == 0, "evaluation blocks take whole rounds");
constexpr int kEvRowBytes = 64;   // A tile bytes per row of a polynomial (256 rows x 64 coefficients)

// Decodes this lane's keystream column (16 words, 2 bits per row, 256 rows)
// into int8 entries stored as rows of columns: the dword holding row m of the
// four columns of this lane's quad goes to base + m * row_stride + 4 * (k / 4),
// k = the lane's column. Four rows at a time: their four codes are spread one
// per byte and mapped through a byte table (00 -> +1, 01 -> -1, else 0) by
// v_perm; then a 4 x 4 byte transpose across the lane quad (two DPP
// exchanges) leaves each lane one dword holding four columns of one row,
// stored whole: lane r of the quad stores rows 4 g + r. Rows >= `rows` are not
// stored. All 64 lanes of the wave must be active (the exchanges read
// neighbours); `base` + offsets must be 4-byte aligned.
__device__ __forceinline__ void jl_quad_store(const uint32_t (&blk)[16], int8_t* base,
                                              uint64_t row_stride, uint32_t k, uint32_t rows) {
  const uint32_t r = k & 3, a4 = k & ~3u;
  const uint32_t sel1 = (r & 1) ? 0x03070105u : 0x06020400u;  // pair columns (r, r ^ 1)
  const uint32_t sel2 = (r & 2) ? 0x03020706u : 0x05040100u;  // then pairs (r, r ^ 2)
#pragma unroll
  for (int g = 0; g < 64; g++) {
    const uint32_t x = (blk[g >> 2] >> (8 * (g & 3))) & 0xffu;
    const uint32_t y = (x | (x << 6) | (x << 12) | (x << 18)) & 0x03030303u;
    const uint32_t e = __builtin_amdgcn_perm(0u, 0x0000FF01u, y);
    const uint32_t pn = (uint32_t)__builtin_amdgcn_mov_dpp((int)e, 0xB1, 0xF, 0xF, false);
    const uint32_t z = __builtin_amdgcn_perm(pn, e, sel1);
    const uint32_t qn = (uint32_t)__builtin_amdgcn_mov_dpp((int)z, 0x4E, 0xF, 0xF, false);
    const uint32_t w = __builtin_amdgcn_perm(qn, z, sel2);
    const uint32_t m = 4 * g + r;
    if (m < rows) *reinterpret_cast<uint32_t*>(base + (uint64_t)m * row_stride + a4) = w;
  }
}

// ---- matrix-core transposes of keystream columns ------------------------------
//
// A keystream block holds one column for all 256 rows; an MFMA A fragment
// holds one row across 16 columns. An MFMA of the decoded keystream against a
// constant 0/1 matrix turns one into the other in registers: its D layout puts
// register r of lane half h on A row 8 (r / 4) + 4 h + r % 4, so the lanes draw
// their columns in the order that makes those rows consecutive, and the int32
// results are packed back to bytes.

// 16 ternary codes of a keystream word -> 16 int8 entries, byte 4 i + o
// holding code 4 o + i (00 -> +1, 01 -> -1, else 0)
__device__ __forceinline__ v4i jl_decode16(uint32_t x) {
  v4i r;
  r[0] = (int)__builtin_amdgcn_perm(0u, 0x0000FF01u, x & 0x03030303u);
  r[1] = (int)__builtin_amdgcn_perm(0u, 0x0000FF01u, (x >> 2) & 0x03030303u);
  r[2] = (int)__builtin_amdgcn_perm(0u, 0x0000FF01u, (x >> 4) & 0x03030303u);
  r[3] = (int)__builtin_amdgcn_perm(0u, 0x0000FF01u, (x >> 6) & 0x03030303u);
  return r;
}

// Column (of 64) drawn by `lane`: lanes 0-31 the first 32, 32-63 the second,
// lane i of a half on 16 ((i >> 2) & 1) + 4 (i >> 3) + (i & 3): a transpose then
// leaves register r of lane half h on column 16 h + r.
__device__ __forceinline__ uint32_t jl_tcoeff(uint32_t lane) {
  const uint32_t i = lane & 31;
  return 32 * (lane >> 5) + 16 * ((i >> 2) & 1) + 4 * (i >> 3) + (i & 3);
}

// B operand of the keystream transpose: a 1 where K position (jl_decode16's
// order, m = 4 o + i at byte 4 i + o of each 16-group) holds row j = lane & 31
__device__ __forceinline__ v4i jl_tperm(uint32_t lane) {
  const uint32_t j = lane & 31, h = lane >> 5;
  v4i r = {0, 0, 0, 0};
  if ((j >> 4) == h) {
    const uint32_t jj = j & 15, b = 4 * (jj & 3) + (jj >> 2);
    r[b >> 2] = (int)(1u << (8 * (b & 3)));
  }
  return r;
}

// B operand of the digit transpose: the 32 x 32 identity
__device__ __forceinline__ v4i jl_tident(uint32_t lane) {
  const uint32_t j = lane & 31, h = lane >> 5;
  v4i r = {0, 0, 0, 0};
  if ((j >> 4) == h) r[(j & 15) >> 2] = (int)(1u << (8 * (j & 3)));
  return r;
}

// the low bytes of 16 small int32 sums as an int8 operand, register r -> byte r
__device__ __forceinline__ v4i jl_pack(const v16i& d) {
  v4i r;
#pragma unroll
  for (int k = 0; k < 4; k++)
    r[k] = (int)(__builtin_amdgcn_perm((uint32_t)d[4 * k + 1], (uint32_t)d[4 * k], 0x0C0C0400u) |
                 __builtin_amdgcn_perm((uint32_t)d[4 * k + 3], (uint32_t)d[4 * k + 2], 0x04000C0Cu));
  return r;
}

// grid (ceil(n / 256), d, parts), 256 threads; rows <= 256
__global__ __launch_bounds__(256) void k_jl_flat(ChaKey key0, uint32_t rows, uint32_t num_parts,
                                                 uint64_t n, uint32_t d, int8_t* __restrict__ flat) {
  __shared__ uint32_t ks[256 * 17];
  const uint32_t t = threadIdx.x;
  const uint64_t poly = (uint64_t)blockIdx.x * 256 + t;
  const uint32_t coeff = blockIdx.y, part = blockIdx.z;
  ChaKey key = key0;
  key.n0 = part;
  const uint64_t col = poly * d + coeff;
  const uint32_t pk = (rows + 3) / 4;
  const uint64_t row_stride = (uint64_t)num_parts * d * n;
  int8_t* dst = flat + ((uint64_t)part * d + coeff) * n + poly;
  if (pk == 64 && n % 4 == 0 && ((uintptr_t)flat & 3) == 0 &&
      (uint64_t)blockIdx.x * 256 + 256 <= n) {
    // every polynomial of the block exists and rows start 4-byte aligned:
    // dword stores of four polynomials per row through the quad transpose
    uint32_t blk[16];
    chacha_block(key, (uint32_t)col, blk);
    jl_quad_store(blk, flat + ((uint64_t)part * d + coeff) * n + (uint64_t)blockIdx.x * 256,
                  row_stride, t, rows);
    return;
  }
  if (pk == 64) {
    if (poly >= n) return;
    uint32_t blk[16];
    chacha_block(key, (uint32_t)col, blk);
#pragma unroll
    for (int m = 0; m < 256; m++) {
      if ((uint32_t)m < rows)
        dst[(uint64_t)m * row_stride] = (int8_t)jl_entry((blk[m >> 4] >> (2 * (m & 15))) & 3);
    }
    return;
  }
  uint32_t* my = ks + t * 17;
  if (poly < n) jl_column_lds(key, col, pk, my);
  if (poly >= n) return;
  const uint8_t* b = reinterpret_cast<const uint8_t*>(my);
  for (uint32_t m = 0; m < rows; m++)
    dst[(uint64_t)m * row_stride] = (int8_t)jl_entry((b[m >> 2] >> (2 * (m & 3))) & 3);
}

template <class W>
__device__ __forceinline__ W conj_word(int e, bool k0, W q) {
  // coefficient k of sigma_{-1}(pi): e at k = 0, -e elsewhere, as a residue
  const int v = k0 ? e : -e;
  return v == 0 ? W(0) : (v > 0 ? W(1) : q - W(1));
}

// grid (ceil(n / 4), parts), 256 threads = 4 polynomials x 64 columns; d = 64
template <class W>
__global__ __launch_bounds__(256) void k_jl_conj_seed(ChaKey key0, uint32_t rows,
                                                      uint32_t num_parts, uint64_t n, W q,
                                                      W* __restrict__ out) {
  __shared__ uint32_t ks[256 * 17];
  const uint32_t t = threadIdx.x;
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (t >> 6);
  const uint32_t kc = t & 63, part = blockIdx.y;
  ChaKey key = key0;
  key.n0 = part;
  const uint64_t col = j * 64 + kc;
  const uint32_t pk = (rows + 3) / 4;
  const uint32_t kout = (64 - kc) & 63;
  // output polynomial (m, part, j) at ((m * parts + part) * n + j) * 64
  W* dst = out + ((uint64_t)part * n + j) * 64 + kout;
  const uint64_t row_stride = (uint64_t)num_parts * n * 64;
  if (pk == 64) {
    if (j >= n) return;
    uint32_t blk[16];
    chacha_block(key, (uint32_t)col, blk);
#pragma unroll
    for (int m = 0; m < 256; m++) {
      if ((uint32_t)m < rows)
        dst[(uint64_t)m * row_stride] =
            conj_word<W>(jl_entry((blk[m >> 4] >> (2 * (m & 15))) & 3), kc == 0, q);
    }
    return;
  }
  uint32_t* my = ks + t * 17;
  if (j < n) jl_column_lds(key, col, pk, my);
  if (j >= n) return;
  const uint8_t* b = reinterpret_cast<const uint8_t*>(my);
  for (uint32_t m = 0; m < rows; m++)
    dst[(uint64_t)m * row_stride] = conj_word<W>(jl_entry((b[m >> 2] >> (2 * (m & 3))) & 3), kc == 0, q);
}

// grid (ceil(n / 64), parts, rows), 256 threads; d = 64
template <class W>
__global__ __launch_bounds__(256) void k_jl_conj_flat(const int8_t* __restrict__ flat,
                                                      uint32_t num_parts, uint64_t n, W q,
                                                      W* __restrict__ out) {
  constexpr int kStride = 68;  // bytes per coefficient row of the tile (17 dwords)
  __shared__ int8_t tile[64 * kStride];
  const uint32_t t = threadIdx.x;
  const uint64_t j0 = (uint64_t)blockIdx.x * 64;
  const uint32_t part = blockIdx.y, m = blockIdx.z;
  const int8_t* src = flat + ((uint64_t)m * num_parts + part) * 64 * n;
  for (int e = 0; e < 16; e++) {
    const uint32_t idx = e * 256 + t, c = idx >> 6, jj = idx & 63;
    tile[c * kStride + jj] = j0 + jj < n ? src[(uint64_t)c * n + j0 + jj] : (int8_t)0;
  }
  __syncthreads();
  W* dst = out + (((uint64_t)m * num_parts + part) * n + j0) * 64;
  for (int e = 0; e < 16; e++) {
    const uint32_t idx = e * 256 + t, jj = idx >> 6, k = idx & 63;
    if (j0 + jj >= n) continue;
    const uint32_t srck = (64 - k) & 63;
    dst[(uint64_t)jj * 64 + k] = conj_word<W>(tile[srck * kStride + jj], k == 0, q);
  }
}

// ---- JL evaluations on the int8 matrix cores ---------------------------------


// Centred lift of a residue into the window that D signed base-256 digits
// cover exactly ([-128 (256^D - 1) / 255, 127 (256^D - 1) / 255], 2^(8D)
// integers), then its digits, least significant first.
template <int D>
__device__ __forceinline__ void signed_digits(int64_t v, int8_t (&dig)[D]) {
#pragma unroll
  for (int b = 0; b < D; b++) {
    const int64_t lo = (int64_t)(int8_t)(uint8_t)(v & 0xff);  // in [-128, 127]
    dig[b] = (int8_t)lo;
    v = (v - lo) >> 8;
  }
}

template <int D>
__device__ __forceinline__ void lift_digits(uint64_t x, uint64_t q, int8_t (&dig)[D]) {
  int64_t v;
  if (D == 4) {
    v = x <= 2139062143ull ? (int64_t)x : (int64_t)x - (int64_t)q;
  } else {
    v = x <= 0x7F7F7F7F7F7F7F7Full ? (int64_t)x : (int64_t)(x - q);
  }
  signed_digits<D>(v, dig);
}

// Signed base-256 digits the lifted word v needs: D digits cover
// [-128 (256^D - 1) / 255, 127 (256^D - 1) / 255], nested ranges, so the count
// is one plus the number of ranges below DT digits that v falls outside.
template <int DT>
__device__ __forceinline__ uint32_t digits_needed(int64_t v) {
  uint32_t d = 1;
  int64_t lo = -128, hi = 127;
#pragma unroll
  for (int i = 1; i < DT; i++) {
    d += (v < lo || v > hi) ? 1u : 0u;
    if (i + 1 < DT) {
      lo = lo * 256 - 128;
      hi = hi * 256 + 127;
    }
  }
  return d;
}

// Digits the lifted word x (or its negation: the evaluations multiply by the
// digits of s and of -s) needs.
template <class W, int DT>
__device__ __forceinline__ uint32_t word_span(uint64_t x, uint64_t q) {
  const uint64_t xn = x == 0 ? 0 : q - x;
  int64_t v, vn;
  if (DT == 4) {
    v = x <= 2139062143ull ? (int64_t)x : (int64_t)x - (int64_t)q;
    vn = xn <= 2139062143ull ? (int64_t)xn : (int64_t)xn - (int64_t)q;
  } else {
    v = x <= 0x7F7F7F7F7F7F7F7Full ? (int64_t)x : (int64_t)(x - q);
    vn = xn <= 0x7F7F7F7F7F7F7F7Full ? (int64_t)xn : (int64_t)(xn - q);
  }
  return max(digits_needed<DT>(v), digits_needed<DT>(vn));
}

// One pass over ND digits dbase .. dbase + ND - 1 of DT (ND = 1, 2 or 4):
// block of 8 waves, wave w on digit dbase + w / (8 / ND) (both 32-wide t
// halves) and m-tiles ND (w % (8 / ND)) .. + ND - 1 (32 rows each): 2 ND
// accumulator tiles of 32 x 32, every wave busy whatever ND is.
//
// The polynomials go through in steps of kEvStep = 4, double-buffered in LDS:
// A tiles (the decoded keystream, 256 x 64 int8 per polynomial) and digit
// windows E ([poly][digit][128] bytes: digits of s_u, then of -s_u). Threads
// t / 256 = g draw the keystream columns (one each: polynomial (t / 64) % 4,
// column t % 64) of every other step. During step s, with every wave running
// the MFMAs of step s from buffer s % 2:
//   the half that drew step s + 1 stores its A tile and E into buffer
//   (s + 1) % 2 first, then multiplies;
//   the other half multiplies first, then draws step s + 2 (keystream and
//   witness word held in registers until it stores them in step s + 1).
// Waves w and w + 4 share a SIMD, so on each SIMD one wave's vector work runs
// beside the other's matrix-core work. Partial sums leave per block.
constexpr int kEvDigits = 4;  // digit rows of E per polynomial (one pass)
constexpr int kEvStep = 4;    // polynomials per pipeline step (two LDS buffers)
static_assert(2 * kEvStep == kEvPolys, "the two step buffers fill the LDS tiles");

template <class W, int DT, int ND>
__device__ __forceinline__ void evals_pass(ChaKey key0, const W* __restrict__ parts, uint64_t n,
                                           uint64_t g_base, uint64_t total_polys,
                                           uint64_t polys_per_block,
                                           uint64_t q, uint32_t dbase, int32_t* __restrict__ partial,
                                           int8_t* a_tile, int8_t* e_tile, uint32_t (&blk)[16],
                                           uint64_t& xw) {
  constexpr int kN = 64 * DT;
  constexpr int kWavesPerDigit = 8 / ND;
  // The keystream columns become A fragments on the matrix cores (as in
  // k_jl_proj: the lanes draw their coefficients in jl_tcoeff order, and an
  // MFMA against a 0/1 matrix transposes them), stored fragment-major:
  // [m-tile][k-step][lane] x 16 bytes per polynomial.
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t digit = wave / kWavesPerDigit, mgrp = wave % kWavesPerDigit;
  const uint32_t half = t >> 8, ps_own = (t >> 6) & 3;
  const uint32_t u = jl_tcoeff(lane);  // this lane's coefficient
  const v4i tperm = jl_tperm(lane);
  const uint64_t g_begin = g_base + (uint64_t)blockIdx.x * polys_per_block;
  const uint64_t g_end = std::min<uint64_t>(g_begin + polys_per_block, total_polys);
  const uint32_t nsteps = (uint32_t)((g_end - g_begin + kEvStep - 1) / kEvStep);
  v16i acc[ND][2];
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int h = 0; h < 2; h++) acc[i][h] = v16i{};

  // this thread's keystream column and witness word of step s (code 10, entry
  // 0, and a zero word past the block's end)
  auto draw = [&](uint32_t step) {
    const uint64_t g = g_begin + (uint64_t)step * kEvStep + ps_own;  // wave-uniform
    if (g < g_end) {
      xw = (uint64_t)parts[g * 64 + u];
      ChaKey key = key0;
      key.n0 = (uint32_t)(g / n);
      chacha_block(key, (uint32_t)((g % n) * 64 + u), blk);
    } else {
      xw = 0;
#pragma unroll
      for (int i = 0; i < 16; i++) blk[i] = 0xAAAAAAAAu;
    }
  };
  // store what draw() left into buffer step % 2
  auto store = [&](uint32_t step) {
    const uint32_t b = step & 1;
    int8_t* tp = a_tile + (b * kEvStep + ps_own) * 256 * kEvRowBytes;
#pragma unroll
    for (int mc = 0; mc < 8; mc++) {  // rows 32 mc .. 32 mc + 31, both k-steps
      const auto sw = __builtin_amdgcn_permlane32_swap(blk[2 * mc], blk[2 * mc + 1], false, false);
      const v4i a1 = jl_pack(
          __builtin_amdgcn_mfma_i32_32x32x32_i8(jl_decode16(sw[0]), tperm, v16i{}, 0, 0, 0));
      const v4i a2 = jl_pack(
          __builtin_amdgcn_mfma_i32_32x32x32_i8(jl_decode16(sw[1]), tperm, v16i{}, 0, 0, 0));
      *reinterpret_cast<v4i*>(tp + ((2 * mc) * 64 + lane) * 16) = a1;
      *reinterpret_cast<v4i*>(tp + ((2 * mc + 1) * 64 + lane) * 16) = a2;
    }
    int8_t dp[DT], dn[DT];
    lift_digits<DT>(xw, q, dp);
    lift_digits<DT>(xw == 0 ? 0 : q - xw, q, dn);
    int8_t* e = e_tile + (b * kEvStep + ps_own) * kEvDigits * 128;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      e[d * 128 + u] = dp[dbase + d];
      e[d * 128 + 64 + u] = dn[dbase + d];
    }
  };
  const uint32_t row = lane & 31, kg = lane >> 5;
  auto mfma_step = [&](uint32_t step) {
    const uint32_t b = step & 1;
    const uint64_t g0 = g_begin + (uint64_t)step * kEvStep;
    for (int ps = 0; ps < kEvStep; ps++) {
      if (g0 + ps >= g_end) break;
      const int8_t* at = a_tile + (b * kEvStep + ps) * 256 * kEvRowBytes;
      const uint32_t* ew =
          reinterpret_cast<const uint32_t*>(e_tile + ((b * kEvStep + ps) * kEvDigits + digit) * 128);
#pragma unroll
      for (int kstep = 0; kstep < 2; kstep++) {
        const uint32_t k0 = kstep * 32 + kg * 16;
        v4i bfrag[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          // bytes E[t + k0 .. t + k0 + 15], t = 32 h + row: five aligned words
          const uint32_t o = 32 * h + row + k0, wi = o >> 2, sh = o & 3;
          uint32_t w5[5];
#pragma unroll
          for (int x = 0; x < 5; x++) w5[x] = ew[wi + x];
#pragma unroll
          for (int x = 0; x < 4; x++)
            bfrag[h][x] = (int)__builtin_amdgcn_alignbyte(w5[x + 1], w5[x], sh);
        }
#pragma unroll
        for (int i = 0; i < ND; i++) {
          const v4i afrag = *reinterpret_cast<const v4i*>(
              at + ((2 * (mgrp * ND + i) + kstep) * 64 + lane) * 16);
#pragma unroll
          for (int h = 0; h < 2; h++)
            acc[i][h] = __builtin_amdgcn_mfma_i32_32x32x32_i8(afrag, bfrag[h], acc[i][h], 0, 0, 0);
        }
      }
    }
  };

  // the caller drew steps 0 (half 0) and 1 (half 1)
  if (half == 0) store(0);
  __syncthreads();
  for (uint32_t s = 0; s < nsteps; s++) {  // block-uniform
    if (half != (s & 1)) {
      if (s + 1 < nsteps) store(s + 1);
      mfma_step(s);
    } else {
      mfma_step(s);
      if (s + 2 < nsteps) draw(s + 2);
    }
    __syncthreads();
  }
  // -- partial sums: partial[block][m][n], n = digit * 64 + t
  int32_t* pb = partial + (uint64_t)blockIdx.x * 256 * kN;
  const uint32_t col_in = lane & 31, rgrp = lane >> 5;
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const uint32_t m = (mgrp * ND + i) * 32 + 8 * (r >> 2) + 4 * rgrp + (r & 3);
        const uint32_t nn = (dbase + digit) * 64 + h * 32 + col_in;
        pb[(uint64_t)m * kN + nn] = acc[i][h][r];
      }
}

// Digits dbase .. dbase + 3 of DT, of which only those below the block's
// digit span (the most any of its witness words needs) can be nonzero: the
// pass runs over 1, 2 or 4 of them (3 as 4), or not at all. The dbase = 0 pass
// finds the span (its loads in flight while the first keystream is drawn) and
// records it in span[block] for the later passes and the reduction.
template <class W, int DT>
__global__ __launch_bounds__(512) void k_jl_evals(ChaKey key0, const W* __restrict__ parts,
                                                  uint64_t n, uint64_t g_base,
                                                  uint64_t total_polys,
                                                  uint64_t polys_per_block, uint64_t q,
                                                  uint32_t dbase, uint32_t* __restrict__ span,
                                                  int32_t* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) int8_t a_tile[kEvPolys * 256 * kEvRowBytes];
  __shared__ __attribute__((aligned(16))) int8_t e_tile[kEvPolys * kEvDigits * 128];
  __shared__ uint32_t wmax[8];
  const uint32_t t = threadIdx.x;
  const uint64_t g_begin = g_base + (uint64_t)blockIdx.x * polys_per_block;
  const uint64_t g_end = std::min<uint64_t>(g_begin + polys_per_block, total_polys);
  uint32_t deff;
  uint32_t blk[16];
  uint64_t xw;
  // steps 0 and 1 of the pass (polynomials g_begin + t / 64): keystream
  // column and witness word of coefficient jl_tcoeff(lane)
  auto first_keystream = [&]() {
    const uint64_t g = g_begin + (t >> 6);  // wave-uniform
    const uint32_t u = jl_tcoeff(t & 63);
    if (g < g_end) {
      xw = (uint64_t)parts[g * 64 + u];
      ChaKey key = key0;
      key.n0 = (uint32_t)(g / n);
      chacha_block(key, (uint32_t)((g % n) * 64 + u), blk);
    } else {
      xw = 0;
#pragma unroll
      for (int i = 0; i < 16; i++) blk[i] = 0xAAAAAAAAu;
    }
  };
  if (dbase == 0) {
    const uint64_t w_end = g_end * 64;
    uint32_t d = 1;
    for (uint64_t i0 = g_begin * 64 + t; i0 < w_end; i0 += 16 * 512) {
      uint64_t x[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const uint64_t i = i0 + u * 512;
        x[u] = i < w_end ? (uint64_t)parts[i] : 0;
      }
      if (i0 == g_begin * 64 + t) first_keystream();
#pragma unroll
      for (int u = 0; u < 16; u++) d = max(d, word_span<W, DT>(x[u], q));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d = max(d, (uint32_t)__shfl_xor((int)d, o));
    if ((t & 63) == 0) wmax[t >> 6] = d;
    __syncthreads();
    deff = wmax[0];
#pragma unroll
    for (int w = 1; w < 8; w++) deff = max(deff, wmax[w]);
    if (t == 0) span[blockIdx.x] = deff;
  } else {
    deff = span[blockIdx.x];
    if (dbase >= deff) return;  // block-uniform
    first_keystream();
  }
  const uint32_t nd = deff - dbase;
  if (nd == 1)
    evals_pass<W, DT, 1>(key0, parts, n, g_base, total_polys, polys_per_block, q, dbase, partial, a_tile,
                         e_tile, blk, xw);
  else if (nd == 2)
    evals_pass<W, DT, 2>(key0, parts, n, g_base, total_polys, polys_per_block, q, dbase, partial, a_tile,
                         e_tile, blk, xw);
  else
    evals_pass<W, DT, 4>(key0, parts, n, g_base, total_polys, polys_per_block, q, dbase, partial, a_tile,
                         e_tile, blk, xw);
}

// Goldilocks: 256 r mod p for r < p (2^64 = 2^32 - 1 mod p)
__device__ __forceinline__ uint64_t gold_shift8(uint64_t r) {
  const uint64_t p = 0xFFFFFFFF00000001ull;
  const uint64_t lo = r << 8, t = (r >> 56) * 0xFFFFFFFFull;
  uint64_t s = lo + t;
  if (s < lo) s += 0xFFFFFFFFull;
  return s >= p ? s - p : s;
}

// evals[m][t] = sum_b 256^b sum_blocks partial[block][m][b * 64 + t] mod q.
// Block m, one thread per digit column n = b * 64 + t summing its column over
// the blocks (eight loads in flight), then 64 threads recombine the digits
// from the top one (Horner). Digit sums are below 2^41 in magnitude (below
// 2^31 per block partial, at most kEvMaxBlocks blocks). With `accumulate` the
// result is added mod q to what `out` holds (the chunks of a long witness).
constexpr uint32_t kEvMaxBlocks = 1024;  // evaluation blocks (the reduction lists them)

template <class W, int D>
__global__ __launch_bounds__(512) void k_jl_evals_fin(const int32_t* __restrict__ partial,
                                                      uint32_t nblk, uint64_t q,
                                                      const uint32_t* __restrict__ span,
                                                      int accumulate, W* __restrict__ out) {
  constexpr int kN = 64 * D, kWaves = 8, kU = 16;
  __shared__ uint16_t blocks[D][kEvMaxBlocks];
  __shared__ uint32_t counts[D];
  __shared__ int64_t sums[kWaves][kN];
  const uint32_t m = blockIdx.x, t = threadIdx.x, lane = t & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  // plane pl lists the blocks whose span covers it (the others did not
  // compute it); wave pl builds the list
  if (wave < (uint32_t)D) {
    const uint32_t pl = wave;
    uint32_t cnt = 0;
    for (uint32_t b0 = 0; b0 < nblk; b0 += 64) {
      const uint32_t b = b0 + lane;
      const bool has = b < nblk && pl < span[b];
      const uint64_t mask = __ballot(has);
      if (has) blocks[pl][cnt + __popcll(mask & ((1ull << lane) - 1))] = (uint16_t)b;
      cnt += (uint32_t)__popcll(mask);
    }
    if (lane == 0) counts[pl] = cnt;
  }
  __syncthreads();
  // every wave takes every 8th listed block of each plane, kU loads in flight
  const uint64_t stride = (uint64_t)256 * kN;
  for (int pl = 0; pl < D; pl++) {
    const int32_t* p = partial + (uint64_t)m * kN + pl * 64 + lane;
    const uint16_t* list = blocks[pl];
    const uint32_t cnt = counts[pl];
    int64_t acc[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) acc[u] = 0;
    uint32_t i = wave;
    for (; i + (kU - 1) * kWaves < cnt; i += kU * kWaves) {
#pragma unroll
      for (int u = 0; u < kU; u++) acc[u] += p[(uint64_t)list[i + u * kWaves] * stride];
    }
    for (; i < cnt; i += kWaves) acc[0] += p[(uint64_t)list[i] * stride];
    int64_t s = 0;
#pragma unroll
    for (int u = 0; u < kU; u++) s += acc[u];
    sums[wave][pl * 64 + lane] = s;
  }
  __syncthreads();
  if (t >= 64) return;
  const uint32_t nn = t;
  uint64_t r = 0;
#pragma unroll
  for (int b = D - 1; b >= 0; b--) {
    int64_t sb = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) sb += sums[w][b * 64 + nn];
    if (D == 4) {
      const int64_t v = ((int64_t)r * 256 + sb) % (int64_t)q;
      r = (uint64_t)(v < 0 ? v + (int64_t)q : v);
    } else {
      r = gold_shift8(r);
      const uint64_t a = sb < 0 ? (uint64_t)(-sb) : (uint64_t)sb;  // < 2^40 < q
      if (sb >= 0) {
        const uint64_t x = r + a;
        r = (x < r || x >= q) ? x - q : x;
      } else {
        r = r >= a ? r - a : r + (q - a);
      }
    }
  }
  if (accumulate) {  // a later chunk of the witness: add to the earlier chunks' sum
    const uint64_t o = (uint64_t)out[(uint64_t)m * 64 + nn];
    const uint64_t x = r + o;  // both below q; wraps 2^64 only for Goldilocks
    r = (x < r || x >= q) ? x - q : x;
  }
  out[(uint64_t)m * 64 + nn] = (W)r;
}

// ---- JL aggregation on the int8 matrix cores ----------------------------------
//
// phi_ij += sum_m c_m * sigma_{-1}(pi_ij^(m)). With sigma(pi) = sum_k e(k) X^-k
// and (c X^-k)[t] = c[t + k] for t + k < 64, -c[t + k - 64] otherwise, this is
// phi[g][t] += sum_(k, m) Pi[m][g, k] * R[w = t + k][m], R[w][m] the digits of
// c_m[w] (w < 64) or of -c_m[w - 64]: a GEMM with the polynomials g as M,
// (k, m) as K (32 m per step) and t as N. The A operand comes straight from
// the keystream in registers: lane l draws column (g, k) of its polynomial,
// l < 32 for even k and l >= 32 for odd k; one v_permlane32_swap per step
// hands each half the word it needs, so every block is drawn once. Within a
// 16-row group the K order is m = 4 o + i at byte 4 i + o, which is what four
// masked shifts of a keystream word give, and R is stored in the same order.
// R sits in LDS.

constexpr int kAggPolys = 128;   // polynomials per block: 4 tiles of 32
constexpr int kAggRow = 272;     // bytes per R row: 256 m + 16 pad (bank spread)

// sum_b 256^(dbase + b) acc[b] mod q for four digit sums of magnitude < 2^22
template <bool GOLD>
__device__ __forceinline__ uint64_t agg_combine(const int32_t (&a)[4], uint32_t dbase, uint64_t q) {
  const int64_t v = (int64_t)a[0] + ((int64_t)a[1] << 8) + ((int64_t)a[2] << 16) +
                    ((int64_t)a[3] << 24);  // below 2^47 in magnitude
  if (!GOLD) {
    const int64_t r = v % (int64_t)q;
    return (uint64_t)(r < 0 ? r + (int64_t)q : r);
  }
  uint64_t r = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
  for (uint32_t b = 0; b < dbase; b++) r = gold_shift8(r);
  return r;
}

// Grid: one block per 128 polynomials, 8 waves (two per SIMD, so one wave's
// keystream can run beside the other's matrix-core ops); wave w
// multiplies tile w & 3 over half w >> 2 of the coefficients (k-pairs
// 16 h .. 16 h + 15) for digits dbase .. dbase + 3; the halves meet in LDS and
// the first adds its sums to phi. Lanes past the last polynomial draw
// keystream like the others: an A row only reaches the D row of the same
// polynomial, which is not stored.
template <class W, int DT>
__global__ __launch_bounds__(512) void k_jl_agg(ChaKey key0, const W* __restrict__ scale,
                                                uint64_t n, uint64_t total_polys, uint64_t q,
                                                uint32_t dbase, W* __restrict__ phi) {
  constexpr int D = 4;  // digits per pass
  constexpr bool GOLD = DT == 8;
  __shared__ __attribute__((aligned(16))) int8_t tab[D * 128 * kAggRow];  // 139,264 B
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;

  // R: row (b, w) = digit dbase + b of c_m[w] or -c_m[w - 64]; within each
  // 16-m group byte 4 i + o holds m = 4 o + i. A task writes one word: i of
  // group c, m = 16 c + i + 4 o for o = 0..3.
  for (uint32_t task = t; task < 64 * 64; task += 512) {
    const uint32_t u = task & 63, c = task >> 8, i = (task >> 6) & 3;
    uint32_t pw[D] = {0, 0, 0, 0}, nw[D] = {0, 0, 0, 0};
#pragma unroll
    for (int o = 0; o < 4; o++) {
      const uint64_t x = (uint64_t)scale[(uint64_t)(16 * c + i + 4 * o) * 64 + u];
      int8_t dp[DT], dn[DT];
      lift_digits<DT>(x, q, dp);
      lift_digits<DT>(x == 0 ? 0 : q - x, q, dn);
#pragma unroll
      for (int b = 0; b < D; b++) {
        pw[b] |= (uint32_t)(uint8_t)dp[dbase + b] << (8 * o);
        nw[b] |= (uint32_t)(uint8_t)dn[dbase + b] << (8 * o);
      }
    }
#pragma unroll
    for (int b = 0; b < D; b++) {
      *reinterpret_cast<uint32_t*>(tab + (b * 128 + u) * kAggRow + 16 * c + 4 * i) = pw[b];
      *reinterpret_cast<uint32_t*>(tab + (b * 128 + 64 + u) * kAggRow + 16 * c + 4 * i) = nw[b];
    }
  }
  __syncthreads();

  const uint32_t tile = wave & 3, half = wave >> 2, hi = lane >> 5, col = lane & 31;
  const uint64_t g = (uint64_t)blockIdx.x * kAggPolys + tile * 32 + col;  // this lane's A row
  const bool live = g < total_polys;
  ChaKey key = key0;
  key.n0 = live ? (uint32_t)(g / n) : 0u;
  const uint32_t jcol = live ? (uint32_t)((g % n) * 64) : 0u;
  v16i acc[2][D];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int b = 0; b < D; b++) acc[h][b] = v16i{};

  // one K step: coefficient pair (k0, k0 + 1), rows m = 32 s .. 32 s + 31, the
  // half hi of the wave on m = 32 s + 16 hi .. + 15 (word 2 s + hi of a block)
  auto step = [&](const uint32_t (&blk)[16], uint32_t k0, int s) {
    const auto sw = __builtin_amdgcn_permlane32_swap(blk[2 * s], blk[2 * s + 1], false, false);
    const v4i a0 = jl_decode16(sw[0]);  // coefficient k0
    const v4i a1 = jl_decode16(sw[1]);  // coefficient k0 + 1
    const int8_t* rb = tab + 32 * s + 16 * hi;
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const v4i& a = kk ? a1 : a0;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t w = 32 * h + col + k0 + kk;  // t + k, below 128
#pragma unroll
        for (int b = 0; b < D; b++) {
          const v4i bf = *reinterpret_cast<const v4i*>(rb + (b * 128 + w) * kAggRow);
          acc[h][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bf, acc[h][b], 0, 0, 0);
        }
      }
    }
  };

  const uint32_t kp_begin = 16 * half;
  for (uint32_t kp = kp_begin; kp < kp_begin + 16; kp++) {
    uint32_t blk[16];
    chacha_block(key, jcol + 2 * kp + hi, blk);
#pragma unroll
    for (int s = 0; s < 8; s++) step(blk, 2 * kp, s);
    // matrix-core order for the 128 MFMAs of the pair: each followed by the R
    // read four MFMAs ahead, so four reads are in flight instead of one
#pragma unroll
    for (int i = 0; i < 4; i++) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
#pragma unroll
    for (int i = 0; i < 124; i++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
    }
#pragma unroll
    for (int i = 0; i < 4; i++) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
  }

  // the second half's sums through LDS (R is no longer read), lane-major
  __syncthreads();
  int32_t* xs = reinterpret_cast<int32_t*>(tab) + (uint64_t)tile * (2 * D * 16 * 64);
  if (half == 1) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int b = 0; b < D; b++)
#pragma unroll
        for (int r = 0; r < 16; r++) xs[((h * D + b) * 16 + r) * 64 + lane] = acc[h][b][r];
  }
  __syncthreads();
  if (half == 1) return;
  const uint64_t g0 = (uint64_t)blockIdx.x * kAggPolys + tile * 32;
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const uint64_t gg = g0 + 8 * (r >> 2) + 4 * hi + (r & 3);  // D-tile row
      int32_t sums[4];
#pragma unroll
      for (int b = 0; b < D; b++) sums[b] = acc[h][b][r] + xs[((h * D + b) * 16 + r) * 64 + lane];
      if (gg >= total_polys) continue;
      W* p = phi + gg * 64 + 32 * h + col;
      const uint64_t add = agg_combine<GOLD>(sums, dbase, q);
      const uint64_t cur_w = (uint64_t)*p;
      const uint64_t s2 = cur_w + add;  // both below q; wraps 2^64 only for Goldilocks
      *p = (W)((s2 < cur_w || s2 >= q) ? s2 - q : s2);
    }
}

// ---- JL projection on the int8 matrix cores -----------------------------------
//
// p[m] = sum_c Pi[m][c] * s_c as an int8 GEMM: Pi (256 rows x the columns) as
// A, one B column per digit of s_c: the centred lift in 5 signed digits for
// the 32-bit prime, so the int64 sum is the reference's exact p
// (jl.rs:388-508), or the Goldilocks window lift in 8 digits, reduced mod p.
// The output is small (256 rows x 8 digit columns), so every wave holds all of
// it and works through whole polynomials of its own, with no LDS tile: lane l
// draws the keystream of one coefficient, and the keystream words, which hold
// one column and many rows, are turned into A fragments, which hold one row
// and many columns, by a matrix-core transpose (an MFMA against a constant 0/1
// matrix; the lanes draw their coefficients in the order that makes the
// transposed columns come out natural). The witness digits are transposed
// the same way into the B fragments.
constexpr int kProjDigits = 8;  // digit slots per witness word (B columns used)
constexpr int kProjWaves = 4;   // waves per block

// Grid: blocks of 4 waves, wave w of block b on polynomials
// [(4 b + w) ppw, + ppw); per polynomial: 16 keystream transposes, 2 digit
// transposes and 16 accumulating MFMAs (8 row tiles x 2 coefficient halves).
// The waves' sums meet in LDS; one partial [256][8] per block.
template <class W, bool GOLD>
__global__ __launch_bounds__(256, 2) void k_jl_proj(ChaKey key0, const W* __restrict__ parts,
                                                 uint64_t n, uint64_t total_polys,
                                                 uint64_t polys_per_wave, uint64_t q,
                                                 int32_t* __restrict__ partial) {
  __shared__ int32_t red[kProjWaves - 1][256 * kProjDigits];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t g_begin = ((uint64_t)blockIdx.x * kProjWaves + wave) * polys_per_wave;
  const uint64_t g_end = std::min<uint64_t>(g_begin + polys_per_wave, total_polys);
  const uint32_t coeff = jl_tcoeff(lane);
  const v4i tperm = jl_tperm(lane), tident = jl_tident(lane);
  v16i acc[8];
#pragma unroll
  for (int mc = 0; mc < 8; mc++) acc[mc] = v16i{};

  for (uint64_t g = g_begin; g < g_end; g++) {  // wave-uniform
    ChaKey key = key0;
    key.n0 = (uint32_t)(g / n);
    uint32_t blk[16];
    chacha_block(key, (uint32_t)((g % n) * 64 + coeff), blk);
    const uint64_t x = (uint64_t)parts[g * 64 + coeff];
    int64_t v;
    if (GOLD)
      v = x <= 0x7F7F7F7F7F7F7F7Full ? (int64_t)x : (int64_t)(x - q);
    else
      v = x > q / 2 ? (int64_t)x - (int64_t)q : (int64_t)x;  // centred, as p_exact
    int8_t dig[kProjDigits];
    signed_digits<kProjDigits>(v, dig);
    uint32_t dw[2] = {0, 0};
#pragma unroll
    for (int d = 0; d < kProjDigits; d++) dw[d >> 2] |= (uint32_t)(uint8_t)dig[d] << (8 * (d & 3));
    // digit rows: the low half of the wave for the first 32 coefficients, the
    // high half's digits brought down for the second
    const auto s0 = __builtin_amdgcn_permlane32_swap(dw[0], 0u, false, false);
    const auto s1 = __builtin_amdgcn_permlane32_swap(dw[1], 0u, false, false);
    const v4i bd1 = jl_pack(__builtin_amdgcn_mfma_i32_32x32x32_i8(
        v4i{(int)s0[0], (int)s1[0], 0, 0}, tident, v16i{}, 0, 0, 0));
    const v4i bd2 = jl_pack(__builtin_amdgcn_mfma_i32_32x32x32_i8(
        v4i{(int)s0[1], (int)s1[1], 0, 0}, tident, v16i{}, 0, 0, 0));
#pragma unroll
    for (int mc = 0; mc < 8; mc++) {
      // rows 32 mc .. 32 mc + 31: word 2 mc + hi of each block, one swap per pair
      const auto sw = __builtin_amdgcn_permlane32_swap(blk[2 * mc], blk[2 * mc + 1], false, false);
      const v4i a1 = jl_pack(__builtin_amdgcn_mfma_i32_32x32x32_i8(jl_decode16(sw[0]), tperm,
                                                                   v16i{}, 0, 0, 0));
      const v4i a2 = jl_pack(__builtin_amdgcn_mfma_i32_32x32x32_i8(jl_decode16(sw[1]), tperm,
                                                                   v16i{}, 0, 0, 0));
      acc[mc] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, bd1, acc[mc], 0, 0, 0);
      acc[mc] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a2, bd2, acc[mc], 0, 0, 0);
    }
  }

  // digit column d = lane & 31 < 8; row m = 32 mc + 8 (r / 4) + 4 (lane / 32) + r % 4
  const uint32_t d = lane & 31, hi = lane >> 5;
  if (wave > 0 && d < kProjDigits) {
#pragma unroll
    for (int mc = 0; mc < 8; mc++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const uint32_t m = 32 * mc + 8 * (r >> 2) + 4 * hi + (r & 3);
        red[wave - 1][m * kProjDigits + d] = acc[mc][r];
      }
  }
  __syncthreads();
  if (wave == 0 && d < kProjDigits) {
#pragma unroll
    for (int mc = 0; mc < 8; mc++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const uint32_t m = 32 * mc + 8 * (r >> 2) + 4 * hi + (r & 3);
        int32_t sum = acc[mc][r];
#pragma unroll
        for (int w = 0; w < kProjWaves - 1; w++) sum += red[w][m * kProjDigits + d];
        partial[((uint64_t)blockIdx.x * 256 + m) * kProjDigits + d] = sum;
      }
  }
}

// p[m] from the block partials: 32 lanes per digit sum the blocks, a tree
// in LDS, then the digits recombine (exactly for the 32-bit prime, mod p for
// Goldilocks). Block m, 256 threads.
template <bool GOLD>
__global__ __launch_bounds__(256) void k_jl_proj_fin(const int32_t* __restrict__ partial,
                                                     uint32_t nblk, uint64_t q, void* p_modq,
                                                     int64_t* __restrict__ p_exact) {
  __shared__ int64_t red[kProjDigits][32];
  const uint32_t m = blockIdx.x, b = threadIdx.x >> 5, l = threadIdx.x & 31;
  int64_t s = 0;
  for (uint32_t blk = l; blk < nblk; blk += 32)
    s += partial[((uint64_t)blk * 256 + m) * kProjDigits + b];
  red[b][l] = s;
  __syncthreads();
  for (uint32_t h = 16; h > 0; h >>= 1) {
    if (l < h) red[b][l] += red[b][l + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (!GOLD) {
    // p mod q by Horner over the digit sums (each below 2^49 in magnitude for
    // any accepted size); the int64 p itself only when the caller asks for it,
    // which the entry point allows for fewer than 2^26 polynomials
    // (|p| <= total * 64 * (q - 1) / 2 < 2^63)
    int64_t r = 0;
#pragma unroll
    for (int d = kProjDigits - 1; d >= 0; d--) {
      r = (r * 256 + red[d][0] % (int64_t)q) % (int64_t)q;
      if (r < 0) r += (int64_t)q;
    }
    ((uint32_t*)p_modq)[m] = (uint32_t)r;
    if (p_exact) {
      int64_t e = 0;
#pragma unroll
      for (int d = kProjDigits - 1; d >= 0; d--) e = e * 256 + red[d][0];
      p_exact[m] = e;
    }
  } else {
    uint64_t r = 0;
#pragma unroll
    for (int d = kProjDigits - 1; d >= 0; d--) {
      const int64_t sb = red[d][0];
      r = gold_shift8(r);
      const uint64_t a = sb < 0 ? (uint64_t)(-sb) : (uint64_t)sb;
      if (sb >= 0) {
        const uint64_t x = r + a;
        r = (x < r || x >= q) ? x - q : x;
      } else {
        r = r >= a ? r - a : r + (q - a);
      }
    }
    ((uint64_t*)p_modq)[m] = r;
  }
}

}  // namespace

void jl_proj_shape(uint64_t total, int num_cus, JlShape* s) {
  // two waves per SIMD, whole polynomials per wave, at most kProjMaxPpw each
  const uint64_t waves = (uint64_t)std::max(num_cus, 1) * 2 * kProjWaves;
  uint64_t ppw = std::max<uint64_t>(1, (total + waves - 1) / waves);
  ppw = std::min<uint64_t>(ppw, kProjMaxPpw);
  uint64_t nblk = (total + ppw * kProjWaves - 1) / (ppw * kProjWaves);
  if (nblk == 0) nblk = 1;
  s->chunk_polys = total;
  s->per_unit = ppw;
  s->blocks = nblk;
  s->chunks = 1;
  s->terms = (uint64_t)kProjWaves * 64 * ppw;
  s->bound = s->terms * 128;
}

// one chunk of the evaluations: blocks (at most one per CU) and polynomials
// per block (a multiple of kEvPolys); polys <= kEvMaxPpb * the block cap
static void jl_evals_chunk(uint64_t polys, int num_cus, uint64_t* nblk, uint64_t* ppb) {
  const uint64_t nmax = std::min<uint64_t>((uint64_t)std::max(num_cus, 1), kEvMaxBlocks);
  uint64_t nb = std::min<uint64_t>(nmax, (polys + kEvPolys - 1) / kEvPolys);
  if (nb == 0) nb = 1;
  uint64_t p = (polys + nb - 1) / nb;
  p = (p + kEvPolys - 1) / kEvPolys * kEvPolys;
  if (p == 0) p = kEvPolys;
  nb = (polys + p - 1) / p;
  if (nb == 0) nb = 1;
  *nblk = nb;
  *ppb = p;
}

void jl_evals_shape(uint64_t total, int num_cus, JlShape* s) {
  const uint64_t nmax = std::min<uint64_t>((uint64_t)std::max(num_cus, 1), kEvMaxBlocks);
  const uint64_t chunk = nmax * kEvMaxPpb;
  const uint64_t first = std::min<uint64_t>(total, chunk);
  uint64_t nblk, ppb;
  jl_evals_chunk(first, num_cus, &nblk, &ppb);
  s->chunk_polys = chunk;
  s->per_unit = ppb;
  s->blocks = nblk;
  s->chunks = total == 0 ? 1 : (total + chunk - 1) / chunk;
  s->terms = 64 * ppb;
  s->bound = s->terms * 128;
}

int launch_jl_proj(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
                   uint32_t num_parts, uint64_t n, void* p_modq, int64_t* p_exact) {
  gh_ctx* ctx = ring->ctx;
  const bool gold = ring->kind == GH_RING_GOLDILOCKS;
  const uint64_t total = (uint64_t)num_parts * n;
  if (p_exact && total >= kJlExactMaxPolys) return GH_E_INVALID_ARGUMENT;
  JlShape shape;
  jl_proj_shape(total, ctx->num_cus, &shape);
  const uint64_t ppw = shape.per_unit, nblk = shape.blocks;
  if (nblk > 0x7fffffffull) return GH_E_INVALID_ARGUMENT;
  void* ws;
  int st = ctx_workspace(ctx, (size_t)nblk * 256 * kProjDigits * sizeof(int32_t), &ws);
  if (st) return st;
  const ChaKey key = chacha_key(seed);
  if (!gold)
    hipLaunchKernelGGL((k_jl_proj<uint32_t, false>), dim3((unsigned)nblk), dim3(64 * kProjWaves),
                       0, ctx->stream, key, (const uint32_t*)parts, n, total, ppw, ring->q,
                       (int32_t*)ws);
  else
    hipLaunchKernelGGL((k_jl_proj<uint64_t, true>), dim3((unsigned)nblk), dim3(64 * kProjWaves),
                       0, ctx->stream, key, (const uint64_t*)parts, n, total, ppw, ring->q,
                       (int32_t*)ws);
  GH_HIP(hipGetLastError());
  if (!gold)
    hipLaunchKernelGGL((k_jl_proj_fin<false>), dim3(rows), dim3(256), 0, ctx->stream,
                       (const int32_t*)ws, (uint32_t)nblk, ring->q, p_modq, p_exact);
  else
    hipLaunchKernelGGL((k_jl_proj_fin<true>), dim3(rows), dim3(256), 0, ctx->stream,
                       (const int32_t*)ws, (uint32_t)nblk, ring->q, p_modq, nullptr);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_jl_flat(gh_ring* ring, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                   uint64_t cols, int8_t* flat) {
  gh_ctx* ctx = ring->ctx;
  const uint64_t n = cols / ring->d;
  if (n == 0 || num_parts == 0) return GH_OK;
  dim3 grid((unsigned)((n + 255) / 256), ring->d, num_parts);
  hipLaunchKernelGGL(k_jl_flat, grid, dim3(256), 0, ctx->stream, chacha_key(seed), rows,
                     num_parts, n, ring->d, flat);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_jl_conj_seed(gh_ring* ring, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                        uint64_t n, void* out) {
  gh_ctx* ctx = ring->ctx;
  if (n == 0 || num_parts == 0) return GH_OK;
  dim3 grid((unsigned)((n + 3) / 4), num_parts);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_jl_conj_seed<uint32_t>, grid, dim3(256), 0, ctx->stream,
                       chacha_key(seed), rows, num_parts, n, (uint32_t)ring->q, (uint32_t*)out);
  else
    hipLaunchKernelGGL(k_jl_conj_seed<uint64_t>, grid, dim3(256), 0, ctx->stream,
                       chacha_key(seed), rows, num_parts, n, (uint64_t)ring->q, (uint64_t*)out);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_jl_conj_flat(gh_ring* ring, const int8_t* flat, uint32_t rows, uint32_t num_parts,
                        uint64_t n, void* out) {
  gh_ctx* ctx = ring->ctx;
  if (n == 0 || num_parts == 0 || rows == 0) return GH_OK;
  dim3 grid((unsigned)((n + 63) / 64), num_parts, rows);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_jl_conj_flat<uint32_t>, grid, dim3(256), 0, ctx->stream, flat,
                       num_parts, n, (uint32_t)ring->q, (uint32_t*)out);
  else
    hipLaunchKernelGGL(k_jl_conj_flat<uint64_t>, grid, dim3(256), 0, ctx->stream, flat,
                       num_parts, n, (uint64_t)ring->q, (uint64_t*)out);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_jl_aggregate(gh_ring* ring, const uint8_t seed[32], const void* scale,
                        uint32_t num_parts, uint64_t n, void* phi) {
  gh_ctx* ctx = ring->ctx;
  const uint64_t total = (uint64_t)num_parts * n;
  if (total == 0) return GH_OK;
  const uint64_t nblk = (total + kAggPolys - 1) / kAggPolys;
  if (nblk > 0x7fffffffull) return GH_E_INVALID_ARGUMENT;
  const ChaKey key = chacha_key(seed);
  if (ring->kind == GH_RING_Q32) {
    hipLaunchKernelGGL((k_jl_agg<uint32_t, 4>), dim3((unsigned)nblk), dim3(512), 0, ctx->stream,
                       key, (const uint32_t*)scale, n, total, ring->q, 0u, (uint32_t*)phi);
    GH_HIP(hipGetLastError());
  } else {
    for (uint32_t dbase = 0; dbase < 8; dbase += 4) {
      hipLaunchKernelGGL((k_jl_agg<uint64_t, 8>), dim3((unsigned)nblk), dim3(512), 0,
                         ctx->stream, key, (const uint64_t*)scale, n, total, ring->q, dbase,
                         (uint64_t*)phi);
      GH_HIP(hipGetLastError());
    }
  }
  return GH_OK;
}

int launch_jl_evals(gh_ring* ring, const uint8_t seed[32], const void* parts, uint32_t num_parts,
                    uint64_t n, void* out) {
  gh_ctx* ctx = ring->ctx;
  const bool gold = ring->kind == GH_RING_GOLDILOCKS;
  const int D = gold ? 8 : 4;  // digits of a lifted witness word
  const uint64_t total = (uint64_t)num_parts * n;
  // one block per CU (the accumulators fill half the register file), polys
  // split evenly in rounds of 8, at most kEvMaxPpb per block: longer
  // witnesses run in chunks added up mod q
  JlShape shape;
  jl_evals_shape(total, ctx->num_cus, &shape);
  const size_t p_bytes = (size_t)shape.blocks * 256 * 64 * D * sizeof(int32_t);
  void* ws;
  int st = ctx_workspace(ctx, p_bytes + shape.blocks * sizeof(uint32_t), &ws);
  if (st) return st;
  // per-block digit spans after the partial sums: small-norm witnesses (the
  // LaBRADOR case) need one digit of four, and the passes skip the zero planes
  uint32_t* span = (uint32_t*)((char*)ws + p_bytes);
  const ChaKey key = chacha_key(seed);
  for (uint64_t c0 = 0; c0 < total; c0 += shape.chunk_polys) {
    const uint64_t polys = std::min<uint64_t>(total - c0, shape.chunk_polys);
    uint64_t nblk, ppb;
    jl_evals_chunk(polys, ctx->num_cus, &nblk, &ppb);
    const uint64_t end = c0 + polys;
    const int acc = c0 > 0;
    if (!gold) {
      hipLaunchKernelGGL((k_jl_evals<uint32_t, 4>), dim3((unsigned)nblk), dim3(512), 0,
                         ctx->stream, key, (const uint32_t*)parts, n, c0, end, ppb, ring->q, 0u,
                         span, (int32_t*)ws);
      GH_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_jl_evals_fin<uint32_t, 4>), dim3(256), dim3(512), 0, ctx->stream,
                         (const int32_t*)ws, (uint32_t)nblk, ring->q, span, acc, (uint32_t*)out);
    } else {
      for (uint32_t dbase = 0; dbase < 8; dbase += 4) {
        hipLaunchKernelGGL((k_jl_evals<uint64_t, 8>), dim3((unsigned)nblk), dim3(512), 0,
                           ctx->stream, key, (const uint64_t*)parts, n, c0, end, ppb, ring->q,
                           dbase, span, (int32_t*)ws);
        GH_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL((k_jl_evals_fin<uint64_t, 8>), dim3(256), dim3(512), 0, ctx->stream,
                         (const int32_t*)ws, (uint32_t)nblk, ring->q, span, acc, (uint64_t*)out);
    }
    GH_HIP(hipGetLastError());
  }
  return GH_OK;
}

}  // namespace gh
