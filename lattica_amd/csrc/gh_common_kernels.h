// Device helpers shared by the commitment and LaBRADOR kernels.
#pragma once

#include "gh_ntt.h"

namespace gh {

// Exact accumulation of a*b into a 96-bit (lo64, hi32) pair: one
// v_mad_u64_u32 with carry-out plus one add-with-carry. The add reads the
// carry that the multiply wrote to VCC; the two wait states between them are
// the ones the compiler itself puts between a VALU write of VCC and a VALU
// read of it as a carry on gfx950.
__device__ __forceinline__ void mac96(uint64_t& lo, uint32_t& hi, uint32_t a, uint32_t b) {
  asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %1, vcc, 0, %1, vcc"
      : "+v"(lo), "+v"(hi)
      : "v"(a), "v"(b)
      : "vcc");
}

// Four independent mac96: the four multiplies first, each carry in its own
// SGPR pair, then the four carry additions, each three instructions after its
// carry is written (no wait state needed).
__device__ __forceinline__ void mac96x4(uint64_t& l0, uint64_t& l1, uint64_t& l2, uint64_t& l3,
                                        uint32_t& h0, uint32_t& h1, uint32_t& h2, uint32_t& h3,
                                        uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                        uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) {
  uint64_t c0, c1, c2, c3;  // carries (SGPR pairs)
  asm("v_mad_u64_u32 %0, %8, %12, %16, %0\n\t"
      "v_mad_u64_u32 %1, %9, %13, %17, %1\n\t"
      "v_mad_u64_u32 %2, %10, %14, %18, %2\n\t"
      "v_mad_u64_u32 %3, %11, %15, %19, %3\n\t"
      "v_addc_co_u32 %4, %8, 0, %4, %8\n\t"
      "v_addc_co_u32 %5, %9, 0, %5, %9\n\t"
      "v_addc_co_u32 %6, %10, 0, %6, %10\n\t"
      "v_addc_co_u32 %7, %11, 0, %7, %11"
      : "+v"(l0), "+v"(l1), "+v"(l2), "+v"(l3), "+v"(h0), "+v"(h1), "+v"(h2), "+v"(h3),
        "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3)
      : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(b0), "v"(b1), "v"(b2), "v"(b3));
}
// the same on four-element accumulator arrays
__device__ __forceinline__ void mac96x4(uint64_t (&lo)[4], uint32_t (&hi)[4], uint32_t a0,
                                        uint32_t a1, uint32_t a2, uint32_t a3, uint32_t b0,
                                        uint32_t b1, uint32_t b2, uint32_t b3) {
  mac96x4(lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3], a0, a1, a2, a3, b0, b1, b2, b3);
}
// lo[r] += a[r] * b for r < N, four at a time
template <int N>
__device__ __forceinline__ void mac96_rows(uint64_t (&lo)[N], uint32_t (&hi)[N],
                                           const uint32_t (&a)[N], uint32_t b) {
#pragma unroll
  for (int r = 0; r + 3 < N; r += 4)
    mac96x4(lo[r], lo[r + 1], lo[r + 2], lo[r + 3], hi[r], hi[r + 1], hi[r + 2], hi[r + 3], a[r],
            a[r + 1], a[r + 2], a[r + 3], b, b, b, b);
#pragma unroll
  for (int r = N & ~3; r < N; r++) mac96(lo[r], hi[r], a[r], b);
}

// Exact lazy accumulation of Goldilocks products: x*y split into 32-bit halves,
// x0*y0 into a (weight 1), x0*y1 + x1*y0 into b (weight 2^32), x1*y1 into d
// (weight 2^64), each a 96-bit (lo64, hi32) sum. Four multiply-accumulates
// per term instead of a 128-bit product, a reduction and a modular add; exact
// for fewer than 2^32 terms.
struct GoldAcc {
  uint64_t a = 0, b = 0, d = 0;
  uint32_t ah = 0, bh = 0, dh = 0;
};
// The four multiply-accumulates are interleaved in one block, each carry in
// its own SGPR pair and read at least two instructions after it is written,
// so no wait state is needed between them.
__device__ __forceinline__ void gold_mac(GoldAcc& s, uint64_t x, uint64_t y) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
  uint64_t c0, c1, c2, c3;  // carries (SGPR pairs)
  asm("v_mad_u64_u32 %0, %6, %10, %12, %0\n\t"   // a += x0 y0
      "v_mad_u64_u32 %1, %7, %10, %13, %1\n\t"   // b += x0 y1
      "v_mad_u64_u32 %2, %8, %11, %13, %2\n\t"   // d += x1 y1
      "v_addc_co_u32 %3, %6, 0, %3, %6\n\t"
      "v_mad_u64_u32 %1, %9, %11, %12, %1\n\t"   // b += x1 y0
      "v_addc_co_u32 %5, %8, 0, %5, %8\n\t"
      "v_addc_co_u32 %4, %7, 0, %4, %7\n\t"
      "v_addc_co_u32 %4, %9, 0, %4, %9"
      : "+v"(s.a), "+v"(s.b), "+v"(s.d), "+v"(s.ah), "+v"(s.bh), "+v"(s.dh), "=&s"(c0),
        "=&s"(c1), "=&s"(c2), "=&s"(c3)
      : "v"(x0), "v"(x1), "v"(y0), "v"(y1));
}
// canonical a + b*2^32 + d*2^64 mod p, with 2^64 = 2^32 - 1 mod p
__device__ __forceinline__ uint64_t gold_reduce(const FGold& f, const GoldAcc& s) {
  const uint64_t ra = f.reduce128(s.a, s.ah), rb = f.reduce128(s.b, s.bh);
  const uint64_t rd = f.reduce128(s.d, s.dh);
  return f.add(f.add(ra, f.mul(rb, 1ull << 32)), f.mul(rd, 0xffffffffull));
}

// t[row] = INTT(sum_p partial[p][row]) with the untwist table `post`, by one
// block of G waves; `red` holds G * 64 words of LDS.
template <class F, int G>
__device__ __forceinline__ void finalize_row_g(F f, const typename F::W* __restrict__ tab,
                                               const typename F::W* __restrict__ partial,
                                               uint32_t nparts, uint32_t rows, uint32_t row,
                                               int post, typename F::W* __restrict__ t_out,
                                               typename F::W* red) {
  using W = typename F::W;
  const int k = threadIdx.x & 63, g = threadIdx.x >> 6;
  // up to 32 independent loads in flight per thread: one round trip (16
  // waves) or a few (4 waves) for the partial counts a device-wide MAC
  // produces (<= 512)
  W acc = 0;
  for (uint32_t p0 = g; p0 < nparts; p0 += G * 32) {
    W v[32];
#pragma unroll
    for (int u = 0; u < 32; u++) {
      const uint32_t p = p0 + G * u;
      v[u] = p < nparts ? partial[((size_t)p * rows + row) * 64 + k] : W(0);
    }
#pragma unroll
    for (int u = 0; u < 32; u++) acc = f.add(acc, v[u]);
  }
  red[g * 64 + k] = acc;
  __syncthreads();
  if (g == 0) {
    W s = red[k];
#pragma unroll
    for (int i = 1; i < G; i++) s = f.add(s, red[i * 64 + k]);
    t_out[(size_t)row * 64 + k] = wave_inv64(f, s, tab, post, k);
  }
}

// ---- held-back finalize of a device-memory commitment, in two stages ----
// A fused commitment writes its partial sums [row][part][slot]. Stage 1 runs
// inside the next commitment's launch: every MAC block sums one slice of
// `per` consecutive partials of one row (one 16-byte or 32-byte load per
// thread, 16 threads per partial) into a group sum [row][group][slot]. Stage 2
// runs inside the launch after: one wave per row sums the row's group sums
// and applies the inverse transform. A standalone finalize (k_finalize_rows)
// does both at once. Everything is field arithmetic on canonical residues, so
// results do not depend on the grouping.
struct FinStage1 {
  const void* src;  // partial sums [rows][n][64]; never null (a dummy when ntasks = 0)
  void* dst;        // group sums [rows][groups][64]
  uint32_t n, rows, groups, per;  // per <= threads / 16
  uint32_t ntasks;                // rows * groups
  // tasks [0, mac_tasks) run in MAC blocks: MAC block (x, y) takes task
  // y * nparts + x
  uint32_t mac_tasks;
  // blocks per row block, after the stage-2 blocks and before the MAC blocks,
  // that run only stage-1 tasks, `tpb` each (at most kStage1MaxTpb): block
  // (x, y) takes tasks mac_tasks + (y * extra + x) * tpb + i
  uint32_t extra, tpb;
};
struct FinStage2 {
  const void* src;  // group sums [rows][n][64]
  void* out;        // t_out, rows polynomials
  uint32_t n, rows, nblk;  // nblk leading blocks of the launch, one wave per row
};

// four consecutive words at p (16-byte aligned)
template <class W>
__device__ __forceinline__ void load4(const W* p, W (&v)[4]) {
  if constexpr (sizeof(W) == 4) {
    const u32x4 x = *reinterpret_cast<const u32x4*>(p);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
  } else {
    const u32x4 x = *reinterpret_cast<const u32x4*>(p);
    const u32x4 y = *reinterpret_cast<const u32x4*>(p + 2);
    v[0] = x.x | ((uint64_t)x.y << 32); v[1] = x.z | ((uint64_t)x.w << 32);
    v[2] = y.x | ((uint64_t)y.y << 32); v[3] = y.z | ((uint64_t)y.w << 32);
  }
}
__device__ __forceinline__ uint32_t shfl_xor_w(uint32_t v, int m) {
  return (uint32_t)__shfl_xor((int)v, m, 64);
}
__device__ __forceinline__ uint64_t shfl_xor_w(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Stage 1, issue: this thread's four words of task `bid` (zero when bid >=
// limit, a block-uniform test, or the partial lies past the group).
template <class W>
__device__ __forceinline__ void stage1_load(const FinStage1& s, uint32_t bid, uint32_t limit,
                                            W (&v)[4]) {
  v[0] = v[1] = v[2] = v[3] = 0;
  if (bid >= limit) return;
  const uint32_t r = bid % s.rows, g = bid / s.rows;
  const uint32_t j = g * s.per + (threadIdx.x >> 4);
  const bool live = (threadIdx.x >> 4) < s.per && j < s.n;
  const W* p = reinterpret_cast<const W*>(s.src) +
               ((size_t)r * s.n + (live ? j : 0)) * 64 + 4 * (threadIdx.x & 15);
  load4(p, v);
  if (!live) v[0] = v[1] = v[2] = v[3] = 0;
}
// Stage 1, sum within each wave (lanes l, l^16, l^32, l^48 hold the same
// slots); lanes 0..15 leave their sums in lds[wave][64]
template <class F>
__device__ __forceinline__ void stage1_wave_sum(const F& f, typename F::W (&v)[4],
                                                typename F::W* lds) {
#pragma unroll
  for (int m = 16; m <= 32; m *= 2)
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = f.add(v[i], shfl_xor_w(v[i], m));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < 16)
#pragma unroll
    for (int i = 0; i < 4; i++) lds[wave * 64 + 4 * lane + i] = v[i];
}
// Stage 1, after a block barrier: the first 64 threads add the waves' sums
// and store the group sum of task bid (< limit)
template <class F>
__device__ __forceinline__ void stage1_store(const F& f, const FinStage1& s, uint32_t bid,
                                             uint32_t limit, const typename F::W* lds,
                                             int waves) {
  using W = typename F::W;
  if (bid >= limit || threadIdx.x >= 64) return;
  const uint32_t r = bid % s.rows, g = bid / s.rows;
  W acc = lds[threadIdx.x];
  for (int w = 1; w < waves; w++) acc = f.add(acc, lds[w * 64 + threadIdx.x]);
  reinterpret_cast<W*>(s.dst)[((size_t)r * s.groups + g) * 64 + threadIdx.x] = acc;
}

// A block that only runs stage-1 tasks: tasks first .. first + tpb - 1 (those
// below ntasks), every load issued before the first sum; `lds` holds
// waves * 64 * tpb words.
template <class F>
__device__ __forceinline__ void stage1_only(const F& f, const FinStage1& s, uint32_t first,
                                            typename F::W* lds, int waves) {
  using W = typename F::W;
  W v[kStage1MaxTpb][4];
#pragma unroll
  for (uint32_t i = 0; i < kStage1MaxTpb; i++)
    if (i < s.tpb) stage1_load(s, first + i, s.ntasks, v[i]);
#pragma unroll
  for (uint32_t i = 0; i < kStage1MaxTpb; i++)
    if (i < s.tpb) stage1_wave_sum(f, v[i], lds + i * waves * 64);
  __syncthreads();
  for (uint32_t i = 0; i < s.tpb; i++)
    stage1_store(f, s, first + i, s.ntasks, lds + i * waves * 64, waves);
}

// Sum of the n polynomials [row][0..n)[64] of `src` by G waves (lane = slot),
// then the inverse transform with the untwist table `post` into out[row];
// `red` holds G * 64 words of LDS (unused for G = 1).
template <class F, int G>
__device__ __forceinline__ void sum_rows_inv(const F& f, const typename F::W* __restrict__ tab,
                                             int post, const typename F::W* __restrict__ src,
                                             uint32_t n, uint32_t row, int wave,
                                             typename F::W* __restrict__ out,
                                             typename F::W* red) {
  using W = typename F::W;
  const int k = threadIdx.x & 63;
  const W* p = src + (size_t)row * n * 64 + k;
  W acc = 0;
  constexpr int U = G == 1 ? 16 : 32;  // loads in flight per thread
  for (uint32_t j0 = wave; j0 < n; j0 += G * U) {
    W v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t j = j0 + G * u;
      v[u] = j < n ? p[(size_t)j * 64] : W(0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc = f.add(acc, v[u]);
  }
  if constexpr (G > 1) {
    red[wave * 64 + k] = acc;
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int i = 1; i < G; i++) acc = f.add(acc, red[i * 64 + k]);
  }
  out[(size_t)row * 64 + k] = wave_inv64(f, acc, tab, post, k);
}

// Standalone finalize of up to two held-back commitments: block b < rows_a
// finalizes row b of segment a, the next rows_b blocks segment b. Each
// segment's sums are laid out [row][n][64].
struct FinSeg {
  const void* src;
  void* out;
  uint32_t n, rows;
};
template <class F, int G>
__global__ __launch_bounds__(64 * G) void k_finalize_rows(F f, const typename F::W* __restrict__ tab,
                                                          int post, FinSeg a, FinSeg b) {
  using W = typename F::W;
  __shared__ W red[G * 64];
  const bool in_a = blockIdx.x < a.rows;
  const FinSeg& s = in_a ? a : b;
  const uint32_t row = in_a ? blockIdx.x : blockIdx.x - a.rows;
  sum_rows_inv<F, G>(f, tab, post, reinterpret_cast<const W*>(s.src), s.n, row,
                     (int)(threadIdx.x >> 6), reinterpret_cast<W*>(s.out), red);
}

template <class F>
__device__ __forceinline__ void finalize_row(F f, const typename F::W* __restrict__ tab,
                                             const typename F::W* __restrict__ partial,
                                             uint32_t nparts, uint32_t rows, uint32_t row, int post,
                                             typename F::W* __restrict__ t_out) {
  __shared__ typename F::W red[16 * 64];
  finalize_row_g<F, 16>(f, tab, partial, nparts, rows, row, post, t_out, red);
}

template <class F>
__global__ __launch_bounds__(1024) void k_finalize(F f, const typename F::W* __restrict__ tab,
                                                   const typename F::W* __restrict__ partial,
                                                   uint32_t nparts, uint32_t rows, int post,
                                                   typename F::W* __restrict__ t_out) {
  finalize_row(f, tab, partial, nparts, rows, blockIdx.x, post, t_out);
}

// The same with 4 waves per row, for many rows and few partials (the Gram
// matrix's 528 polynomials): at 104-128 VGPRs a 16-wave block fills a CU, so
// 528 rows would take three rounds of 256 blocks; 4-wave blocks take one.
template <class F>
__global__ __launch_bounds__(256) void k_finalize4(F f, const typename F::W* __restrict__ tab,
                                                   const typename F::W* __restrict__ partial,
                                                   uint32_t nparts, uint32_t rows, int post,
                                                   typename F::W* __restrict__ t_out) {
  __shared__ typename F::W red[4 * 64];
  finalize_row_g<F, 4>(f, tab, partial, nparts, rows, blockIdx.x, post, t_out, red);
}

}  // namespace gh
