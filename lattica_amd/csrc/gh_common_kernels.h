// Device helpers shared by the commitment and LaBRADOR kernels.
#pragma once

#include "gh_ntt.h"

namespace gh {

// Exact accumulation of a*b into a 96-bit (lo64, hi32) pair: one
// v_mad_u64_u32 with carry-out plus one add-with-carry.
__device__ __forceinline__ void mac96(uint64_t& lo, uint32_t& hi, uint32_t a, uint32_t b) {
  asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
      "v_addc_co_u32 %1, vcc, 0, %1, vcc"
      : "+v"(lo), "+v"(hi)
      : "v"(a), "v"(b)
      : "vcc");
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
__device__ __forceinline__ void gold_mac(GoldAcc& s, uint64_t x, uint64_t y) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
  mac96(s.a, s.ah, x0, y0);
  mac96(s.b, s.bh, x0, y1);
  mac96(s.b, s.bh, x1, y0);
  mac96(s.d, s.dh, x1, y1);
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
