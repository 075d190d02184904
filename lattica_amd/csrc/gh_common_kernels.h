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

// t[row] = INTT(sum_p partial[p][row]) with the untwist table `post`, by one
// block of 1024 threads.
template <class F>
__device__ __forceinline__ void finalize_row(F f, const typename F::W* __restrict__ tab,
                                             const typename F::W* __restrict__ partial,
                                             uint32_t nparts, uint32_t rows, uint32_t row, int post,
                                             typename F::W* __restrict__ t_out) {
  using W = typename F::W;
  const int k = threadIdx.x & 63, g = threadIdx.x >> 6;
  // up to 32 independent loads in flight per thread: one round trip for the
  // partial counts a device-wide MAC produces (<= 512)
  W acc = 0;
  for (uint32_t p0 = g; p0 < nparts; p0 += 16 * 32) {
    W v[32];
#pragma unroll
    for (int u = 0; u < 32; u++) {
      const uint32_t p = p0 + 16 * u;
      v[u] = p < nparts ? partial[((size_t)p * rows + row) * 64 + k] : W(0);
    }
#pragma unroll
    for (int u = 0; u < 32; u++) acc = f.add(acc, v[u]);
  }
  __shared__ W red[16][64];
  red[g][k] = acc;
  __syncthreads();
  if (g == 0) {
    W s = red[0][k];
#pragma unroll
    for (int i = 1; i < 16; i++) s = f.add(s, red[i][k]);
    t_out[(size_t)row * 64 + k] = wave_inv64(f, s, tab, post, k);
  }
}

template <class F>
__global__ __launch_bounds__(1024) void k_finalize(F f, const typename F::W* __restrict__ tab,
                                                   const typename F::W* __restrict__ partial,
                                                   uint32_t nparts, uint32_t rows, int post,
                                                   typename F::W* __restrict__ t_out) {
  finalize_row(f, tab, partial, nparts, rows, blockIdx.x, post, t_out);
}

}  // namespace gh
