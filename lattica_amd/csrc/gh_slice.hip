// Ring slice hooks on device words: elementwise add, subtract, scalar
// multiply, scaled add, and the dot product (algebra/src/arith/ring.rs:75-135,
// PrimeField overrides at algebra/src/arith/prime.rs:433-491).
//
// Slices are `len` canonical words of any length at any word alignment. The
// elementwise kernels are HBM bound: a thread moves one 16-byte vector of each
// operand per step when both pointers are 16-byte aligned (a scalar loop
// otherwise), and the words past the last whole vector are done by the first
// threads of the grid.
//
// The dot product sums exact products: q < 2^32 products (< 2^64) into a
// 64-bit sum plus a 32-bit carry count per thread, Goldilocks products as four
// 32 x 32-bit partial products into three such sums (GoldAcc). Each thread
// reduces its sum once, a block adds its threads' residues mod q, and a
// second one-block kernel adds the block residues into the result word.
#include "gh_internal.h"
#include "gh_common_kernels.h"

namespace gh {
namespace {

// One word of a slice operation. For q32 the scalar is the pair
// (v * 2^32 mod q, that * q^-1 mod 2^32) so that mul_tw gives x * v mod q;
// for Goldilocks c[0] is v itself.
template <class F, int OP>
__device__ __forceinline__ typename F::W slice_word(const F& f, typename F::W d,
                                                    typename F::W s,
                                                    const typename F::W* c) {
  if constexpr (OP == kSliceAdd) return f.add(d, s);
  if constexpr (OP == kSliceSub) return f.sub(d, s);
  if constexpr (OP == kSliceScale) return f.mul_tw(d, c);
  return f.add(d, f.mul_tw(s, c));
}

template <class F, int OP, bool VEC>
__global__ void __launch_bounds__(256) k_slice(F f, typename F::W* dst,
                                               const typename F::W* src, typename F::W c0,
                                               typename F::W c1, size_t len) {
  using W = typename F::W;
  const W c[2] = {c0, c1};
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t done = 0;
  if constexpr (VEC) {
    constexpr int V = 16 / sizeof(W);
    typedef W Vec __attribute__((ext_vector_type(V)));
    const size_t nvec = len / V;
    Vec* dv = reinterpret_cast<Vec*>(dst);
    const Vec* sv = reinterpret_cast<const Vec*>(src);
    for (size_t i = tid; i < nvec; i += stride) {
      Vec d = dv[i];
      Vec s = d;
      if constexpr (OP != kSliceScale) s = sv[i];
#pragma unroll
      for (int j = 0; j < V; j++) d[j] = slice_word<F, OP>(f, d[j], s[j], c);
      dv[i] = d;
    }
    done = nvec * V;
  }
  for (size_t i = done + tid; i < len; i += stride) {
    const W s = OP == kSliceScale ? W(0) : src[i];
    dst[i] = slice_word<F, OP>(f, dst[i], s, c);
  }
}

// Exact per-thread sum of products, reduced to a canonical residue.
struct DotQ32 {
  uint64_t lo = 0;
  uint32_t hi = 0;
  __device__ __forceinline__ void mac(uint32_t a, uint32_t b) {
    const uint64_t p = (uint64_t)a * b;
    lo += p;
    hi += lo < p;
  }
  // (hi * 2^64 + lo) mod q: REDC by 2^64, then a Montgomery multiply by
  // r3 = 2^96 mod q restores the factor (x * 2^96 * 2^-32 = x * 2^64)
  __device__ __forceinline__ uint32_t value(const FQ32& f, uint32_t r3) const {
    const uint32_t h2 = hi >= f.q ? hi - f.q : hi;
    const uint32_t x = f.reduce96_r2((uint32_t)lo, (uint32_t)(lo >> 32), h2);
    return f.mont(x, r3);
  }
};

template <bool VEC>
__device__ __forceinline__ uint32_t dot_thread_q32(const FQ32& f, uint32_t r3,
                                                   const uint32_t* a, const uint32_t* b,
                                                   size_t len) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  DotQ32 acc;
  size_t done = 0;
  if constexpr (VEC) {
    const size_t nvec = len / 4;
    const u32x4* av = reinterpret_cast<const u32x4*>(a);
    const u32x4* bv = reinterpret_cast<const u32x4*>(b);
    for (size_t i = tid; i < nvec; i += stride) {
      const u32x4 x = av[i], y = bv[i];
#pragma unroll
      for (int j = 0; j < 4; j++) acc.mac(x[j], y[j]);
    }
    done = nvec * 4;
  }
  for (size_t i = done + tid; i < len; i += stride) acc.mac(a[i], b[i]);
  return acc.value(f, r3);
}

template <bool VEC>
__device__ __forceinline__ uint64_t dot_thread_gold(const FGold& f, const uint64_t* a,
                                                    const uint64_t* b, size_t len) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  GoldAcc acc;
  size_t done = 0;
  if constexpr (VEC) {
    typedef uint64_t V2 __attribute__((ext_vector_type(2)));
    const size_t nvec = len / 2;
    const V2* av = reinterpret_cast<const V2*>(a);
    const V2* bv = reinterpret_cast<const V2*>(b);
    for (size_t i = tid; i < nvec; i += stride) {
      const V2 x = av[i], y = bv[i];
      gold_mac(acc, x[0], y[0]);
      gold_mac(acc, x[1], y[1]);
    }
    done = nvec * 2;
  }
  for (size_t i = done + tid; i < len; i += stride) gold_mac(acc, a[i], b[i]);
  return gold_reduce(f, acc);
}

// Sum of the block's 256 residues mod q; the result is valid in thread 0.
template <class F>
__device__ __forceinline__ typename F::W block_sum(const F& f, typename F::W v) {
  using W = typename F::W;
  __shared__ W red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = f.add(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) v = f.add(f.add(red[0], red[1]), f.add(red[2], red[3]));
  return v;
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_dot_q32(FQ32 f, uint32_t r3, const uint32_t* a,
                                                 const uint32_t* b, size_t len,
                                                 uint32_t* partial) {
  const uint32_t v = block_sum(f, dot_thread_q32<VEC>(f, r3, a, b, len));
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_dot_gold(FGold f, const uint64_t* a, const uint64_t* b,
                                                  size_t len, uint64_t* partial) {
  const uint64_t v = block_sum(f, dot_thread_gold<VEC>(f, a, b, len));
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// out[0] = sum of n canonical residues mod q (one block)
template <class F>
__global__ void __launch_bounds__(256) k_sum_words(F f, const typename F::W* in, uint32_t n,
                                                   typename F::W* out) {
  using W = typename F::W;
  W v = 0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) v = f.add(v, in[i]);
  v = block_sum(f, v);
  if (threadIdx.x == 0) out[0] = v;
}

unsigned slice_grid(size_t units, int num_cus) {
  const size_t cap = (size_t)num_cus * 8;  // 8 blocks of 256 threads per CU
  const size_t g = (units + 255) / 256;
  return (unsigned)(g == 0 ? 1 : (g < cap ? g : cap));
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <class F, int OP>
int slice_impl(gh_ring* ring, F f, void* dst, const void* src, typename F::W c0,
               typename F::W c1, size_t len) {
  using W = typename F::W;
  gh_ctx* ctx = ring->ctx;
  constexpr size_t V = 16 / sizeof(W);
  const bool vec = aligned16(dst) && (OP == kSliceScale || aligned16(src));
  const unsigned grid = slice_grid(vec ? len / V + 1 : len, ctx->num_cus);
  if (vec)
    hipLaunchKernelGGL((k_slice<F, OP, true>), dim3(grid), dim3(256), 0, ctx->stream, f, (W*)dst,
                       (const W*)src, c0, c1, len);
  else
    hipLaunchKernelGGL((k_slice<F, OP, false>), dim3(grid), dim3(256), 0, ctx->stream, f, (W*)dst,
                       (const W*)src, c0, c1, len);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <int OP>
int slice_dispatch(gh_ring* ring, void* dst, const void* src, uint64_t scalar, size_t len) {
  if (ring->kind == GH_RING_Q32) {
    const uint32_t q = (uint32_t)ring->q;
    // (v * 2^32 mod q, that * q^-1 mod 2^32): mul_tw(x) = x * v mod q
    const uint32_t w = (uint32_t)((scalar << 32) % q);
    const uint32_t wq = w * ring->qinv32;
    return slice_impl<FQ32, OP>(ring, FQ32{q, ring->qinv32}, dst, src, w, wq, len);
  }
  return slice_impl<FGold, OP>(ring, FGold{}, dst, src, scalar, 0, len);
}

}  // namespace

int launch_slice(gh_ring* ring, int op, void* dst, const void* src, uint64_t scalar,
                 size_t len) {
  if (len == 0) return GH_OK;
  switch (op) {
    case kSliceAdd: return slice_dispatch<kSliceAdd>(ring, dst, src, scalar, len);
    case kSliceSub: return slice_dispatch<kSliceSub>(ring, dst, src, scalar, len);
    case kSliceScale: return slice_dispatch<kSliceScale>(ring, dst, src, scalar, len);
    case kSliceAxpy: return slice_dispatch<kSliceAxpy>(ring, dst, src, scalar, len);
  }
  return GH_E_INVALID_ARGUMENT;
}

size_t dot_partial_words(gh_ring* ring, size_t len) {
  return slice_grid(len, ring->ctx->num_cus);
}

int launch_dot(gh_ring* ring, const void* a, const void* b, size_t len, void* partial,
               void* out) {
  gh_ctx* ctx = ring->ctx;
  const unsigned grid = slice_grid(len, ctx->num_cus);
  const bool vec = aligned16(a) && aligned16(b);
  if (ring->kind == GH_RING_Q32) {
    const uint32_t q = (uint32_t)ring->q;
    const FQ32 f{q, ring->qinv32};
    const uint32_t r3 = (uint32_t)powmod(2, 96, q);
    if (len) {
      if (vec)
        hipLaunchKernelGGL((k_dot_q32<true>), dim3(grid), dim3(256), 0, ctx->stream, f, r3,
                           (const uint32_t*)a, (const uint32_t*)b, len, (uint32_t*)partial);
      else
        hipLaunchKernelGGL((k_dot_q32<false>), dim3(grid), dim3(256), 0, ctx->stream, f, r3,
                           (const uint32_t*)a, (const uint32_t*)b, len, (uint32_t*)partial);
    }
    hipLaunchKernelGGL((k_sum_words<FQ32>), dim3(1), dim3(256), 0, ctx->stream, f,
                       (const uint32_t*)partial, len ? grid : 0u, (uint32_t*)out);
  } else {
    const FGold f{};
    if (len) {
      if (vec)
        hipLaunchKernelGGL((k_dot_gold<true>), dim3(grid), dim3(256), 0, ctx->stream, f,
                           (const uint64_t*)a, (const uint64_t*)b, len, (uint64_t*)partial);
      else
        hipLaunchKernelGGL((k_dot_gold<false>), dim3(grid), dim3(256), 0, ctx->stream, f,
                           (const uint64_t*)a, (const uint64_t*)b, len, (uint64_t*)partial);
    }
    hipLaunchKernelGGL((k_sum_words<FGold>), dim3(1), dim3(256), 0, ctx->stream, f,
                       (const uint64_t*)partial, len ? grid : 0u, (uint64_t*)out);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // namespace gh
