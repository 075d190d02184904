// Batched transforms, negacyclic products, key layout and the Ajtai commitment
// for d = 64 on gfx950.
//
// Transforms run one polynomial per thread: a block stages a tile of
// polynomials through LDS with coalesced 16-byte global accesses (row stride 65
// words, so both the coalesced tile fill and the per-thread polynomial reads are
// bank-conflict free), each thread then runs the whole 64-point transform in
// registers (gh_ntt.h) and the tile leaves the same way.
//
// The commitment t = A*s streams the resident NTT-domain key from HBM:
//   1. twisted forward NTT of the witness (thread per polynomial);
//   2. k_mac: lane = slot, one column of 16 rows per wave step; the key layout
//      [col][row][slot] makes every wave load a contiguous 256-byte row slice
//      and a block sweep contiguous 4 KB columns. Products are accumulated
//      exactly (96-bit for q < 2^32) and reduced once per block;
//   3. k_finalize: sums the per-block partials and runs the inverse transform
//      with one slot per lane.
#include "gh_internal.h"
#include "gh_ntt.h"
#include "gh_common_kernels.h"

namespace gh {

constexpr int kStride = 65;  // LDS row stride in words

// ---------------------------------------------------------------------------
// LDS tile staging
// ---------------------------------------------------------------------------

template <class W, int TILE>
__device__ __forceinline__ void tile_load(W* __restrict__ lds, const W* __restrict__ g,
                                          size_t words) {
  constexpr int V = 16 / sizeof(W);  // words per 16-byte access
  const int tid = threadIdx.x;
  const size_t vecs = words / V;     // words is a multiple of 64
  using Vec = uint4;
  const Vec* gv = reinterpret_cast<const Vec*>(g);
#pragma unroll 4
  for (size_t i = tid; i < vecs; i += TILE) {
    Vec v = gv[i];
    const W* w = reinterpret_cast<const W*>(&v);
    const size_t e = i * V;
#pragma unroll
    for (int c = 0; c < V; c++) {
      const size_t ec = e + c;
      lds[(ec >> 6) * kStride + (ec & 63)] = w[c];
    }
  }
}

template <class W, int TILE>
__device__ __forceinline__ void tile_store(W* __restrict__ g, const W* __restrict__ lds,
                                           size_t words) {
  constexpr int V = 16 / sizeof(W);
  const int tid = threadIdx.x;
  const size_t vecs = words / V;
  using Vec = uint4;
  Vec* gv = reinterpret_cast<Vec*>(g);
#pragma unroll 4
  for (size_t i = tid; i < vecs; i += TILE) {
    Vec v;
    W* w = reinterpret_cast<W*>(&v);
    const size_t e = i * V;
#pragma unroll
    for (int c = 0; c < V; c++) {
      const size_t ec = e + c;
      w[c] = lds[(ec >> 6) * kStride + (ec & 63)];
    }
    gv[i] = v;
  }
}

template <class W>
__device__ __forceinline__ void lds_read_poly(const W* lds, int t, W (&x)[64]) {
#pragma unroll
  for (int k = 0; k < 64; k++) x[k] = lds[t * kStride + k];
}
template <class W>
__device__ __forceinline__ void lds_write_poly(W* lds, int t, const W (&x)[64]) {
#pragma unroll
  for (int k = 0; k < 64; k++) lds[t * kStride + k] = x[k];
}

// ---------------------------------------------------------------------------
// Batched transforms
// ---------------------------------------------------------------------------

template <class F, int MODE, int TILE>
__global__ __launch_bounds__(TILE) void k_xform(F f, const typename F::W* __restrict__ tab,
                                                const typename F::W* src,
                                                typename F::W* dst, size_t count) {
  using W = typename F::W;
  __shared__ W lds[TILE * kStride];
  const int t = threadIdx.x;
  for (size_t tile = blockIdx.x; tile * TILE < count; tile += gridDim.x) {
    const size_t p0 = tile * TILE;
    const size_t np = count - p0 < (size_t)TILE ? count - p0 : (size_t)TILE;
    tile_load<W, TILE>(lds, src + p0 * 64, np * 64);
    __syncthreads();
    if ((size_t)t < np) {
      W x[64];
      lds_read_poly(lds, t, x);
      if (MODE == kTwistedFwd) fwd64<true>(f, x, tab);
      else if (MODE == kCyclicFwd) fwd64<false>(f, x, tab);
      else if (MODE == kTwistedInv) inv64(f, x, tab, kTabUntwist);
      else inv64(f, x, tab, kTabUnscale);
      lds_write_poly(lds, t, x);
    }
    __syncthreads();
    tile_store<W, TILE>(dst + p0 * 64, lds, np * 64);
    __syncthreads();
  }
}

template <class F, int TILE>
__global__ __launch_bounds__(TILE) void k_negacyclic_mul(F f, const typename F::W* __restrict__ tab,
                                                         const typename F::W* a,
                                                         const typename F::W* b,
                                                         typename F::W* out, size_t count) {
  using W = typename F::W;
  __shared__ W lds[TILE * kStride];
  const int t = threadIdx.x;
  for (size_t tile = blockIdx.x; tile * TILE < count; tile += gridDim.x) {
    const size_t p0 = tile * TILE;
    const size_t np = count - p0 < (size_t)TILE ? count - p0 : (size_t)TILE;
    const bool active = (size_t)t < np;
    W x[64], y[64];
    tile_load<W, TILE>(lds, a + p0 * 64, np * 64);
    __syncthreads();
    if (active) lds_read_poly(lds, t, x);
    __syncthreads();
    tile_load<W, TILE>(lds, b + p0 * 64, np * 64);
    __syncthreads();
    if (active) {
      lds_read_poly(lds, t, y);
      fwd64<true>(f, x, tab);
      fwd64<true>(f, y, tab);
#pragma unroll
      for (int k = 0; k < 64; k++) x[k] = f.mont(x[k], y[k]);
      inv64(f, x, tab, kTabUntwistMul);
      lds_write_poly(lds, t, x);
    }
    __syncthreads();
    tile_store<W, TILE>(out + p0 * 64, lds, np * 64);
    __syncthreads();
  }
}

// out = a (.) b, canonical: Montgomery product then a Montgomery multiply by
// R^2 (stored as the kTabUntwistMul-free constant r2m) restores the R factor.
template <class F>
__global__ void k_pointwise(F f, typename F::W r2m, const typename F::W* a,
                            const typename F::W* b, typename F::W* out, size_t words) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < words; i += stride) {
    typename F::W x = f.mont(a[i], b[i]);
    out[i] = f.mont(x, r2m);
  }
}

// ---------------------------------------------------------------------------
// Key layout: row-major [row][col][slot] -> [col][row][slot]
// ---------------------------------------------------------------------------

template <class W>
__global__ void k_key_layout(const W* __restrict__ src, W* __restrict__ dst, uint32_t rows,
                             uint32_t cols, uint32_t col_offset, uint32_t total_cols) {
  const size_t polys = (size_t)rows * cols;
  const int lane = threadIdx.x & 63;
  size_t p = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 6;
  for (; p < polys; p += stride) {
    const uint32_t row = (uint32_t)(p / cols), col = (uint32_t)(p % cols);
    dst[(((size_t)(col + col_offset)) * rows + row) * 64 + lane] = src[p * 64 + lane];
  }
}

// ---------------------------------------------------------------------------
// Commitment MAC
// ---------------------------------------------------------------------------

constexpr int kMacWaves = 16;


// Partial t_ntt for rows [row0, row0+RB) over this block's column range,
// written as partial[blockIdx.x][row][slot] (q32: carries the factor 2^-64).
template <int RB>
__global__ __launch_bounds__(64 * kMacWaves) void k_mac_q32(
    FQ32 f, const uint32_t* __restrict__ A, const uint32_t* __restrict__ s_ntt,
    uint32_t* __restrict__ partial, uint32_t rows, uint32_t cols, uint32_t cols_per_block) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t row0 = blockIdx.y * RB;
  const uint32_t c_begin = blockIdx.x * cols_per_block;
  uint32_t c_end = c_begin + cols_per_block;
  if (c_end > cols) c_end = cols;

  uint64_t lo[RB];
  uint32_t hi[RB];
#pragma unroll
  for (int r = 0; r < RB; r++) { lo[r] = 0; hi[r] = 0; }

  const size_t col_stride = (size_t)rows * 64;
  uint32_t c = c_begin + wave;
  // two columns per step for memory-level parallelism
  for (; c + kMacWaves < c_end; c += 2 * kMacWaves) {
    const uint32_t* a0 = A + (size_t)c * col_stride + (size_t)row0 * 64 + lane;
    const uint32_t* a1 = a0 + kMacWaves * col_stride;
    uint32_t s0 = s_ntt[(size_t)c * 64 + lane];
    uint32_t s1 = s_ntt[(size_t)(c + kMacWaves) * 64 + lane];
    uint32_t va[RB], vb[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) va[r] = __builtin_nontemporal_load(a0 + r * 64);
#pragma unroll
    for (int r = 0; r < RB; r++) vb[r] = __builtin_nontemporal_load(a1 + r * 64);
#pragma unroll
    for (int r = 0; r < RB; r++) mac96(lo[r], hi[r], va[r], s0);
#pragma unroll
    for (int r = 0; r < RB; r++) mac96(lo[r], hi[r], vb[r], s1);
  }
  if (c < c_end) {
    const uint32_t* a0 = A + (size_t)c * col_stride + (size_t)row0 * 64 + lane;
    uint32_t s0 = s_ntt[(size_t)c * 64 + lane];
#pragma unroll
    for (int r = 0; r < RB; r++) mac96(lo[r], hi[r], __builtin_nontemporal_load(a0 + r * 64), s0);
  }

  __shared__ uint32_t red[kMacWaves][RB][64];
#pragma unroll
  for (int r = 0; r < RB; r++)
    red[wave][r][lane] = f.reduce96_r2((uint32_t)lo[r], (uint32_t)(lo[r] >> 32), hi[r]);
  __syncthreads();
  for (int o = threadIdx.x; o < RB * 64; o += 64 * kMacWaves) {
    const int r = o >> 6, k = o & 63;
    uint32_t acc = red[0][r][k];
#pragma unroll
    for (int w = 1; w < kMacWaves; w++) acc = f.add(acc, red[w][r][k]);
    partial[((size_t)blockIdx.x * rows + row0 + r) * 64 + k] = acc;
  }
}

// Goldilocks: modular multiply-add per term, canonical partials.
template <int RB>
__global__ __launch_bounds__(64 * kMacWaves) void k_mac_gold(
    FGold f, const uint64_t* __restrict__ A, const uint64_t* __restrict__ s_ntt,
    uint64_t* __restrict__ partial, uint32_t rows, uint32_t cols, uint32_t cols_per_block) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t row0 = blockIdx.y * RB;
  const uint32_t c_begin = blockIdx.x * cols_per_block;
  uint32_t c_end = c_begin + cols_per_block;
  if (c_end > cols) c_end = cols;
  uint64_t acc[RB];
#pragma unroll
  for (int r = 0; r < RB; r++) acc[r] = 0;
  const size_t col_stride = (size_t)rows * 64;
  for (uint32_t c = c_begin + wave; c < c_end; c += kMacWaves) {
    const uint64_t* a0 = A + (size_t)c * col_stride + (size_t)row0 * 64 + lane;
    const uint64_t s0 = s_ntt[(size_t)c * 64 + lane];
    uint64_t va[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) va[r] = __builtin_nontemporal_load(a0 + r * 64);
#pragma unroll
    for (int r = 0; r < RB; r++) acc[r] = f.add(acc[r], f.mul(va[r], s0));
  }
  __shared__ uint64_t red[kMacWaves][RB][64];
#pragma unroll
  for (int r = 0; r < RB; r++) red[wave][r][lane] = acc[r];
  __syncthreads();
  for (int o = threadIdx.x; o < RB * 64; o += 64 * kMacWaves) {
    const int r = o >> 6, k = o & 63;
    uint64_t a = red[0][r][k];
#pragma unroll
    for (int w = 1; w < kMacWaves; w++) a = f.add(a, red[w][r][k]);
    partial[((size_t)blockIdx.x * rows + row0 + r) * 64 + k] = a;
  }
}

// Flags any word >= q (input validation for canonical residues).
template <class W>
__global__ void k_check_canonical(const W* __restrict__ p, size_t words, W q, int* bad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  int found = 0;
  for (; i < words; i += stride) found |= p[i] >= q;
  if (found) atomicOr(bad, 1);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline unsigned grid_for(size_t tiles, int num_cus, int per_cu) {
  size_t cap = (size_t)num_cus * per_cu;
  size_t g = tiles < cap ? tiles : cap;
  return (unsigned)(g ? g : 1);
}

template <class F, int MODE>
static int xform_impl(gh_ring* ring, F f, const void* src, void* dst, size_t count) {
  using W = typename F::W;
  gh_ctx* ctx = ring->ctx;
  const W* tab = (const W*)ring->dev_tab;
  constexpr int BIG = sizeof(W) == 4 ? 256 : 128;
  if (count >= (size_t)BIG * ctx->num_cus * 2) {
    unsigned grid = grid_for((count + BIG - 1) / BIG, ctx->num_cus, 2);
    hipLaunchKernelGGL((k_xform<F, MODE, BIG>), dim3(grid), dim3(BIG), 0, ctx->stream, f, tab,
                       (const W*)src, (W*)dst, count);
  } else {
    unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, 8);
    hipLaunchKernelGGL((k_xform<F, MODE, 64>), dim3(grid), dim3(64), 0, ctx->stream, f, tab,
                       (const W*)src, (W*)dst, count);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <class F>
static int xform_mode(gh_ring* ring, F f, int mode, const void* src, void* dst, size_t count) {
  switch (mode) {
    case kTwistedFwd: return xform_impl<F, kTwistedFwd>(ring, f, src, dst, count);
    case kTwistedInv: return xform_impl<F, kTwistedInv>(ring, f, src, dst, count);
    case kCyclicFwd: return xform_impl<F, kCyclicFwd>(ring, f, src, dst, count);
    case kCyclicInv: return xform_impl<F, kCyclicInv>(ring, f, src, dst, count);
  }
  return GH_E_INVALID_ARGUMENT;
}

int launch_xform(gh_ring* ring, int mode, const void* src, void* dst, size_t count) {
  if (count == 0) return GH_OK;
  if (ring->kind == GH_RING_Q32)
    return xform_mode(ring, FQ32{(uint32_t)ring->q, ring->qinv32}, mode, src, dst, count);
  return xform_mode(ring, FGold{}, mode, src, dst, count);
}

template <class F>
static int negacyclic_impl(gh_ring* ring, F f, const void* a, const void* b, void* out,
                           size_t count) {
  using W = typename F::W;
  gh_ctx* ctx = ring->ctx;
  const W* tab = (const W*)ring->dev_tab;
  constexpr int BIG = sizeof(W) == 4 ? 256 : 128;
  if (count >= (size_t)BIG * ctx->num_cus * 2) {
    unsigned grid = grid_for((count + BIG - 1) / BIG, ctx->num_cus, 2);
    hipLaunchKernelGGL((k_negacyclic_mul<F, BIG>), dim3(grid), dim3(BIG), 0, ctx->stream, f, tab,
                       (const W*)a, (const W*)b, (W*)out, count);
  } else {
    unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, 8);
    hipLaunchKernelGGL((k_negacyclic_mul<F, 64>), dim3(grid), dim3(64), 0, ctx->stream, f, tab,
                       (const W*)a, (const W*)b, (W*)out, count);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_negacyclic_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count) {
  if (count == 0) return GH_OK;
  if (ring->kind == GH_RING_Q32)
    return negacyclic_impl(ring, FQ32{(uint32_t)ring->q, ring->qinv32}, a, b, out, count);
  return negacyclic_impl(ring, FGold{}, a, b, out, count);
}

int launch_pointwise_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count) {
  if (count == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  const size_t words = count * 64;
  unsigned grid = grid_for((words + 255) / 256, ctx->num_cus, 8);
  if (ring->kind == GH_RING_Q32) {
    // mont(a, b) = a*b/R and mont(a*b/R, R^2 mod q) = a*b
    const uint64_t r = (1ull << 32) % ring->q;
    const uint32_t r2 = (uint32_t)(r * r % ring->q);
    hipLaunchKernelGGL((k_pointwise<FQ32>), dim3(grid), dim3(256), 0, ctx->stream,
                       FQ32{(uint32_t)ring->q, ring->qinv32}, r2, (const uint32_t*)a,
                       (const uint32_t*)b, (uint32_t*)out, words);
  } else {
    hipLaunchKernelGGL((k_pointwise<FGold>), dim3(grid), dim3(256), 0, ctx->stream, FGold{},
                       (uint64_t)1, (const uint64_t*)a, (const uint64_t*)b, (uint64_t*)out, words);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_key_layout(gh_ring* ring, const void* src, uint32_t rows, uint32_t cols,
                      uint32_t col_offset, uint32_t total_cols, void* dst, int domain) {
  (void)domain;
  if ((size_t)rows * cols == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  const size_t polys = (size_t)rows * cols;
  unsigned grid = grid_for((polys + 3) / 4, ctx->num_cus, 16);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_key_layout<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t*)src, (uint32_t*)dst, rows, cols, col_offset, total_cols);
  else
    hipLaunchKernelGGL(k_key_layout<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint64_t*)src, (uint64_t*)dst, rows, cols, col_offset, total_cols);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_check_canonical(gh_ring* ring, const void* p, size_t words, int* bad) {
  if (words == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  unsigned grid = grid_for((words + 255) / 256, ctx->num_cus, 4);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_check_canonical<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t*)p, words, (uint32_t)ring->q, bad);
  else
    hipLaunchKernelGGL(k_check_canonical<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint64_t*)p, words, (uint64_t)ring->q, bad);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <int RB>
static int mac_launch(gh_key* key, const void* s_ntt, uint32_t cols, void* partial,
                      uint32_t nparts, uint32_t cpb) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  dim3 grid(nparts, key->rows / RB);
  if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_mac_q32<RB>), grid, dim3(64 * kMacWaves), 0, ctx->stream,
                       FQ32{(uint32_t)ring->q, ring->qinv32}, (const uint32_t*)key->dev_a,
                       (const uint32_t*)s_ntt, (uint32_t*)partial, key->rows, cols, cpb);
  else
    hipLaunchKernelGGL((k_mac_gold<RB>), grid, dim3(64 * kMacWaves), 0, ctx->stream, FGold{},
                       (const uint64_t*)key->dev_a, (const uint64_t*)s_ntt, (uint64_t*)partial,
                       key->rows, cols, cpb);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// t = A[:, :cols] * s; s (cols polys) and t (rows polys) are device pointers.
int launch_commit(gh_key* key, const void* s_dev, uint32_t cols, int domain, void* t_dev) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const uint32_t rows = key->rows;
  // partial count: about one block per CU, at least one column per wave
  uint32_t nparts = (uint32_t)ctx->num_cus;
  uint32_t max_parts = (cols + kMacWaves - 1) / kMacWaves;
  if (nparts > max_parts) nparts = max_parts;
  if (nparts == 0) nparts = 1;
  uint32_t cpb = (cols + nparts - 1) / nparts;
  nparts = (cols + cpb - 1) / cpb;

  const size_t s_bytes = (size_t)cols * 64 * wb;
  const size_t p_bytes = (size_t)nparts * rows * 64 * wb;
  void* ws;
  int st = ctx_workspace(ctx, s_bytes + p_bytes, &ws);
  if (st) return st;
  void* s_ntt = ws;
  void* partial = (char*)ws + s_bytes;
  hipEvent_t e0 = prof_mark(ctx);
  if (domain == GH_DOMAIN_COEFF) {
    st = launch_xform(ring, kTwistedFwd, s_dev, s_ntt, cols);
    if (st) return st;
  } else {
    s_ntt = const_cast<void*>(s_dev);
  }
  hipEvent_t e1 = prof_mark(ctx);
  int rb = rows % 16 == 0 ? 16 : rows % 8 == 0 ? 8 : rows % 4 == 0 ? 4 : rows % 2 == 0 ? 2 : 1;
  switch (rb) {
    case 16: st = mac_launch<16>(key, s_ntt, cols, partial, nparts, cpb); break;
    case 8: st = mac_launch<8>(key, s_ntt, cols, partial, nparts, cpb); break;
    case 4: st = mac_launch<4>(key, s_ntt, cols, partial, nparts, cpb); break;
    case 2: st = mac_launch<2>(key, s_ntt, cols, partial, nparts, cpb); break;
    default: st = mac_launch<1>(key, s_ntt, cols, partial, nparts, cpb); break;
  }
  if (st) return st;
  hipEvent_t e2 = prof_mark(ctx);
  if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_finalize<FQ32>), dim3(rows), dim3(1024), 0, ctx->stream,
                       FQ32{(uint32_t)ring->q, ring->qinv32}, (const uint32_t*)ring->dev_tab,
                       (const uint32_t*)partial, nparts, rows, (int)kTabUntwistMac,
                       (uint32_t*)t_dev);
  else
    hipLaunchKernelGGL((k_finalize<FGold>), dim3(rows), dim3(1024), 0, ctx->stream, FGold{},
                       (const uint64_t*)ring->dev_tab, (const uint64_t*)partial, nparts, rows,
                       (int)kTabUntwist, (uint64_t*)t_dev);
  GH_HIP(hipGetLastError());
  hipEvent_t e3 = prof_mark(ctx);
  prof_span(ctx, 0, e0, e1);
  prof_span(ctx, 1, e1, e2);
  prof_span(ctx, 2, e2, e3);
  return GH_OK;
}

}  // namespace gh
