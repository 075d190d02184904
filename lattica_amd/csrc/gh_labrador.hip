// LaBRADOR hot calls on the GPU: Gram matrix g_ij = <s_i, s_j> and the
// ternary JL projection p = sum_part Pi_part * s_part.
//
// Gram (labrador/src/main_protocol/garbage.rs:76-93): every s_i is transformed
// once, then a SYRK-like sweep over the NTT domain: one wave per (4 x 4 tile of
// (i, j) pairs, column chunk), lane = slot, exact 96-bit accumulation for
// q < 2^32; per-chunk partials are summed and inverse-transformed by
// k_finalize (gh_kernels.hip). The reference's coefficient-domain dot product
// does one NTT multiply per term; the result is the same ring element.
//
// JL (labrador/src/jl.rs:388-508): column col of part p is ChaCha20 keystream
// bytes [col*pk, (col+1)*pk), pk = ceil(rows/4), key = seed, nonce = LE u32(p)
// || 0^8, counter from 0 (RFC 8439 block function below). A block of threads
// generates the keystream of 256 columns into LDS (one column per thread),
// then thread m accumulates row m over those columns. For q < 2^32 the
// accumulation is exact in int64 over centred lifts (|sum| <= cols * q/2 <
// 2^63 for cols < 2^32) and p mod q is derived from it; for Goldilocks it is
// modular.
#include "gh_internal.h"
#include "gh_ntt.h"
#include "gh_common_kernels.h"
#include "gh_chacha.h"

namespace gh {

// ---------------------------------------------------------------------------
// Gram
// ---------------------------------------------------------------------------

constexpr int kGramTile = 4;

__device__ __forceinline__ uint32_t pair_index(uint32_t r, uint32_t i, uint32_t j) {
  // garbage_index (main_protocol/mod.rs:664-670): row i starts at r*i - i*(i-1)/2
  return r * i - (i * (i - 1)) / 2 + (j - i);
}

// tile t of the upper triangle of ntile x ntile tiles -> (a, b), a <= b
__device__ __forceinline__ void tile_coords(uint32_t t, uint32_t ntile, uint32_t& a, uint32_t& b) {
  a = 0;
  uint32_t rem = t;
  while (rem >= ntile - a) { rem -= ntile - a; a++; }
  b = a + rem;
}

// one wave: pair tile `tile`, column chunk `ci` of the Gram matrix of s_ntt
__device__ __forceinline__ void gram_tile_q32(FQ32 f, const uint32_t* __restrict__ s_ntt,
                                              uint32_t r, uint32_t n, uint32_t chunk, uint32_t ci,
                                              uint32_t tile, uint32_t* __restrict__ partial) {
  const int lane = threadIdx.x;
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  uint32_t ta, tb;
  tile_coords(tile, ntile, ta, tb);
  const uint32_t c0 = ci * chunk;
  uint32_t c1 = c0 + chunk;
  if (c1 > n) c1 = n;
  uint64_t lo[kGramTile][kGramTile];
  uint32_t hi[kGramTile][kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) { lo[x][y] = 0; hi[x][y] = 0; }
  uint32_t ii[kGramTile], jj[kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++) {
    uint32_t i = ta * kGramTile + x, j = tb * kGramTile + x;
    ii[x] = i < r ? i : r - 1;  // clamp: out-of-range rows are computed but not stored
    jj[x] = j < r ? j : r - 1;
  }
  for (uint32_t c = c0; c < c1; c++) {
    uint32_t si[kGramTile], sj[kGramTile];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) si[x] = s_ntt[((size_t)ii[x] * n + c) * 64 + lane];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) sj[x] = s_ntt[((size_t)jj[x] * n + c) * 64 + lane];
    static_assert(kGramTile == 4, "four accumulators per mac96x4");
#pragma unroll
    for (int x = 0; x < kGramTile; x++)
      mac96x4(lo[x], hi[x], si[x], si[x], si[x], si[x], sj[0], sj[1], sj[2], sj[3]);
  }
  const uint32_t npairs = r * (r + 1) / 2;
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) {
      const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + y;
      if (i < r && j < r && i <= j) {
        uint32_t v = f.reduce96_r2((uint32_t)lo[x][y], (uint32_t)(lo[x][y] >> 32), hi[x][y]);
        partial[((size_t)ci * npairs + pair_index(r, i, j)) * 64 + lane] = v;
      }
    }
}

__device__ __forceinline__ void gram_tile_gold(FGold f, const uint64_t* __restrict__ s_ntt,
                                               uint32_t r, uint32_t n, uint32_t chunk, uint32_t ci,
                                               uint32_t tile, uint64_t* __restrict__ partial) {
  const int lane = threadIdx.x;
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  uint32_t ta, tb;
  tile_coords(tile, ntile, ta, tb);
  const uint32_t c0 = ci * chunk;
  uint32_t c1 = c0 + chunk;
  if (c1 > n) c1 = n;
  GoldAcc acc[kGramTile][kGramTile];  // exact lazy sums, reduced once per chunk
  uint32_t ii[kGramTile], jj[kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++) {
    uint32_t i = ta * kGramTile + x, j = tb * kGramTile + x;
    ii[x] = i < r ? i : r - 1;
    jj[x] = j < r ? j : r - 1;
  }
  for (uint32_t c = c0; c < c1; c++) {
    uint64_t si[kGramTile], sj[kGramTile];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) si[x] = s_ntt[((size_t)ii[x] * n + c) * 64 + lane];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) sj[x] = s_ntt[((size_t)jj[x] * n + c) * 64 + lane];
#pragma unroll
    for (int x = 0; x < kGramTile; x++)
#pragma unroll
      for (int y = 0; y < kGramTile; y++) gold_mac(acc[x][y], si[x], sj[y]);
  }
  const uint32_t npairs = r * (r + 1) / 2;
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) {
      const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + y;
      if (i < r && j < r && i <= j)
        partial[((size_t)ci * npairs + pair_index(r, i, j)) * 64 + lane] = gold_reduce(f, acc[x][y]);
    }
}

__global__ __launch_bounds__(64) void k_gram_q32(FQ32 f, const uint32_t* __restrict__ s_ntt,
                                                 uint32_t r, uint32_t n, uint32_t chunk,
                                                 uint32_t* __restrict__ partial) {
  gram_tile_q32(f, s_ntt, r, n, chunk, blockIdx.y, blockIdx.x, partial);
}

__global__ __launch_bounds__(64) void k_gram_gold(FGold f, const uint64_t* __restrict__ s_ntt,
                                                  uint32_t r, uint32_t n, uint32_t chunk,
                                                  uint64_t* __restrict__ partial) {
  gram_tile_gold(f, s_ntt, r, n, chunk, blockIdx.y, blockIdx.x, partial);
}

// Gram matrices of several witnesses in one launch: one wave per work entry
__global__ __launch_bounds__(64) void k_gram_batch_q32(FQ32 f, const GramDesc* __restrict__ desc,
                                                       const GramWork* __restrict__ work) {
  const GramWork w = work[blockIdx.x];
  const GramDesc d = desc[w.g];
  gram_tile_q32(f, reinterpret_cast<const uint32_t*>(d.s_ntt), d.r, d.n, d.chunk, w.chunk, w.tile,
                reinterpret_cast<uint32_t*>(d.partial));
}

__global__ __launch_bounds__(64) void k_gram_batch_gold(FGold f, const GramDesc* __restrict__ desc,
                                                        const GramWork* __restrict__ work) {
  const GramWork w = work[blockIdx.x];
  const GramDesc d = desc[w.g];
  gram_tile_gold(f, reinterpret_cast<const uint64_t*>(d.s_ntt), d.r, d.n, d.chunk, w.chunk, w.tile,
                 reinterpret_cast<uint64_t*>(d.partial));
}

// one 4-wave block per Gram entry (a batch has hundreds or thousands; 16-wave
// blocks would run one per CU at a time)
template <class F>
__global__ __launch_bounds__(256) void k_gram_fin_batch(F f, const typename F::W* __restrict__ tab,
                                                        const GramDesc* __restrict__ desc,
                                                        const GramFin* __restrict__ fin, int post) {
  using W = typename F::W;
  __shared__ W red[4 * 64];
  const GramFin e = fin[blockIdx.x];
  const GramDesc d = desc[e.g];
  finalize_row_g<F, 4>(f, tab, reinterpret_cast<const W*>(d.partial), d.nchunks,
                       d.r * (d.r + 1) / 2, e.pair, post, reinterpret_cast<W*>(d.dst), red);
}

uint32_t gram_tiles(uint32_t r) {
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  return ntile * (ntile + 1) / 2;
}

// column chunks: enough waves to cover the device a few times
void gram_chunking(int num_cus, uint32_t r, uint32_t n, uint32_t* chunk, uint32_t* nchunks) {
  const uint32_t tiles = gram_tiles(r);
  uint32_t want = (uint32_t)(4 * num_cus * 4 / (tiles ? tiles : 1));
  if (want < 1) want = 1;
  if (want > n) want = n;
  *chunk = (n + want - 1) / want;
  *nchunks = (n + *chunk - 1) / *chunk;
}

int enqueue_gram_batch(gh_ring* ring, hipStream_t stream, const GramDesc* desc,
                       const GramWork* work, uint32_t nwork, const GramFin* fin, uint32_t nfin) {
  if (nwork == 0) return GH_OK;
  if (ring->kind == GH_RING_Q32) {
    FQ32 f{(uint32_t)ring->q, ring->qinv32};
    hipLaunchKernelGGL(k_gram_batch_q32, dim3(nwork), dim3(64), 0, stream, f, desc, work);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_gram_fin_batch<FQ32>), dim3(nfin), dim3(256), 0, stream, f,
                       (const uint32_t*)ring->dev_tab, desc, fin, (int)kTabUntwistMac);
  } else {
    hipLaunchKernelGGL(k_gram_batch_gold, dim3(nwork), dim3(64), 0, stream, FGold{}, desc, work);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_gram_fin_batch<FGold>), dim3(nfin), dim3(256), 0, stream, FGold{},
                       (const uint64_t*)ring->dev_tab, desc, fin, (int)kTabUntwist);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_gram(gh_ring* ring, const void* s_dev, uint32_t r, uint32_t n, int domain,
                void* g_dev) {
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  const uint32_t tiles = ntile * (ntile + 1) / 2;
  const uint32_t npairs = r * (r + 1) / 2;
  uint32_t chunk, nchunks;
  gram_chunking(ctx->num_cus, r, n, &chunk, &nchunks);
  const size_t s_bytes = (size_t)r * n * 64 * wb;
  const size_t p_bytes = (size_t)nchunks * npairs * 64 * wb;
  void* ws;
  int st = ctx_workspace(ctx, s_bytes + p_bytes, &ws);
  if (st) return st;
  void* s_ntt = ws;
  void* partial = (char*)ws + s_bytes;
  if (domain == GH_DOMAIN_COEFF) {
    st = launch_xform(ring, kTwistedFwd, s_dev, s_ntt, (size_t)r * n);
    if (st) return st;
  } else {
    s_ntt = const_cast<void*>(s_dev);
  }
  dim3 grid(tiles, nchunks);
  if (ring->kind == GH_RING_Q32) {
    FQ32 f{(uint32_t)ring->q, ring->qinv32};
    hipLaunchKernelGGL(k_gram_q32, grid, dim3(64), 0, ctx->stream, f, (const uint32_t*)s_ntt, r,
                       n, chunk, (uint32_t*)partial);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_finalize4<FQ32>), dim3(npairs), dim3(256), 0, ctx->stream, f,
                       (const uint32_t*)ring->dev_tab, (const uint32_t*)partial, nchunks, npairs,
                       (int)kTabUntwistMac, (uint32_t*)g_dev);
  } else {
    hipLaunchKernelGGL(k_gram_gold, grid, dim3(64), 0, ctx->stream, FGold{},
                       (const uint64_t*)s_ntt, r, n, chunk, (uint64_t*)partial);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_finalize4<FGold>), dim3(npairs), dim3(256), 0, ctx->stream, FGold{},
                       (const uint64_t*)ring->dev_tab, (const uint64_t*)partial, nchunks, npairs,
                       (int)kTabUntwist, (uint64_t*)g_dev);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// ---------------------------------------------------------------------------
// LaBRADOR step around the commitment: balanced decomposition, garbage h,
// amortization
// ---------------------------------------------------------------------------

// decompose_coeff_balanced (labrador/src/main_protocol/mod.rs:258-289), one
// thread per coefficient; out[(poly * limbs + l) * 64 + k] as decompose_polys
// (mod.rs:731-745). |carry| <= q/2 and base <= 2^32 keep every step in int64;
// % and / truncate toward zero as the reference's i128 arithmetic does.
template <class W>
__global__ __launch_bounds__(256) void k_decompose(const W* __restrict__ in, size_t words,
                                                   uint64_t q, int64_t base, uint32_t limbs,
                                                   W* __restrict__ out) {
  const int64_t half = base / 2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t p = i >> 6, k = i & 63;
    const uint64_t c = (uint64_t)in[i];
    int64_t carry = c > q / 2 ? -(int64_t)(q - c) : (int64_t)c;
    for (uint32_t l = 0; l < limbs; l++) {
      int64_t r = carry % base;
      if (r > half) r -= base;
      else if (r < -half) r += base;
      out[(p * limbs + l) * 64 + k] = (W)(r >= 0 ? (uint64_t)r : q - (uint64_t)(-r));
      carry = (carry - r) / base;
    }
  }
}

// h_ij = (<phi_i, s_j> + <phi_j, s_i>) / 2 (garbage.rs:101-133): the Gram
// sweep's 4x4 pair tiles with two inputs, both NTT domain; the halving is
// folded into the inverse transform's table.
__device__ __forceinline__ void garbage_tile_q32(FQ32 f, const uint32_t* __restrict__ phi,
                                                 const uint32_t* __restrict__ s, uint32_t r,
                                                 uint32_t n, uint32_t chunk, uint32_t ci,
                                                 uint32_t tile, uint32_t* __restrict__ partial) {
  const int lane = threadIdx.x;
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  uint32_t ta, tb;
  tile_coords(tile, ntile, ta, tb);
  const uint32_t c0 = ci * chunk;
  uint32_t c1 = c0 + chunk;
  if (c1 > n) c1 = n;
  uint64_t lo[kGramTile][kGramTile];
  uint32_t hi[kGramTile][kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) { lo[x][y] = 0; hi[x][y] = 0; }
  uint32_t ii[kGramTile], jj[kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++) {
    const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + x;
    ii[x] = i < r ? i : r - 1;
    jj[x] = j < r ? j : r - 1;
  }
  for (uint32_t c = c0; c < c1; c++) {
    uint32_t pi[kGramTile], si[kGramTile], pj[kGramTile], sj[kGramTile];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) {
      pi[x] = phi[((size_t)ii[x] * n + c) * 64 + lane];
      si[x] = s[((size_t)ii[x] * n + c) * 64 + lane];
      pj[x] = phi[((size_t)jj[x] * n + c) * 64 + lane];
      sj[x] = s[((size_t)jj[x] * n + c) * 64 + lane];
    }
#pragma unroll
    for (int x = 0; x < kGramTile; x++) {
      mac96x4(lo[x], hi[x], pi[x], pi[x], pi[x], pi[x], sj[0], sj[1], sj[2], sj[3]);
      mac96x4(lo[x], hi[x], pj[0], pj[1], pj[2], pj[3], si[x], si[x], si[x], si[x]);
    }
  }
  const uint32_t npairs = r * (r + 1) / 2;
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) {
      const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + y;
      if (i < r && j < r && i <= j)
        partial[((size_t)ci * npairs + pair_index(r, i, j)) * 64 + lane] =
            f.reduce96_r2((uint32_t)lo[x][y], (uint32_t)(lo[x][y] >> 32), hi[x][y]);
    }
}

__device__ __forceinline__ void garbage_tile_gold(FGold f, const uint64_t* __restrict__ phi,
                                                  const uint64_t* __restrict__ s, uint32_t r,
                                                  uint32_t n, uint32_t chunk, uint32_t ci,
                                                  uint32_t tile, uint64_t* __restrict__ partial) {
  const int lane = threadIdx.x;
  const uint32_t ntile = (r + kGramTile - 1) / kGramTile;
  uint32_t ta, tb;
  tile_coords(tile, ntile, ta, tb);
  const uint32_t c0 = ci * chunk;
  uint32_t c1 = c0 + chunk;
  if (c1 > n) c1 = n;
  GoldAcc acc[kGramTile][kGramTile];
  uint32_t ii[kGramTile], jj[kGramTile];
#pragma unroll
  for (int x = 0; x < kGramTile; x++) {
    const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + x;
    ii[x] = i < r ? i : r - 1;
    jj[x] = j < r ? j : r - 1;
  }
  for (uint32_t c = c0; c < c1; c++) {
    uint64_t pi[kGramTile], si[kGramTile], pj[kGramTile], sj[kGramTile];
#pragma unroll
    for (int x = 0; x < kGramTile; x++) {
      pi[x] = phi[((size_t)ii[x] * n + c) * 64 + lane];
      si[x] = s[((size_t)ii[x] * n + c) * 64 + lane];
      pj[x] = phi[((size_t)jj[x] * n + c) * 64 + lane];
      sj[x] = s[((size_t)jj[x] * n + c) * 64 + lane];
    }
#pragma unroll
    for (int x = 0; x < kGramTile; x++)
#pragma unroll
      for (int y = 0; y < kGramTile; y++) {
        gold_mac(acc[x][y], pi[x], sj[y]);
        gold_mac(acc[x][y], pj[y], si[x]);
      }
  }
  const uint32_t npairs = r * (r + 1) / 2;
#pragma unroll
  for (int x = 0; x < kGramTile; x++)
#pragma unroll
    for (int y = 0; y < kGramTile; y++) {
      const uint32_t i = ta * kGramTile + x, j = tb * kGramTile + y;
      if (i < r && j < r && i <= j)
        partial[((size_t)ci * npairs + pair_index(r, i, j)) * 64 + lane] = gold_reduce(f, acc[x][y]);
    }
}

__global__ __launch_bounds__(64) void k_garbage_q32(FQ32 f, const uint32_t* __restrict__ phi,
                                                    const uint32_t* __restrict__ s, uint32_t r,
                                                    uint32_t n, uint32_t chunk,
                                                    uint32_t* __restrict__ partial) {
  garbage_tile_q32(f, phi, s, r, n, chunk, blockIdx.y, blockIdx.x, partial);
}

__global__ __launch_bounds__(64) void k_garbage_gold(FGold f, const uint64_t* __restrict__ phi,
                                                     const uint64_t* __restrict__ s, uint32_t r,
                                                     uint32_t n, uint32_t chunk,
                                                     uint64_t* __restrict__ partial) {
  garbage_tile_gold(f, phi, s, r, n, chunk, blockIdx.y, blockIdx.x, partial);
}

int launch_decompose(gh_ring* ring, const void* in, size_t polys, uint64_t base, uint32_t limbs,
                     void* out) {
  gh_ctx* ctx = ring->ctx;
  const size_t words = polys * 64;
  if (words == 0) return GH_OK;
  const size_t want = (words + 255) / 256;
  const unsigned grid = (unsigned)(want < (size_t)ctx->num_cus * 8 ? want : (size_t)ctx->num_cus * 8);
  if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_decompose<uint32_t>), dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t*)in, words, ring->q, (int64_t)base, limbs, (uint32_t*)out);
  else
    hipLaunchKernelGGL((k_decompose<uint64_t>), dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint64_t*)in, words, ring->q, (int64_t)base, limbs, (uint64_t*)out);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// phi, s: r x n polys each in `domain`; h_dev: r(r+1)/2 polys
int launch_garbage_h(gh_ring* ring, const void* phi, const void* s, uint32_t r, uint32_t n,
                     int domain, void* h_dev) {
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const uint32_t tiles = gram_tiles(r);
  const uint32_t npairs = r * (r + 1) / 2;
  uint32_t chunk, nchunks;
  gram_chunking(ctx->num_cus, r, n, &chunk, &nchunks);
  const size_t in_bytes = (size_t)r * n * 64 * wb;
  const size_t p_bytes = (size_t)nchunks * npairs * 64 * wb;
  void* ws;
  int st = ctx_workspace(ctx, (domain == GH_DOMAIN_COEFF ? 2 * in_bytes : 0) + p_bytes, &ws);
  if (st) return st;
  const void* pn = phi;
  const void* sn = s;
  void* partial = ws;
  if (domain == GH_DOMAIN_COEFF) {
    st = launch_xform(ring, kTwistedFwd, phi, ws, (size_t)r * n);
    if (st) return st;
    st = launch_xform(ring, kTwistedFwd, s, (char*)ws + in_bytes, (size_t)r * n);
    if (st) return st;
    pn = ws;
    sn = (char*)ws + in_bytes;
    partial = (char*)ws + 2 * in_bytes;
  }
  dim3 grid(tiles, nchunks);
  if (ring->kind == GH_RING_Q32) {
    FQ32 f{(uint32_t)ring->q, ring->qinv32};
    hipLaunchKernelGGL(k_garbage_q32, grid, dim3(64), 0, ctx->stream, f, (const uint32_t*)pn,
                       (const uint32_t*)sn, r, n, chunk, (uint32_t*)partial);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_finalize4<FQ32>), dim3(npairs), dim3(256), 0, ctx->stream, f,
                       (const uint32_t*)ring->dev_tab, (const uint32_t*)partial, nchunks, npairs,
                       (int)kTabUntwistMacHalf, (uint32_t*)h_dev);
  } else {
    hipLaunchKernelGGL(k_garbage_gold, grid, dim3(64), 0, ctx->stream, FGold{},
                       (const uint64_t*)pn, (const uint64_t*)sn, r, n, chunk, (uint64_t*)partial);
    GH_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_finalize4<FGold>), dim3(npairs), dim3(256), 0, ctx->stream, FGold{},
                       (const uint64_t*)ring->dev_tab, (const uint64_t*)partial, nchunks, npairs,
                       (int)kTabUntwistMacHalf, (uint64_t*)h_dev);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// z[k] = sum_i c_i * s_i[k]. The NTT-domain witness laid out [i][k][slot] is a
// commitment key's [column][row][slot]: z = "A" c with A[k][i] = s_i[k], one
// launch of the commitment path with c as the witness.
int launch_amortize(gh_ring* ring, const void* c, const void* s, uint32_t r, uint32_t n,
                    int domain, void* z_dev, void* s_scratch) {
  const void* sn = s;
  if (domain == GH_DOMAIN_COEFF) {
    int st = launch_xform(ring, kTwistedFwd, s, s_scratch, (size_t)r * n);
    if (st) return st;
    sn = s_scratch;
  }
  gh_key view;
  view.ring = ring;
  view.rows = n;
  view.msg_len = r;
  view.open_len = 0;
  view.dev_a = const_cast<void*>(sn);
  return launch_commit(&view, c, r, domain, z_dev);
}

// ---------------------------------------------------------------------------
// JL projection
// ---------------------------------------------------------------------------

constexpr int kJlCols = 256;  // columns per LDS sweep (one per thread)

// grid: (chunks, num_parts); block: 256 threads. Writes per-block row sums to
// part_acc[(part * chunks + chunk) * rows + m].
template <class W, bool GOLD>
__global__ __launch_bounds__(256) void k_jl(ChaKey key0, uint32_t rows, uint64_t q,
                                            const W* __restrict__ parts, uint64_t cols,
                                            uint64_t chunk_cols,
                                            int64_t* __restrict__ part_acc) {
  __shared__ uint8_t ks[kJlCols][64];
  __shared__ int64_t sval[kJlCols];
  __shared__ uint64_t sres[kJlCols];
  const int t = threadIdx.x;
  const uint32_t part = blockIdx.y;
  const uint32_t pk = (rows + 3) / 4;
  ChaKey key = key0;
  key.n0 = part;
  const W* sp = parts + (size_t)part * cols;
  const uint64_t c_begin = (uint64_t)blockIdx.x * chunk_cols;
  uint64_t c_end = c_begin + chunk_cols;
  if (c_end > cols) c_end = cols;
  int64_t acc = 0;
  uint64_t accm = 0;  // Goldilocks: modular
  for (uint64_t c0 = c_begin; c0 < c_end; c0 += kJlCols) {
    const uint64_t col = c0 + t;
    if (col < c_end) {
      // keystream bytes [col*pk, col*pk + pk): at most two 64-byte blocks
      const uint64_t off = col * pk;
      const uint32_t b0 = (uint32_t)(off >> 6);
      const uint32_t o0 = (uint32_t)(off & 63);
      uint32_t blk[16];
      chacha_block(key, b0, blk);
      const uint8_t* bytes = reinterpret_cast<const uint8_t*>(blk);
      uint32_t first = 64 - o0 < pk ? 64 - o0 : pk;
      for (uint32_t i = 0; i < first; i++) ks[t][i] = bytes[o0 + i];
      if (first < pk) {
        chacha_block(key, b0 + 1, blk);
        for (uint32_t i = first; i < pk; i++) ks[t][i] = bytes[i - first];
      }
      const W v = sp[col];
      sres[t] = (uint64_t)v;
      sval[t] = (uint64_t)v > q / 2 ? (int64_t)((uint64_t)v - q) : (int64_t)(uint64_t)v;
    }
    __syncthreads();
    const uint64_t ncol = c_end - c0 < (uint64_t)kJlCols ? c_end - c0 : (uint64_t)kJlCols;
    if ((uint32_t)t < rows) {
      const int byte_i = t >> 2, sh = 2 * (t & 3);
      for (uint64_t c = 0; c < ncol; c++) {
        const int code = (ks[c][byte_i] >> sh) & 3;
        if (GOLD) {
          const uint64_t v = sres[c];
          if (code == 0) {
            uint64_t s = accm + v;
            accm = s < accm ? s + kGoldEps : (s >= kGoldilocks ? s - kGoldilocks : s);
          } else if (code == 1) {
            accm = accm >= v ? accm - v : accm - v - kGoldEps;
          }
        } else {
          const int64_t v = sval[c];
          acc += code == 0 ? v : (code == 1 ? -v : 0);
        }
      }
    }
    __syncthreads();
  }
  if ((uint32_t)t < rows)
    part_acc[((size_t)part * gridDim.x + blockIdx.x) * rows + t] = GOLD ? (int64_t)accm : acc;
}

// p[m] = sum over blocks. Goldilocks: modular. q32: each block partial is the
// exact int64 sum of its columns (below 2^26 * 2^31 in magnitude); p mod q
// sums them reduced mod q, and the exact p (requested only below 2^26
// polynomials, |p| < 2^63) sums them in int64. One block per row: 256
// threads sum strided partials, then a tree in LDS.
template <bool GOLD>
__global__ __launch_bounds__(256) void k_jl_reduce(const int64_t* __restrict__ part_acc,
                                                   uint32_t nblk, uint32_t rows, uint64_t q,
                                                   int64_t* __restrict__ p_exact,
                                                   void* __restrict__ p_modq) {
  __shared__ int64_t red[256];
  __shared__ uint64_t redm[256];
  const uint32_t m = blockIdx.x, t = threadIdx.x;
  int64_t a = 0;
  uint64_t am = 0;  // q32: the sum mod q
  for (uint32_t b = t; b < nblk; b += 256) {
    const int64_t v = part_acc[(size_t)b * rows + m];
    if (GOLD) {
      const uint64_t x = (uint64_t)a, y = (uint64_t)v, s = x + y;
      a = (int64_t)(s < x ? s + kGoldEps : (s >= kGoldilocks ? s - kGoldilocks : s));
    } else {
      a += v;
      int64_t r = v % (int64_t)q;
      if (r < 0) r += (int64_t)q;
      am = (am + (uint64_t)r) % q;
    }
  }
  red[t] = a;
  redm[t] = am;
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (t < h) {
      if (GOLD) {
        const uint64_t x = (uint64_t)red[t], y = (uint64_t)red[t + h], s = x + y;
        red[t] = (int64_t)(s < x ? s + kGoldEps : (s >= kGoldilocks ? s - kGoldilocks : s));
      } else {
        red[t] += red[t + h];
        redm[t] = (redm[t] + redm[t + h]) % q;
      }
    }
    __syncthreads();
  }
  if (t != 0) return;
  if (GOLD) {
    ((uint64_t*)p_modq)[m] = (uint64_t)red[0];
  } else {
    if (p_exact) p_exact[m] = red[0];
    ((uint32_t*)p_modq)[m] = (uint32_t)redm[0];
  }
}

int launch_jl(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
              uint32_t num_parts, uint64_t cols, void* p_modq, int64_t* p_exact, int8_t* flat) {
  gh_ctx* ctx = ring->ctx;
  // one keystream block per column (rows 253..256, the reference profile's 256
  // among them): the matrix-core projection; otherwise the row-per-thread kernel
  if ((rows + 3) / 4 == 64) {
    int st = launch_jl_proj(ring, seed, rows, parts, num_parts, cols / ring->d, p_modq, p_exact);
    if (st) return st;
    if (flat) return launch_jl_flat(ring, seed, rows, num_parts, cols, flat);
    return GH_OK;
  }
  const ChaKey key = chacha_key(seed);
  // chunks per part so that the device holds a few blocks per CU
  uint64_t want = (uint64_t)ctx->num_cus * 4 / (num_parts ? num_parts : 1);
  if (want < 1) want = 1;
  uint64_t max_chunks = (cols + kJlCols - 1) / kJlCols;
  if (want > max_chunks) want = max_chunks;
  uint64_t chunk_cols = (cols + want - 1) / want;
  chunk_cols = (chunk_cols + kJlCols - 1) / kJlCols * kJlCols;
  uint32_t chunks = (uint32_t)((cols + chunk_cols - 1) / chunk_cols);
  const uint32_t nblk = chunks * num_parts;
  void* ws;
  int st = ctx_workspace(ctx, (size_t)nblk * rows * sizeof(int64_t), &ws);
  if (st) return st;
  dim3 grid(chunks, num_parts);
  const bool gold = ring->kind == GH_RING_GOLDILOCKS;
  if (!gold)
    hipLaunchKernelGGL((k_jl<uint32_t, false>), grid, dim3(256), 0, ctx->stream, key, rows,
                       ring->q, (const uint32_t*)parts, cols, chunk_cols, (int64_t*)ws);
  else
    hipLaunchKernelGGL((k_jl<uint64_t, true>), grid, dim3(256), 0, ctx->stream, key, rows,
                       ring->q, (const uint64_t*)parts, cols, chunk_cols, (int64_t*)ws);
  GH_HIP(hipGetLastError());
  if (!gold)
    hipLaunchKernelGGL((k_jl_reduce<false>), dim3(rows), dim3(256), 0, ctx->stream,
                       (const int64_t*)ws, nblk, rows, ring->q, p_exact, p_modq);
  else
    hipLaunchKernelGGL((k_jl_reduce<true>), dim3(rows), dim3(256), 0, ctx->stream,
                       (const int64_t*)ws, nblk, rows, ring->q, p_exact, p_modq);
  GH_HIP(hipGetLastError());
  // the entries in the flat layout, written row-contiguous by their own kernel
  if (flat) return launch_jl_flat(ring, seed, rows, num_parts, cols, flat);
  return GH_OK;
}

}  // namespace gh
