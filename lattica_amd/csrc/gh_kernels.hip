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
#include <hip/hip_ext.h>

#include <algorithm>

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

// Twisted forward NTT, one polynomial per quad of lanes (gh_ntt.h quad_fwd64,
// or quad_fwd64_lds for LX). Loads are stride-4 words (the quad together reads
// whole 64-byte segments); each lane stores its 16 output slots with 16-byte
// stores (LX: slots 16*k1 + 4*t .. +3).
template <class F, bool LX>
__global__ __launch_bounds__(256) void k_twisted_fwd_quad(F f, const typename F::W* __restrict__ tab,
                                                          const typename F::W* src,
                                                          typename F::W* dst, size_t count) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  constexpr int V = 16 / sizeof(W);
  __shared__ W qlds[LX ? 64 * kQuadStride<W> : 1];
  const int t = threadIdx.x & 3;
  W ltw[16][TW];
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int w = 0; w < TW; w++) ltw[k][w] = tab[(kTabQuad + 16 * t + k) * TW + w];
  const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 2;
  for (size_t p = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2; p < count; p += stride) {
    W x[16];
    const W* sp = src + p * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
    if constexpr (LX) {
      quad_fwd64_lds(f, x, tab, ltw, t, qlds + (threadIdx.x >> 2) * kQuadStride<W>);
      quad_store_k1(dst + p * 64, x, t);
    } else {
      quad_fwd64(f, x, tab, ltw, t);
      uint4* dp = reinterpret_cast<uint4*>(dst + p * 64 + 16 * bitrev2(t));
#pragma unroll
      for (int v = 0; v < 16 / V; v++) {
        uint4 o;
        W* ow = reinterpret_cast<W*>(&o);
#pragma unroll
        for (int c = 0; c < V; c++) ow[c] = x[v * V + c];
        dp[v] = o;
      }
    }
  }
}

// Negacyclic product, one pair per thread. The first operand's transform
// waits in the thread's own LDS row while the second is transformed (the two
// trade places word by word), so only one operand is ever in registers: about
// 170 VGPRs for q < 2^32 against 270 with both held, which lets two waves
// share each SIMD.
// Twisted inverse NTT, one polynomial per quad of lanes (gh_ntt.h
// quad_inv64): lane t loads the 16 contiguous slots 16*bitrev2(t) + k2 with
// 16-byte loads (LX, quad_inv64_lds: slots 16*k1 + 4*t .. +3) and stores
// coefficients t + 4*i2 (stride-4 words).
template <class F, bool LX>
__global__ __launch_bounds__(256) void k_twisted_inv_quad(F f, const typename F::W* __restrict__ tab,
                                                          const typename F::W* src,
                                                          typename F::W* dst, size_t count) {
  using W = typename F::W;
  constexpr int V = 16 / sizeof(W);
  __shared__ W qlds[LX ? 64 * kQuadStride<W> : 1];
  const int t = threadIdx.x & 3;
  // this lane's psi^(-t(2k2+1)), loaded once (LX)
  constexpr int TW = F::kTwWords;
  W itw[16][TW];
  if constexpr (LX) {
#pragma unroll
    for (int k = 0; k < 16; k++)
#pragma unroll
      for (int w = 0; w < TW; w++) itw[k][w] = tab[(kTabQuadInv + 16 * t + k) * TW + w];
  }
  const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 2;
  for (size_t p = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2; p < count; p += stride) {
    W x[16];
    if constexpr (LX) {
      quad_load_k1(src + p * 64, x, t);
      quad_inv64_lds(f, x, tab, kTabUntwist, t, qlds + (threadIdx.x >> 2) * kQuadStride<W>, itw);
      W* dp = dst + p * 64 + t;
#pragma unroll
      for (int i = 0; i < 16; i++) dp[4 * i] = x[i];
      continue;
    }
    const uint4* sp = reinterpret_cast<const uint4*>(src + p * 64 + 16 * bitrev2(t));
#pragma unroll
    for (int v = 0; v < 16 / V; v++) {
      const uint4 o = sp[v];
      const W* ow = reinterpret_cast<const W*>(&o);
#pragma unroll
      for (int c = 0; c < V; c++) x[v * V + c] = ow[c];
    }
    quad_inv64(f, x, tab, kTabUntwist, t);
    W* dp = dst + p * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) dp[4 * i] = x[i];
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
  W* row = lds + t * kStride;
  for (size_t tile = blockIdx.x; tile * TILE < count; tile += gridDim.x) {
    const size_t p0 = tile * TILE;
    const size_t np = count - p0 < (size_t)TILE ? count - p0 : (size_t)TILE;
    const bool active = (size_t)t < np;
    W x[64];
    tile_load<W, TILE>(lds, a + p0 * 64, np * 64);
    __syncthreads();
    if (active) {
#pragma unroll
      for (int k = 0; k < 64; k++) x[k] = row[k];
      fwd64<true>(f, x, tab);
    }
    __syncthreads();
    tile_load<W, TILE>(lds, b + p0 * 64, np * 64);
    __syncthreads();
    if (active) {
      W y[64];
#pragma unroll
      for (int k = 0; k < 64; k++) {
        y[k] = row[k];
        row[k] = x[k];
      }
      fwd64<true>(f, y, tab);
#pragma unroll
      for (int k = 0; k < 64; k++) y[k] = f.mont(y[k], row[k]);
      inv64(f, y, tab, kTabUntwistMul);
      lds_write_poly(lds, t, y);
    }
    __syncthreads();
    tile_store<W, TILE>(out + p0 * 64, lds, np * 64);
    __syncthreads();
  }
}

// Negacyclic product, one pair per quad of lanes (gh_ntt.h quad_fwd64 and
// quad_inv64): no LDS and 16 coefficients of each operand per lane, so more
// waves fit on a SIMD than with a whole polynomial per thread. Loads and
// stores are stride-4 words; a quad together covers whole 64-byte segments.
template <class F, bool LX>
__global__ __launch_bounds__(256) void k_negacyclic_quad(F f, const typename F::W* __restrict__ tab,
                                                         const typename F::W* a,
                                                         const typename F::W* b,
                                                         typename F::W* out, size_t count) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  __shared__ W qlds[LX ? 64 * kQuadStride<W> : 1];
  const int t = threadIdx.x & 3;
  W ltw[16][TW];
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int w = 0; w < TW; w++) ltw[k][w] = tab[(kTabQuad + 16 * t + k) * TW + w];
  const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 2;
  for (size_t p = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2; p < count; p += stride) {
    W x[16], y[16];
    const W* ap = a + p * 64 + t;
    const W* bp = b + p * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = ap[4 * i];
#pragma unroll
    for (int i = 0; i < 16; i++) y[i] = bp[4 * i];
    if constexpr (LX) {
      W* sc = qlds + (threadIdx.x >> 2) * kQuadStride<W>;
      quad_fwd64_lds(f, x, tab, ltw, t, sc);
      quad_fwd64_lds(f, y, tab, ltw, t, sc);
#pragma unroll
      for (int k = 0; k < 16; k++) x[k] = f.mont(x[k], y[k]);
      quad_inv64_lds(f, x, tab, kTabUntwistMul, t, sc);
    } else {
      quad_fwd64(f, x, tab, ltw, t);
      quad_fwd64(f, y, tab, ltw, t);
#pragma unroll
      for (int k = 0; k < 16; k++) x[k] = f.mont(x[k], y[k]);
      quad_inv64(f, x, tab, kTabUntwistMul, t);
    }
    W* op = out + p * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) op[4 * i] = x[i];
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
    mac96_rows(lo, hi, va, s0);
    mac96_rows(lo, hi, vb, s1);
  }
  if (c < c_end) {
    const uint32_t* a0 = A + (size_t)c * col_stride + (size_t)row0 * 64 + lane;
    uint32_t s0 = s_ntt[(size_t)c * 64 + lane];
    uint32_t va[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) va[r] = __builtin_nontemporal_load(a0 + r * 64);
    mac96_rows(lo, hi, va, s0);
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

// 16-byte variant: one wave load covers four rows of a column (1 KB), lane l
// holds rows 4i + l/16 and slots 4(l%16)..+3, so a 16-row column is four
// global_load_dwordx4 plus one for the witness slots. RB is a multiple of 4.
template <int RB, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_mac4_q32(
    FQ32 f, const uint32_t* __restrict__ A, const uint32_t* __restrict__ s_ntt,
    uint32_t* __restrict__ partial, uint32_t rows, uint32_t cols, uint32_t cols_per_block) {
  constexpr int NI = RB / 4;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rsub = lane >> 4, k0 = 4 * (lane & 15);
  const uint32_t row0 = blockIdx.y * RB;
  const uint32_t c_begin = blockIdx.x * cols_per_block;
  uint32_t c_end = c_begin + cols_per_block;
  if (c_end > cols) c_end = cols;
  uint64_t lo[NI][4];
  uint32_t hi[NI][4];
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) { lo[i][j] = 0; hi[i][j] = 0; }
  const size_t col_stride = (size_t)rows * 64;
  uint32_t c = c_begin + wave;
  for (; c + WAVES < c_end; c += 2 * WAVES) {
    const u32x4* a0 = reinterpret_cast<const u32x4*>(A + (size_t)c * col_stride + (size_t)row0 * 64) + lane;
    const u32x4* a1 = a0 + (size_t)WAVES * (col_stride / 4);
    const u32x4 s0 = *reinterpret_cast<const u32x4*>(s_ntt + (size_t)c * 64 + k0);
    const u32x4 s1 = *reinterpret_cast<const u32x4*>(s_ntt + (size_t)(c + WAVES) * 64 + k0);
    u32x4 va[NI], vb[NI];
#pragma unroll
    for (int i = 0; i < NI; i++) va[i] = __builtin_nontemporal_load(a0 + i * 64);
#pragma unroll
    for (int i = 0; i < NI; i++) vb[i] = __builtin_nontemporal_load(a1 + i * 64);
#pragma unroll
    for (int i = 0; i < NI; i++) {
      mac96x4(lo[i], hi[i], va[i].x, va[i].y, va[i].z, va[i].w, s0.x,
              s0.y, s0.z, s0.w);
    }
#pragma unroll
    for (int i = 0; i < NI; i++) {
      mac96x4(lo[i], hi[i], vb[i].x, vb[i].y, vb[i].z, vb[i].w, s1.x,
              s1.y, s1.z, s1.w);
    }
  }
  if (c < c_end) {
    const u32x4* a0 = reinterpret_cast<const u32x4*>(A + (size_t)c * col_stride + (size_t)row0 * 64) + lane;
    const u32x4 s0 = *reinterpret_cast<const u32x4*>(s_ntt + (size_t)c * 64 + k0);
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const u32x4 v = __builtin_nontemporal_load(a0 + i * 64);
      mac96x4(lo[i], hi[i], v.x, v.y, v.z, v.w, s0.x,
              s0.y, s0.z, s0.w);
    }
  }
  __shared__ uint32_t red[WAVES][RB][64];
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
      red[wave][4 * i + rsub][k0 + j] =
          f.reduce96_r2((uint32_t)lo[i][j], (uint32_t)(lo[i][j] >> 32), hi[i][j]);
  __syncthreads();
  for (int o = threadIdx.x; o < RB * 64; o += 64 * WAVES) {
    const int r = o >> 6, k = o & 63;
    uint32_t acc = red[0][r][k];
#pragma unroll
    for (int w = 1; w < WAVES; w++) acc = f.add(acc, red[w][r][k]);
    partial[((size_t)blockIdx.x * rows + row0 + r) * 64 + k] = acc;
  }
}

// Fused witness transform + MAC (q32). A block owns a contiguous range of at
// most kFuseChunk key columns (the launcher sizes the grid so). In issue order:
// the lane twiddles and the first witness polynomial of each quad (which the
// transform waits for), then the key loads of each wave's first PF columns,
// then the transform of the witness polynomials into LDS (four lanes per
// polynomial) while those key loads are in flight, then the MAC over the
// wave's columns against the LDS copy with PF columns kept in flight: a
// column's registers are reloaded with the column PF further on as soon as
// they are used. PF = max(1, DEPTH * 8 / WAVES) columns per wave: with two
// 4-wave blocks per CU, depth d keeps d * 64 KB of key in flight per CU for
// the whole sweep (128 KB, depth 2, measures fastest at kappa 16).
//
// Vector loads complete in issue order and the wait counter is one number, so
// the code between the prefetch and the MAC is kept free of branches that
// load: the first polynomial's transform is straight-line, loads past the
// block's end are clamped to its last column (and multiplied by a zero
// witness), and with NW = columns per wave known at compile time the MAC is
// fully unrolled, letting each wait name exactly the column it needs. NW = 0
// is the general loop (blocks of any width, conservative waits). Results
// equal k_mac4_q32 on a separately transformed witness.
constexpr int kFuseChunk = 128;
constexpr int kFuseStride = 68;  // LDS row stride of s_ntt (16-byte aligned rows)

template <int NI>
__device__ __forceinline__ void mac_column(uint64_t (&lo)[NI][4], uint32_t (&hi)[NI][4],
                                           const u32x4 (&a)[NI], const u32x4 s0) {
#pragma unroll
  for (int i = 0; i < NI; i++) {
    mac96x4(lo[i], hi[i], a[i].x, a[i].y, a[i].z, a[i].w, s0.x,
            s0.y, s0.z, s0.w);
  }
}

// Held-back finalizes (FinStage1/FinStage2): the first s2.nblk blocks finish
// the commitment before the previous one (one wave per row), and every MAC
// block runs one stage-1 task of the previous commitment (its loads issued
// with the witness loads, summed after the transform). Partial sums are
// written [row][part][slot].
template <int RB, int WAVES, int DEPTH, int NW>
__global__ __launch_bounds__(64 * WAVES) void k_commit_fused_q32(
    FQ32 f, const uint32_t* __restrict__ tab, const uint32_t* __restrict__ A,
    const uint32_t* __restrict__ s_coeff, uint32_t* __restrict__ partial, uint32_t rows,
    uint32_t cols, uint32_t cols_per_block, FinStage1 s1, FinStage2 s2) {
  constexpr int NI = RB / 4;
  constexpr int PF = DEPTH * 8 / WAVES > 0 ? DEPTH * 8 / WAVES : 1;
  // cols_per_block * kFuseStride witness words, then WAVES * 64 stage-1 words;
  // WAVES * RB * 64 words for the block reduction
  extern __shared__ uint32_t lds[];
  const uint32_t lead = s2.nblk + s1.extra;
  const uint32_t nparts = gridDim.x - lead;
  if (blockIdx.x < lead) {
    if (blockIdx.x < s2.nblk) {
      // stage 2: the group sums were complete when this launch began
      const uint32_t row = blockIdx.x * WAVES + (threadIdx.x >> 6);
      if (blockIdx.y == 0 && row < s2.rows)
        sum_rows_inv<FQ32, 1>(f, tab, (int)kTabUntwistMac, (const uint32_t*)s2.src, s2.n, row,
                              0, (uint32_t*)s2.out, nullptr);
    } else {
      stage1_only(f, s1, s1.mac_tasks + (blockIdx.y * s1.extra + (blockIdx.x - s2.nblk)) * s1.tpb,
                  lds, WAVES);
    }
    return;
  }
  const uint32_t bx = blockIdx.x - lead;
  const uint32_t bid = blockIdx.y * nparts + bx;  // stage-1 task
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rsub = lane >> 4, k0 = 4 * (lane & 15);
  const int t = threadIdx.x & 3;
  const uint32_t row0 = blockIdx.y * RB;
  const uint32_t ch = bx * cols_per_block;  // < cols: every block owns a column
  const uint32_t ncol = ch + cols_per_block < cols ? cols_per_block : cols - ch;
  const size_t col_stride = (size_t)rows * 64;
  const u32x4* akey = reinterpret_cast<const u32x4*>(A + (size_t)row0 * 64) + lane;
  const size_t col4 = col_stride / 4;  // one key column, in u32x4
  // column j of this wave is wave + j * WAVES; loads clamp to the block's last
  const uint32_t nwm = NW > 0 ? (uint32_t)NW : (ncol + WAVES - 1) / WAVES;
  auto col_of = [&](uint32_t j) -> uint32_t {
    const uint32_t c = wave + j * WAVES;
    return c < ncol ? c : ncol - 1;
  };

  uint64_t lo[NI][4];
  uint32_t hi[NI][4];
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) { lo[i][j] = 0; hi[i][j] = 0; }

  // (a) loads the transform waits for: lane twiddles, this thread's first
  // witness polynomial (clamped into the block)
  uint32_t ltw[16][2];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    ltw[k][0] = tab[(kTabQuad + 16 * t + k) * 2];
    ltw[k][1] = tab[(kTabQuad + 16 * t + k) * 2 + 1];
  }
  const uint32_t p0 = threadIdx.x >> 2;
  uint32_t x[16];
  {
    const uint32_t pc = p0 < ncol ? p0 : ncol - 1;
    const uint32_t* sp = s_coeff + (size_t)(ch + pc) * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
  }
  uint32_t l1v[4];
  stage1_load(s1, bid, s1.mac_tasks, l1v);
  // (b) key prefetch: this wave's first PF columns
  u32x4 pa[PF][NI];
#pragma unroll
  for (int q = 0; q < PF; q++) {
    if (NW > 0 ? q < NW : (uint32_t)q < nwm) {
      const uint32_t cq = col_of(q);
#pragma unroll
      for (int i = 0; i < NI; i++)
        pa[q][i] = __builtin_nontemporal_load(akey + (size_t)(ch + cq) * col4 + i * 64);
    }
  }
  // (c) witness transform into LDS while (b) is in flight: the first
  // polynomial straight-line, further ones (blocks wider than 16 * WAVES) after
  quad_fwd64(f, x, tab, ltw, t);
  if (p0 < ncol) {
    u32x4* dp = reinterpret_cast<u32x4*>(lds + p0 * kFuseStride + 16 * bitrev2(t));
#pragma unroll
    for (int v = 0; v < 4; v++) dp[v] = u32x4{x[4 * v], x[4 * v + 1], x[4 * v + 2], x[4 * v + 3]};
  }
  for (uint32_t p = p0 + 16 * WAVES; p < ncol; p += 16 * WAVES) {
    const uint32_t* sp = s_coeff + (size_t)(ch + p) * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
    quad_fwd64(f, x, tab, ltw, t);
    u32x4* dp = reinterpret_cast<u32x4*>(lds + p * kFuseStride + 16 * bitrev2(t));
#pragma unroll
    for (int v = 0; v < 4; v++) dp[v] = u32x4{x[4 * v], x[4 * v + 1], x[4 * v + 2], x[4 * v + 3]};
  }
  uint32_t* l1lds = lds + cols_per_block * kFuseStride;
  stage1_wave_sum(f, l1v, l1lds);
  __syncthreads();
  stage1_store(f, s1, bid, s1.mac_tasks, l1lds, WAVES);
  // (d) MAC over the wave's columns, PF in flight; a clamped column meets a
  // zero witness
  auto witness = [&](uint32_t j) -> u32x4 {
    const uint32_t c = wave + j * WAVES;
    const u32x4 s0 = *reinterpret_cast<const u32x4*>(lds + col_of(j) * kFuseStride + k0);
    const uint32_t keep = c < ncol ? 0xffffffffu : 0u;
    return u32x4{s0.x & keep, s0.y & keep, s0.z & keep, s0.w & keep};
  };
  if constexpr (NW > 0) {
#pragma unroll
    for (int j = 0; j < NW; j++) {
      constexpr int kPF = PF;
      const int q = j % kPF;
      mac_column<NI>(lo, hi, pa[q], witness(j));
      if (j + kPF < NW) {
        const uint32_t cn = col_of(j + kPF);
#pragma unroll
        for (int i = 0; i < NI; i++)
          pa[q][i] = __builtin_nontemporal_load(akey + (size_t)(ch + cn) * col4 + i * 64);
      }
    }
  } else {
    for (uint32_t j0 = 0; j0 < nwm; j0 += PF) {
#pragma unroll
      for (int q = 0; q < PF; q++) {
        const uint32_t j = j0 + q;
        if (j < nwm) {
          mac_column<NI>(lo, hi, pa[q], witness(j));
          if (j + PF < nwm) {
            const uint32_t cn = col_of(j + PF);
#pragma unroll
            for (int i = 0; i < NI; i++)
              pa[q][i] = __builtin_nontemporal_load(akey + (size_t)(ch + cn) * col4 + i * 64);
          }
        }
      }
    }
  }
  // (e) block reduction through the same LDS
  __syncthreads();
  uint32_t* red = lds;  // [WAVES][RB][64]
#pragma unroll
  for (int i = 0; i < NI; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
      red[(wave * RB + 4 * i + rsub) * 64 + k0 + j] =
          f.reduce96_r2((uint32_t)lo[i][j], (uint32_t)(lo[i][j] >> 32), hi[i][j]);
  __syncthreads();
  for (int o = threadIdx.x; o < RB * 64; o += 64 * WAVES) {
    uint32_t acc = red[o];
#pragma unroll
    for (int w = 1; w < WAVES; w++) acc = f.add(acc, red[w * RB * 64 + o]);
    partial[((size_t)(row0 + o / 64) * nparts + bx) * 64 + o % 64] = acc;
  }
}

// Fused witness transform + MAC (q32), row-split form for 16-row blocks of
// four waves: wave w owns rows 4w..4w+3 of the block for every column the
// block owns (lane l accumulates row 4w + l / 16, slots 4 (l % 16) .. +3). A
// wave's sums are complete when its sweep ends, so there is no cross-wave
// reduction through LDS and no barrier after the sweep: each lane reduces its
// four 96-bit sums and stores them as one 16-byte word. Per column a wave
// loads its four rows (1 KB, contiguous in the [col][row][slot] key) and
// reads the column's transformed witness from LDS; PF = 8 * DEPTH columns are
// kept in flight per wave. The transform, the held-back stages and the
// clamped prefetch are those of k_commit_fused_q32, as are the results.
// NC > 0: every block owns exactly NC columns (fully unrolled sweep). LX: the
// transform's 4-point DFT over the quad goes through the polynomial's LDS area
// (quad_fwd64_lds) instead of lane exchanges.
template <int DEPTH, int NC, bool LX>
__global__ __launch_bounds__(256) void k_commit_rows_q32(
    FQ32 f, const uint32_t* __restrict__ tab, const uint32_t* __restrict__ A,
    const uint32_t* __restrict__ s_coeff, uint32_t* __restrict__ partial, uint32_t rows,
    uint32_t cols, uint32_t cols_per_block, FinStage1 s1, FinStage2 s2) {
  constexpr int WAVES = 4;
  constexpr int PF = 8 * DEPTH;
  // cols_per_block * kFuseStride witness words, then WAVES * 64 stage-1 words
  extern __shared__ uint32_t lds[];
  const uint32_t lead = s2.nblk + s1.extra;
  const uint32_t nparts = gridDim.x - lead;
  if (blockIdx.x < lead) {
    if (blockIdx.x < s2.nblk) {
      const uint32_t row = blockIdx.x * WAVES + (threadIdx.x >> 6);
      if (blockIdx.y == 0 && row < s2.rows)
        sum_rows_inv<FQ32, 1>(f, tab, (int)kTabUntwistMac, (const uint32_t*)s2.src, s2.n, row,
                              0, (uint32_t*)s2.out, nullptr);
    } else {
      stage1_only(f, s1, s1.mac_tasks + (blockIdx.y * s1.extra + (blockIdx.x - s2.nblk)) * s1.tpb,
                  lds, WAVES);
    }
    return;
  }
  const uint32_t bx = blockIdx.x - lead;
  const uint32_t bid = blockIdx.y * nparts + bx;  // stage-1 task
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int k0 = 4 * (lane & 15);
  const int t = threadIdx.x & 3;
  const uint32_t row = blockIdx.y * 16 + 4 * wave + (lane >> 4);
  const uint32_t ch = bx * cols_per_block;  // < cols: every block owns a column
  const uint32_t ncol =
      NC > 0 ? (uint32_t)NC : (ch + cols_per_block < cols ? cols_per_block : cols - ch);
  // u32x4 of (column c, this lane): c * rows * 16 + (16 * blockIdx.y + 4 * wave) * 16 + lane
  const u32x4* akey = reinterpret_cast<const u32x4*>(A + (size_t)blockIdx.y * 16 * 64) + 64 * wave + lane;
  const size_t col4 = (size_t)rows * 16;  // one key column, in u32x4
  auto col_of = [&](uint32_t c) -> uint32_t { return c < ncol ? c : ncol - 1; };

  uint64_t lo[4] = {0, 0, 0, 0};
  uint32_t hi[4] = {0, 0, 0, 0};
  // (a) lane twiddles and this thread's first witness polynomial
  uint32_t ltw[16][2];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    ltw[k][0] = tab[(kTabQuad + 16 * t + k) * 2];
    ltw[k][1] = tab[(kTabQuad + 16 * t + k) * 2 + 1];
  }
  const uint32_t p0 = threadIdx.x >> 2;
  uint32_t x[16];
  {
    const uint32_t* sp = s_coeff + (size_t)(ch + col_of(p0)) * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
  }
  uint32_t l1v[4];
  stage1_load(s1, bid, s1.mac_tasks, l1v);
  // (b) key prefetch: the first PF columns
  u32x4 pa[PF];
#pragma unroll
  for (int q = 0; q < PF; q++) pa[q] = __builtin_nontemporal_load(akey + (size_t)(ch + col_of(q)) * col4);
  // (c) witness transform into LDS while (b) is in flight (a quad's lanes
  // share p, so a clamped quad skips as a whole)
  auto transform_store = [&](uint32_t p) {
    uint32_t* sc = lds + p * kFuseStride;
    if constexpr (LX) {
      quad_fwd64_lds(f, x, tab, ltw, t, sc);
#pragma unroll
      for (int k1 = 0; k1 < 4; k1++)
        *reinterpret_cast<u32x4*>(sc + 16 * k1 + 4 * t) =
            u32x4{x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]};
    } else {
      quad_fwd64(f, x, tab, ltw, t);
      u32x4* dp = reinterpret_cast<u32x4*>(sc + 16 * bitrev2(t));
#pragma unroll
      for (int v = 0; v < 4; v++) dp[v] = u32x4{x[4 * v], x[4 * v + 1], x[4 * v + 2], x[4 * v + 3]};
    }
  };
  if (p0 < ncol) transform_store(p0);
  for (uint32_t p = p0 + 16 * WAVES; p < ncol; p += 16 * WAVES) {
    const uint32_t* sp = s_coeff + (size_t)(ch + p) * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
    transform_store(p);
  }
  uint32_t* l1lds = lds + cols_per_block * kFuseStride;
  stage1_wave_sum(f, l1v, l1lds);
  __syncthreads();
  stage1_store(f, s1, bid, s1.mac_tasks, l1lds, WAVES);
  // (d) sweep: every column of the block, PF in flight
  auto mac = [&](const u32x4& a, uint32_t c) {
    const u32x4 s0 = *reinterpret_cast<const u32x4*>(lds + c * kFuseStride + k0);
    mac96x4(lo, hi, a.x, a.y, a.z, a.w, s0.x, s0.y,
            s0.z, s0.w);
  };
  if constexpr (NC > 0) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      constexpr int kPF = PF;
      const int q = c % kPF;
      mac(pa[q], (uint32_t)c);
      if (c + kPF < NC) pa[q] = __builtin_nontemporal_load(akey + (size_t)(ch + c + kPF) * col4);
    }
  } else {
    for (uint32_t c0 = 0; c0 < ncol; c0 += PF) {
#pragma unroll
      for (int q = 0; q < PF; q++) {
        const uint32_t c = c0 + q;
        if (c < ncol) {
          mac(pa[q], c);
          if (c + PF < ncol) pa[q] = __builtin_nontemporal_load(akey + (size_t)(ch + c + PF) * col4);
        }
      }
    }
  }
  // (e) this lane's four sums, reduced and stored: [row][part][slot]
  u32x4 out;
  out.x = f.reduce96_r2((uint32_t)lo[0], (uint32_t)(lo[0] >> 32), hi[0]);
  out.y = f.reduce96_r2((uint32_t)lo[1], (uint32_t)(lo[1] >> 32), hi[1]);
  out.z = f.reduce96_r2((uint32_t)lo[2], (uint32_t)(lo[2] >> 32), hi[2]);
  out.w = f.reduce96_r2((uint32_t)lo[3], (uint32_t)(lo[3] >> 32), hi[3]);
  *reinterpret_cast<u32x4*>(partial + ((size_t)row * nparts + bx) * 64 + k0) = out;
}

// Fused witness transform + MAC (Goldilocks), the per-call form: a block
// transforms its contiguous range of at most kFuseChunkGold witness
// polynomials into LDS (four lanes per polynomial), then each wave sums its
// columns' (wave, wave + WAVES, ...) products against that copy exactly and
// lazily (GoldAcc, one reduction per row at the end), two columns of key loads
// in flight, and the waves' sums meet in the same LDS. The first fin.nblk blocks run the previous commitment's
// finalize instead. Results equal k_mac_gold on a separately transformed
// witness.
constexpr int kFuseChunkGold = 64;
constexpr int kFuseStrideGold = 66;  // LDS row stride of the transformed witness, words

template <int RB, int WAVES, bool LX>
__global__ __launch_bounds__(64 * WAVES) void k_commit_fused_gold(
    FGold f, const uint64_t* __restrict__ tab, const uint64_t* __restrict__ A,
    const uint64_t* __restrict__ s_coeff, uint64_t* __restrict__ partial, uint32_t rows,
    uint32_t cols, uint32_t cols_per_block, FinStage1 s1, FinStage2 s2) {
  // cols_per_block * kFuseStrideGold witness words, then WAVES * 64 stage-1
  // words; WAVES * RB * 64 words for the block reduction
  extern __shared__ uint64_t lds64[];
  const uint32_t lead = s2.nblk + s1.extra;
  const uint32_t nparts = gridDim.x - lead;
  if (blockIdx.x < lead) {
    if (blockIdx.x < s2.nblk) {
      const uint32_t row = blockIdx.x * WAVES + (threadIdx.x >> 6);
      if (blockIdx.y == 0 && row < s2.rows)
        sum_rows_inv<FGold, 1>(f, tab, (int)kTabUntwist, (const uint64_t*)s2.src, s2.n, row, 0,
                               (uint64_t*)s2.out, nullptr);
    } else {
      stage1_only(f, s1, s1.mac_tasks + (blockIdx.y * s1.extra + (blockIdx.x - s2.nblk)) * s1.tpb,
                  lds64, WAVES);
    }
    return;
  }
  const uint32_t bx = blockIdx.x - lead;
  const uint32_t bid = blockIdx.y * nparts + bx;  // stage-1 task
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x & 3;
  const uint32_t row0 = blockIdx.y * RB;
  const uint32_t ch = bx * cols_per_block;  // < cols: every block owns a column
  const uint32_t ncol = ch + cols_per_block < cols ? cols_per_block : cols - ch;
  uint64_t ltw[16][1];
#pragma unroll
  for (int k = 0; k < 16; k++) ltw[k][0] = tab[kTabQuad + 16 * t + k];
  uint64_t l1v[4];
  stage1_load(s1, bid, s1.mac_tasks, l1v);
  // whole quads share p, so they drop out together (the transform exchanges
  // values within a quad)
  for (uint32_t p = threadIdx.x >> 2; p < ncol; p += 16 * WAVES) {
    uint64_t x[16];
    const uint64_t* sp = s_coeff + (size_t)(ch + p) * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
    uint64_t* sc = lds64 + p * kFuseStrideGold;
    if constexpr (LX) {
      quad_fwd64_lds(f, x, tab, ltw, t, sc);
      quad_store_k1(sc, x, t);
    } else {
      quad_fwd64(f, x, tab, ltw, t);
      uint64_t* dp = sc + 16 * bitrev2(t);
#pragma unroll
      for (int v = 0; v < 16; v++) dp[v] = x[v];
    }
  }
  uint64_t* l1lds = lds64 + cols_per_block * kFuseStrideGold;
  stage1_wave_sum(f, l1v, l1lds);
  __syncthreads();
  stage1_store(f, s1, bid, s1.mac_tasks, l1lds, WAVES);
  GoldAcc acc[RB];  // exact lazy sums, reduced once per row at the end
  const size_t col_stride = (size_t)rows * 64;
  const uint64_t* akey = A + (size_t)row0 * 64 + lane;
  // this wave's columns wave + j * WAVES, j < nw; the next column's key loads
  // are issued before the current column's products (clamped to the block's
  // last column, so the loads never branch)
  const uint32_t nw = (uint32_t)wave < ncol ? (ncol - wave + WAVES - 1) / WAVES : 0;
  auto load_col = [&](uint64_t (&v)[RB], uint32_t j) {
    const uint32_t c = wave + j * WAVES;
    const uint64_t* a0 = akey + (size_t)(ch + (c < ncol ? c : ncol - 1)) * col_stride;
#pragma unroll
    for (int r = 0; r < RB; r++) v[r] = __builtin_nontemporal_load(a0 + r * 64);
  };
  auto mac_col = [&](const uint64_t (&v)[RB], uint32_t j) {
    const uint64_t s0 = lds64[(wave + j * WAVES) * kFuseStrideGold + lane];
#pragma unroll
    for (int r = 0; r < RB; r++) gold_mac(acc[r], v[r], s0);
  };
  uint64_t va[RB], vb[RB];
  if (nw > 0) load_col(va, 0);
  for (uint32_t j = 0; j < nw; j += 2) {
    load_col(vb, j + 1);
    mac_col(va, j);
    if (j + 1 >= nw) break;
    load_col(va, j + 2);
    mac_col(vb, j + 1);
  }
  __syncthreads();  // the witness copy is no longer read
  uint64_t* red = lds64;  // [WAVES][RB][64]
#pragma unroll
  for (int r = 0; r < RB; r++) red[(wave * RB + r) * 64 + lane] = gold_reduce(f, acc[r]);
  __syncthreads();
  for (int o = threadIdx.x; o < RB * 64; o += 64 * WAVES) {
    uint64_t a = red[o];
#pragma unroll
    for (int w = 1; w < WAVES; w++) a = f.add(a, red[w * RB * 64 + o]);
    partial[((size_t)(row0 + o / 64) * nparts + bx) * 64 + o % 64] = a;
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

// Bandwidth probes: a streaming read (XOR-folded so the loads stay live) and a
// streaming copy of known size. They give the achievable HBM ceiling beside the
// nominal peak, and a reference dispatch for scaling hardware traffic
// counters. Every block reads (copies) one contiguous span of the buffer with
// nontemporal 16-byte loads, four in flight per thread: per-block contiguous
// spans stream faster on MI355X than a grid-wide interleave, whose waves each
// touch a new region every iteration.
__global__ __launch_bounds__(512) void k_stream_read(const u32x4* __restrict__ src, size_t n,
                                                     size_t span, uint32_t* __restrict__ sink) {
  u32x4 acc = {0, 0, 0, 0};
  const size_t b0 = (size_t)blockIdx.x * span;
  const size_t b1 = b0 + span < n ? b0 + span : n;
  const size_t st = blockDim.x;
  size_t i = b0 + threadIdx.x;
  for (; i + 3 * st < b1; i += 4 * st) {
    const u32x4 a = __builtin_nontemporal_load(src + i);
    const u32x4 b = __builtin_nontemporal_load(src + i + st);
    const u32x4 c = __builtin_nontemporal_load(src + i + 2 * st);
    const u32x4 d = __builtin_nontemporal_load(src + i + 3 * st);
    acc ^= a ^ b ^ c ^ d;
  }
  for (; i < b1; i += st) acc ^= __builtin_nontemporal_load(src + i);
  const uint32_t v = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (v == 0x9e3779b9u) sink[blockIdx.x] = v;  // data dependent, practically never taken
}

__global__ __launch_bounds__(512) void k_stream_copy(const u32x4* __restrict__ src, size_t n,
                                                     size_t span, u32x4* __restrict__ dst) {
  const size_t b0 = (size_t)blockIdx.x * span;
  const size_t b1 = b0 + span < n ? b0 + span : n;
  for (size_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Issue-rate probes (gh_measure_issue): each wave runs kIssueIters loop
// iterations of independent work and writes one value per thread (so the
// work stays live) to out[blockIdx.x * 256 + threadIdx.x]:
//   0: four independent v_mfma_i32_32x32x32_i8 per iteration;
//   1: eight independent v_mfma_i32_16x16x64_i8 per iteration (the same
//      int8 multiply-accumulates per iteration as 0);
//   2: four independent ChaCha20 quarter rounds (48 VALU instructions);
//   3: the MFMAs of 0 and the quarter rounds of 2 interleaved in one loop.
constexpr int kIssueIters = 4096;
typedef int pv4i __attribute__((ext_vector_type(4)));
typedef int pv16i __attribute__((ext_vector_type(16)));
#define GH_PQR(a, b, c, d)                                                                   \
  a += b; d ^= a; d = __builtin_amdgcn_alignbit(d, d, 16); c += d; b ^= c;                   \
  b = __builtin_amdgcn_alignbit(b, b, 20); a += b; d ^= a; d = __builtin_amdgcn_alignbit(d, d, 24); \
  c += d; b ^= c; b = __builtin_amdgcn_alignbit(b, b, 25);

template <int KIND>
__global__ __launch_bounds__(256) void k_issue_probe(int* __restrict__ out, int seed) {
  const pv4i a = {seed + (int)threadIdx.x, 1, 2, 3}, b = {3, 2, 1, seed};
  pv16i c0{}, c1{}, c2{}, c3{};
  pv4i d[8];
#pragma unroll
  for (int j = 0; j < 8; j++) d[j] = pv4i{};
  uint32_t x0 = seed + threadIdx.x, x1 = 1, x2 = 2, x3 = 3, x4 = 4, x5 = 5, x6 = 6, x7 = 7;
  uint32_t x8 = 8, x9 = 9, x10 = 10, x11 = 11, x12 = 12, x13 = 13, x14 = 14, x15 = 15;
  for (int i = 0; i < kIssueIters; i++) {
    if constexpr (KIND == 0 || KIND == 3) c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
    if constexpr (KIND >= 2) { GH_PQR(x0, x4, x8, x12) }
    if constexpr (KIND == 0 || KIND == 3) c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
    if constexpr (KIND >= 2) { GH_PQR(x1, x5, x9, x13) }
    if constexpr (KIND == 0 || KIND == 3) c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c2, 0, 0, 0);
    if constexpr (KIND >= 2) { GH_PQR(x2, x6, x10, x14) }
    if constexpr (KIND == 0 || KIND == 3) c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c3, 0, 0, 0);
    if constexpr (KIND >= 2) { GH_PQR(x3, x7, x11, x15) }
    if constexpr (KIND == 1) {
#pragma unroll
      for (int j = 0; j < 8; j++) d[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, d[j], 0, 0, 0);
    }
  }
  int s = (int)(x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7 ^ x8 ^ x9 ^ x10 ^ x11 ^ x12 ^ x13 ^ x14 ^ x15);
#pragma unroll
  for (int r = 0; r < 16; r++) s += c0[r] + c1[r] + c2[r] + c3[r];
#pragma unroll
  for (int j = 0; j < 8; j++) s += d[j][0] + d[j][1] + d[j][2] + d[j][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
#undef GH_PQR

// kind 0..3 on `blocks` blocks of 256 threads; out holds blocks * 256 ints
int launch_issue_probe(gh_ctx* ctx, int kind, uint32_t blocks, int* out, hipEvent_t e0,
                       hipEvent_t e1) {
  const dim3 grid(blocks), block(256);
  if (kind == 0)
    hipExtLaunchKernelGGL((k_issue_probe<0>), grid, block, 0, ctx->stream, e0, e1, 0, out, 1);
  else if (kind == 1)
    hipExtLaunchKernelGGL((k_issue_probe<1>), grid, block, 0, ctx->stream, e0, e1, 0, out, 1);
  else if (kind == 2)
    hipExtLaunchKernelGGL((k_issue_probe<2>), grid, block, 0, ctx->stream, e0, e1, 0, out, 1);
  else
    hipExtLaunchKernelGGL((k_issue_probe<3>), grid, block, 0, ctx->stream, e0, e1, 0, out, 1);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// Issue rate of single VALU instruction kinds (gh_measure_valu): eight
// independent chains per lane, one instruction per chain and iteration,
// written as asm so that the compiler neither fuses nor removes them; no
// memory operation inside the loop. Each thread writes one word to
// out[blockIdx.x * 256 + threadIdx.x].
//   0 v_add_u32, 1 v_mul_hi_u32, 2 v_mul_lo_u32, 3 v_mad_u64_u32 (64-bit
//   accumulate, carry to an SGPR pair), 4 v_lshl_add_u64 (64-bit add),
//   5 v_fma_f64, 6 FQ32::sub_fix (v_sub_co_u32 with the borrow in an SGPR
//   pair, v_cndmask_b32 on it, v_add_u32: three instructions per step, with
//   the hazard waits the compiler puts between them), 7 v_cndmask_b32 (lane
//   mask in an SGPR pair written before the loop), 8 v_xor_b32,
//   9 v_alignbit_b32 (32-bit rotate).
constexpr int kValuIters = 2048;
#define GH_CHAINS8(ST) ST(0) ST(1) ST(2) ST(3) ST(4) ST(5) ST(6) ST(7)
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t* __restrict__ out, uint32_t s) {
  uint32_t res = 0;
  if constexpr (KIND <= 2 || KIND >= 6) {
    uint32_t a[8];
    const FQ32 fq{s | 0x80000000u, 1};
    // lane mask for kind 7: every other lane
    const uint64_t mask = 0x5555555555555555ull ^ (uint64_t)s;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = threadIdx.x + j;
    for (int i = 0; i < kValuIters; i++) {
      if constexpr (KIND == 0) {
#define ST(j) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[j]) : "v"(s));
        GH_CHAINS8(ST)
#undef ST
      } else if constexpr (KIND == 1) {
#define ST(j) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a[j]) : "v"(s));
        GH_CHAINS8(ST)
#undef ST
      } else if constexpr (KIND == 2) {
#define ST(j) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[j]) : "v"(s));
        GH_CHAINS8(ST)
#undef ST
      } else if constexpr (KIND == 6) {
        // compiler-generated, so that its hazard waits are the ones the
        // transforms get; q = s | 2^31 keeps every step's result below 2^32
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = fq.sub_fix(a[j], s);
      } else if constexpr (KIND == 7) {
#define ST(j) asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(a[j]) : "v"(s), "s"(mask));
        GH_CHAINS8(ST)
#undef ST
      } else if constexpr (KIND == 8) {
#define ST(j) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[j]) : "v"(s));
        GH_CHAINS8(ST)
#undef ST
      } else {
#define ST(j) asm volatile("v_alignbit_b32 %0, %0, %0, 25" : "+v"(a[j]));
        GH_CHAINS8(ST)
#undef ST
      }
    }

#pragma unroll
    for (int j = 0; j < 8; j++) res ^= a[j];
  } else if constexpr (KIND <= 4) {
    uint64_t a[8], cy[8];
    const uint64_t b = s;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = threadIdx.x + j;
    for (int i = 0; i < kValuIters; i++) {
      if constexpr (KIND == 3) {
#define ST(j) asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a[j]), "=s"(cy[j]) : "v"(s), "v"(s));
        GH_CHAINS8(ST)
#undef ST
      } else {
#define ST(j) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(a[j]) : "v"(b));
        GH_CHAINS8(ST)
#undef ST
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) res ^= (uint32_t)a[j] ^ (uint32_t)(a[j] >> 32);
  } else {
    double a[8];
    const double b = 1.0000001, c = 0.5;
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = threadIdx.x + j;
    for (int i = 0; i < kValuIters; i++) {
#define ST(j) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(a[j]) : "v"(b), "v"(c));
      GH_CHAINS8(ST)
#undef ST
    }
    double r = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) r += a[j];
    res = (uint32_t)r;
  }
  out[blockIdx.x * 256 + threadIdx.x] = res;
}
#undef GH_CHAINS8

// kind 0..kValuKinds-1 on `blocks` blocks of 256 threads; out holds blocks * 256 words
int launch_valu_probe(gh_ctx* ctx, int kind, uint32_t blocks, uint32_t* out, hipEvent_t e0,
                      hipEvent_t e1) {
  const dim3 grid(blocks), block(256);
  const uint32_t s = 12345u;
  switch (kind) {
    case 0: hipExtLaunchKernelGGL((k_valu_probe<0>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 1: hipExtLaunchKernelGGL((k_valu_probe<1>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 2: hipExtLaunchKernelGGL((k_valu_probe<2>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 3: hipExtLaunchKernelGGL((k_valu_probe<3>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 4: hipExtLaunchKernelGGL((k_valu_probe<4>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 5: hipExtLaunchKernelGGL((k_valu_probe<5>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 6: hipExtLaunchKernelGGL((k_valu_probe<6>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 7: hipExtLaunchKernelGGL((k_valu_probe<7>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 8: hipExtLaunchKernelGGL((k_valu_probe<8>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    case 9: hipExtLaunchKernelGGL((k_valu_probe<9>), grid, block, 0, ctx->stream, e0, e1, 0, out, s); break;
    default: return GH_E_INVALID_ARGUMENT;
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_stream_probe(gh_ctx* ctx, const void* src, void* dst, size_t bytes, int copy,
                        hipEvent_t e0, hipEvent_t e1) {
  const size_t n = bytes / 16;
  const unsigned grid = (unsigned)ctx->num_cus * (copy ? 2 : 1);
  const size_t span = (n + grid - 1) / grid;
  if (copy)
    hipExtLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(512), 0, ctx->stream, e0, e1, 0,
                          (const u32x4*)src, n, span, (u32x4*)dst);
  else
    hipExtLaunchKernelGGL(k_stream_read, dim3(grid), dim3(512), 0, ctx->stream, e0, e1, 0,
                          (const u32x4*)src, n, span, (uint32_t*)dst);
  GH_HIP(hipGetLastError());
  return GH_OK;
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
  // four lanes per polynomial at every size by default (for Goldilocks a
  // whole polynomial per thread needs 267 VGPRs, one wave per SIMD)
  const int xq = ctx->tune.xform_quad;
  const bool quad_all = xq == 1 || (xq < 0 && sizeof(W) == 8);
  if (MODE == kTwistedInv && quad_all) {
    unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, 8);
    if (ctx->tune.quad_lds)
      hipLaunchKernelGGL((k_twisted_inv_quad<F, true>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)src, (W*)dst, count);
    else
      hipLaunchKernelGGL((k_twisted_inv_quad<F, false>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)src, (W*)dst, count);
  } else if (MODE == kTwistedFwd && (quad_all || count <= (size_t)ctx->tune.quad_max)) {
    // four lanes per polynomial: 64 polynomials per 256-thread block
    unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, 8);
    if (ctx->tune.quad_lds)
      hipLaunchKernelGGL((k_twisted_fwd_quad<F, true>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)src, (W*)dst, count);
    else
      hipLaunchKernelGGL((k_twisted_fwd_quad<F, false>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)src, (W*)dst, count);
  } else if (count >= (size_t)BIG * ctx->num_cus * 2) {
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
  const int quad = ctx->tune.negacyclic_quad;
  if (quad == 1 || (quad < 0 && (sizeof(W) == 8 || ctx->tune.quad_lds))) {
    // 64 pairs per 256-thread block
    unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, 8);
    if (ctx->tune.quad_lds)
      hipLaunchKernelGGL((k_negacyclic_quad<F, true>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)a, (const W*)b, (W*)out, count);
    else
      hipLaunchKernelGGL((k_negacyclic_quad<F, false>), dim3(grid), dim3(256), 0, ctx->stream, f,
                         tab, (const W*)a, (const W*)b, (W*)out, count);
    GH_HIP(hipGetLastError());
    return GH_OK;
  }
  // blocks per CU: LDS-bound for q < 2^32 (9 x 16.6 KB), one wave per SIMD
  // for Goldilocks (480 VGPRs)
  constexpr int kPerCu = sizeof(W) == 4 ? 9 : 4;
  unsigned grid = grid_for((count + 63) / 64, ctx->num_cus, kPerCu);
  hipLaunchKernelGGL((k_negacyclic_mul<F, 64>), dim3(grid), dim3(64), 0, ctx->stream, f, tab,
                     (const W*)a, (const W*)b, (W*)out, count);
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
                      uint32_t nparts, uint32_t cpb, int waves) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  dim3 grid(nparts, key->rows / RB);
  if (ring->kind == GH_RING_Q32) {
    FQ32 f{(uint32_t)ring->q, ring->qinv32};
    const uint32_t* A = (const uint32_t*)key->dev_a;
    const uint32_t* S = (const uint32_t*)s_ntt;
    uint32_t* P = (uint32_t*)partial;
    if constexpr (RB % 4 == 0) {
      if (ctx->tune.mac_variant == 1) {
        if (waves == 4)
          hipLaunchKernelGGL((k_mac4_q32<RB, 4>), grid, dim3(256), 0, ctx->stream, f, A, S, P, key->rows, cols, cpb);
        else if (waves == 8)
          hipLaunchKernelGGL((k_mac4_q32<RB, 8>), grid, dim3(512), 0, ctx->stream, f, A, S, P, key->rows, cols, cpb);
        else
          hipLaunchKernelGGL((k_mac4_q32<RB, 16>), grid, dim3(1024), 0, ctx->stream, f, A, S, P, key->rows, cols, cpb);
        GH_HIP(hipGetLastError());
        return GH_OK;
      }
    }
    hipLaunchKernelGGL((k_mac_q32<RB>), grid, dim3(64 * kMacWaves), 0, ctx->stream, f, A, S, P,
                       key->rows, cols, cpb);
  } else {
    hipLaunchKernelGGL((k_mac_gold<RB>), grid, dim3(64 * kMacWaves), 0, ctx->stream, FGold{},
                       (const uint64_t*)key->dev_a, (const uint64_t*)s_ntt, (uint64_t*)partial,
                       key->rows, cols, cpb);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

template <int RB, int WAVES, int DEPTH>
static void fused_launch_nw(gh_ctx* ctx, dim3 grid, size_t lds_bytes, hipEvent_t e0,
                            hipEvent_t e1, FQ32 f, const uint32_t* tab, const uint32_t* A,
                            const uint32_t* S, uint32_t* P, uint32_t rows, uint32_t cols,
                            uint32_t cpb, const FinStage1& s1, const FinStage2& s2) {
  // the unrolled sweep when every full block has the same columns per wave
  // (instantiated for 16-row blocks, kappa a multiple of 16)
  const uint32_t nw = RB == 16 && cpb % WAVES == 0 ? cpb / WAVES : 0;
#define GH_FUSED_NW(N)                                                                          \
  hipExtLaunchKernelGGL((k_commit_fused_q32<RB, WAVES, DEPTH, N>), grid, dim3(64 * WAVES),      \
                        lds_bytes, ctx->stream, e0, e1, 0, f, tab, A, S, P, rows, cols, cpb,    \
                        s1, s2)
  if constexpr (RB == 16) {
    if (nw == 8) GH_FUSED_NW(8);
    else if (nw == 16) GH_FUSED_NW(16);
    else if (nw == 32) GH_FUSED_NW(32);
    else GH_FUSED_NW(0);
  } else {
    GH_FUSED_NW(0);
  }
#undef GH_FUSED_NW
}

template <int RB, int WAVES>
static void fused_launch(gh_ctx* ctx, dim3 grid, size_t lds_bytes, hipEvent_t e0, hipEvent_t e1,
                         FQ32 f, const uint32_t* tab, const uint32_t* A, const uint32_t* S,
                         uint32_t* P, uint32_t rows, uint32_t cols, uint32_t cpb,
                         const FinStage1& s1, const FinStage2& s2) {
  switch (ctx->tune.fused_depth) {
    case 1: fused_launch_nw<RB, WAVES, 1>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, s1, s2); break;
    case 3: fused_launch_nw<RB, WAVES, 3>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, s1, s2); break;
    case 4: fused_launch_nw<RB, WAVES, 4>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, s1, s2); break;
    default: fused_launch_nw<RB, WAVES, 2>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, s1, s2); break;
  }
}

template <int DEPTH, bool LX>
static void rows_launch_nc(gh_ctx* ctx, dim3 grid, size_t lds_bytes, hipEvent_t e0, hipEvent_t e1,
                           FQ32 f, const uint32_t* tab, const uint32_t* A, const uint32_t* S,
                           uint32_t* P, uint32_t rows, uint32_t cols, uint32_t cpb,
                           const FinStage1& s1, const FinStage2& s2) {
  // the unrolled sweep when every block owns 64 columns
  if (cpb == 64 && cols % 64 == 0)
    hipExtLaunchKernelGGL((k_commit_rows_q32<DEPTH, 64, LX>), grid, dim3(256), lds_bytes, ctx->stream,
                          e0, e1, 0, f, tab, A, S, P, rows, cols, cpb, s1, s2);
  else
    hipExtLaunchKernelGGL((k_commit_rows_q32<DEPTH, 0, LX>), grid, dim3(256), lds_bytes, ctx->stream,
                          e0, e1, 0, f, tab, A, S, P, rows, cols, cpb, s1, s2);
}

static void rows_launch(gh_ctx* ctx, dim3 grid, size_t lds_bytes, hipEvent_t e0, hipEvent_t e1,
                        FQ32 f, const uint32_t* tab, const uint32_t* A, const uint32_t* S,
                        uint32_t* P, uint32_t rows, uint32_t cols, uint32_t cpb,
                        const FinStage1& s1, const FinStage2& s2) {
#define GH_ROWS(D, L) rows_launch_nc<D, L>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, s1, s2)
  const bool lx = ctx->tune.quad_lds != 0;
  switch (ctx->tune.fused_depth) {
    case 1: if (lx) GH_ROWS(1, true); else GH_ROWS(1, false); break;
    case 3: if (lx) GH_ROWS(3, true); else GH_ROWS(3, false); break;
    case 4: if (lx) GH_ROWS(4, true); else GH_ROWS(4, false); break;
    default: if (lx) GH_ROWS(2, true); else GH_ROWS(2, false); break;
  }
#undef GH_ROWS
}

// Standalone finalize of up to two commitments of one ring whose sums are
// laid out [row][n][slot]: 16 waves per row for up to one row per CU, 4
// waves per row beyond (all rows in one round).
static int launch_finalize_rows(gh_ctx* ctx, const gh_ring* ring, const FinSeg& a,
                                const FinSeg& b, bool prof = true) {
  const uint32_t nrows = a.rows + b.rows;
  if (nrows == 0) return GH_OK;
  hipEvent_t f0 = nullptr, f1 = nullptr;
  if (prof) prof_pair(ctx, 2, &f0, &f1);
  const bool many = nrows > (uint32_t)ctx->num_cus;
  if (ring->kind == GH_RING_Q32) {
    const FQ32 f{(uint32_t)ring->q, ring->qinv32};
    const uint32_t* tab = (const uint32_t*)ring->dev_tab;
    if (many)
      hipExtLaunchKernelGGL((k_finalize_rows<FQ32, 4>), dim3(nrows), dim3(256), 0, ctx->stream,
                            f0, f1, 0, f, tab, (int)kTabUntwistMac, a, b);
    else
      hipExtLaunchKernelGGL((k_finalize_rows<FQ32, 16>), dim3(nrows), dim3(1024), 0,
                            ctx->stream, f0, f1, 0, f, tab, (int)kTabUntwistMac, a, b);
  } else {
    const uint64_t* tab = (const uint64_t*)ring->dev_tab;
    if (many)
      hipExtLaunchKernelGGL((k_finalize_rows<FGold, 4>), dim3(nrows), dim3(256), 0, ctx->stream,
                            f0, f1, 0, FGold{}, tab, (int)kTabUntwist, a, b);
    else
      hipExtLaunchKernelGGL((k_finalize_rows<FGold, 16>), dim3(nrows), dim3(1024), 0,
                            ctx->stream, f0, f1, 0, FGold{}, tab, (int)kTabUntwist, a, b);
  }
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int ctx_flush(gh_ctx* ctx) {
  PendingFin& pf = ctx->fin;
  if (!pf.any()) return GH_OK;
  const char* w = (const char*)ctx->work;
  FinSeg a{}, b{};
  if (pf.s2.active) a = FinSeg{w + pf.s2.off, pf.s2.t_out, pf.s2.n, pf.s2.rows};
  if (pf.s1.active) b = FinSeg{w + pf.s1.off, pf.s1.t_out, pf.s1.n, pf.s1.rows};
  pf.s1.active = pf.s2.active = false;
  return launch_finalize_rows(ctx, pf.ring, a, b);
}

static bool spans_overlap(const void* a, size_t na, const void* b, size_t nb) {
  const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
  return a0 < b0 + nb && b0 < a0 + na;
}

// The no-allocation form's partial sums (gh_commit_noalloc): one array per
// device, held until the process ends; host code uses the address that
// hipGetSymbolAddress reports for the current device, never the name.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

int scratch_address(void** out) {
  *out = nullptr;
  if (hipGetSymbolAddress(out, HIP_SYMBOL(scratch)) != hipSuccess || !*out) return GH_E_DEVICE;
  return GH_OK;
}

// Shape of a fused launch and the stage work it carries.
struct FusedPlan {
  size_t off = 0;        // this commitment's partial sums in the workspace
  FinStage1 s1{};
  FinStage2 s2{};
};

// Places a fused commitment's partial sums (pbytes; a grid of nparts MAC
// blocks per row block, nrb row blocks, `threads` threads each) and, when it
// defers, the held-back stages its launch runs; a stage-1 task list longer
// than the MAC blocks gets extra blocks. Pending finalizes are issued first
// when the launch cannot carry them: another ring, no deferral, a witness
// overlapping a pending t_out (inside one launch nothing orders the MAC
// blocks' witness reads after the stage-2 writes), or partial sums larger than
// the workspace's areas.
static int fused_plan(gh_ctx* ctx, const gh_ring* ring, const void* s_dev, size_t s_bytes,
                      size_t pbytes, uint32_t nparts, uint32_t nrb, uint32_t threads,
                      uint32_t waves, const void* dummy, bool defer, FusedPlan* plan) {
  PendingFin& pf = ctx->fin;
  const size_t wb = ring->word_bytes;
  const uint32_t per_pass = threads / 16;  // partials a block sums in one stage-1 task
  const size_t cap = (ctx->work_bytes / 4) & ~(size_t)255;
  uint32_t groups = 0;
  bool keep = defer && pf.ring == ring && pbytes <= cap;
  if (keep) {
    for (const PendingStage* st : {&pf.s1, &pf.s2})
      if (st->active && spans_overlap(s_dev, s_bytes, st->t_out, (size_t)st->rows * 64 * wb))
        keep = false;
  }
  if (!keep) {
    if (int st = ctx_flush(ctx)) return st;
    void* ws;
    // room for the four areas of a pipeline (or one area without deferral)
    if (int st = ctx_workspace(ctx, (defer ? 4 : 1) * pbytes, &ws)) return st;
    pf.flip = 0;
  }
  pf.ring = ring;
  const size_t area = defer ? (ctx->work_bytes / 4) & ~(size_t)255 : 0;
  char* w = (char*)ctx->work;
  plan->off = pf.flip * area;
  plan->s1 = FinStage1{dummy, nullptr, 0, 1, 0, 0, 0, 0, 0, 1};
  plan->s2 = FinStage2{nullptr, nullptr, 0, 0, 0};
  if (pf.s2.active)
    plan->s2 = FinStage2{w + pf.s2.off, pf.s2.t_out, pf.s2.n, pf.s2.rows,
                         (pf.s2.rows + waves - 1) / waves};
  if (pf.s1.active) {
    groups = (pf.s1.n + per_pass - 1) / per_pass;
    const uint32_t tasks = pf.s1.rows * groups;
    // stage1_tasks_per_block 0: one task in each MAC block, the rest in
    // stage-1-only blocks of one task; n > 0: every task in stage-1-only
    // blocks of n; -1: such blocks, about one per two CUs. (The kernels fit
    // three blocks per CU; about num_cus / 2 leading blocks leave the MAC
    // blocks spread evenly and measure fastest, while num_cus leading blocks
    // put three MAC blocks on some CUs.)
    const int tune_tpb = ctx->tune.stage1_tasks_per_block;
    const uint32_t half_cus = std::max(1, ctx->num_cus / 2);
    const uint32_t tpb =
        tune_tpb >= 0 ? (uint32_t)tune_tpb
                      : std::min(kStage1MaxTpb, std::max(1u, (tasks + half_cus - 1) / half_cus));
    const uint32_t mac_tasks = tpb ? 0 : std::min(tasks, nparts * nrb);
    const uint32_t eb = tpb ? tpb : 1;
    const uint32_t nextra = (tasks - mac_tasks + eb - 1) / eb;  // stage-1-only blocks
    const uint32_t extra = (nextra + nrb - 1) / nrb;
    const uint32_t per = (pf.s1.n + groups - 1) / groups;
    const size_t goff = (2 + (pf.flip ^ 1)) * area;  // the previous commitment's group sums
    plan->s1 = FinStage1{w + pf.s1.off, w + goff, pf.s1.n, pf.s1.rows, groups, per, tasks,
                         mac_tasks, extra, eb};
    pf.s2 = PendingStage{true, goff, groups, pf.s1.rows, pf.s1.t_out};
  } else {
    pf.s2.active = false;
  }
  pf.s1.active = false;
  return GH_OK;
}

// Plan of a fused launch whose partial sums go to `scratch` (the
// no-allocation form): no stage work, nothing allocated. A held-back finalize
// is issued first (its launch allocates nothing), so an earlier commitment's
// output is complete before this one reads or overwrites it.
static int scratch_plan(gh_ctx* ctx, size_t pbytes, const void* dummy, FusedPlan* plan) {
  if (pbytes > kScratchBytes) return GH_E_INVALID_ARGUMENT;
  if (int e = ctx_flush(ctx)) return e;
  plan->off = 0;
  plan->s1 = FinStage1{dummy, nullptr, 0, 1, 0, 0, 0, 0, 0, 1};
  plan->s2 = FinStage2{nullptr, nullptr, 0, 0, 0};
  return GH_OK;
}

// After a fused launch: with deferral its finalize is held back as stage s1,
// else it runs now (from `scratch` when that holds the partial sums).
static int fused_finish(gh_ctx* ctx, const gh_ring* ring, const FusedPlan& plan, uint32_t nparts,
                        uint32_t rows, void* t_dev, bool defer, void* scratch = nullptr) {
  PendingFin& pf = ctx->fin;
  if (scratch)
    return launch_finalize_rows(ctx, ring, FinSeg{scratch, t_dev, nparts, rows},
                                FinSeg{nullptr, nullptr, 0, 0}, false);
  if (!defer)
    return launch_finalize_rows(ctx, ring, FinSeg{(const char*)ctx->work + plan.off, t_dev, nparts,
                                                  rows},
                                FinSeg{nullptr, nullptr, 0, 0});
  pf.s1 = PendingStage{true, plan.off, nparts, rows, t_dev};
  pf.flip ^= 1;
  return GH_OK;
}

// q32 commitment: the fused transform+MAC kernel, then the finalize. With
// `defer` the finalize is held back and runs in two stages inside the next
// two deferred launches over the same ring (see PendingFin), so back-to-back
// commitments cost one launch each. (A second stream for the finalize was
// measured: the cross-queue event hand-off costs more than it overlaps.)
static int launch_commit_fused(gh_key* key, const void* s_dev, uint32_t cols, void* t_dev,
                               bool defer, void* scratch = nullptr) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  const uint32_t rows = key->rows;
  // fused 2: the row-split kernel (16-row blocks of four waves)
  const bool rowsplit = ctx->tune.fused == 2 && rows % 16 == 0;
  const int waves = rowsplit ? 4 : ctx->tune.mac_waves;
  const int rb = rows % 16 == 0 ? 16 : rows % 8 == 0 ? 8 : 4;
  // blocks: the tuned count per CU, at least one per kFuseChunk columns (a
  // block's witness fits its LDS), at most one per `waves` columns; every block
  // owns at least one column
  uint32_t nparts = (uint32_t)(ctx->num_cus * ctx->tune.mac_blocks_per_cu);
  if (scratch) nparts = std::min<uint32_t>(nparts, kScratchBytes / ((size_t)rows * 64 * 4));
  const uint32_t min_parts = (cols + kFuseChunk - 1) / kFuseChunk;
  if (nparts < min_parts) nparts = min_parts;
  const uint32_t max_parts = (cols + waves - 1) / waves;
  if (nparts > max_parts) nparts = max_parts;
  if (nparts == 0) nparts = 1;
  const uint32_t cpb = (cols + nparts - 1) / nparts;
  nparts = (cols + cpb - 1) / cpb;
  const size_t pbytes = ((size_t)nparts * rows * 64 * 4 + 255) & ~(size_t)255;

  FusedPlan plan;
  if (int st = scratch ? scratch_plan(ctx, pbytes, key->dev_a, &plan)
                       : fused_plan(ctx, ring, s_dev, (size_t)cols * 64 * 4, pbytes, nparts,
                                    rows / rb, 64 * waves, waves, key->dev_a, defer, &plan))
    return st;
  FQ32 f{(uint32_t)ring->q, ring->qinv32};
  const uint32_t* tab = (const uint32_t*)ring->dev_tab;
  const uint32_t* A = (const uint32_t*)key->dev_a;
  const uint32_t* S = (const uint32_t*)s_dev;
  uint32_t* P = scratch ? (uint32_t*)scratch : (uint32_t*)((char*)ctx->work + plan.off);
  dim3 grid(plan.s2.nblk + plan.s1.extra + nparts, rows / rb);
  size_t lds_words = (size_t)cpb * kFuseStride + (size_t)waves * 64;
  if (!rowsplit) lds_words = std::max(lds_words, (size_t)waves * rb * 64);
  lds_words = std::max(lds_words, (size_t)waves * 64 * plan.s1.tpb);
  const size_t lds_bytes = lds_words * 4;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!scratch) prof_pair(ctx, 1, &e0, &e1);
#define GH_FUSED(R, W) fused_launch<R, W>(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, plan.s1, plan.s2)
  if (rowsplit) rows_launch(ctx, grid, lds_bytes, e0, e1, f, tab, A, S, P, rows, cols, cpb, plan.s1, plan.s2);
  else if (rb == 16) { if (waves == 4) GH_FUSED(16, 4); else if (waves == 8) GH_FUSED(16, 8); else GH_FUSED(16, 16); }
  else if (rb == 8) { if (waves == 4) GH_FUSED(8, 4); else if (waves == 8) GH_FUSED(8, 8); else GH_FUSED(8, 16); }
  else { if (waves == 4) GH_FUSED(4, 4); else if (waves == 8) GH_FUSED(4, 8); else GH_FUSED(4, 16); }
#undef GH_FUSED
  GH_HIP(hipGetLastError());
  return fused_finish(ctx, ring, plan, nparts, rows, t_dev, defer, scratch);
}

// Goldilocks commitment, per-call form: k_commit_fused_gold, then the
// finalize, held back with `defer` exactly as for q32.
static int launch_commit_fused_gold(gh_key* key, const void* s_dev, uint32_t cols, void* t_dev,
                                    bool defer, void* scratch = nullptr) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  const uint32_t rows = key->rows;
  constexpr int kWaves = 4;
  // 8-row blocks at most: a block's lazy sums take 9 registers per row, and
  // two row blocks transforming the same witness columns cost less than
  // 16-row blocks at two waves per SIMD
  const int rb = rows % 8 == 0 ? 8 : 4;
  // blocks: the tuned count per CU over all row blocks, at least one per
  // kFuseChunkGold columns, at most one per wave's column; every block owns at
  // least one column
  uint32_t nparts = (uint32_t)(ctx->num_cus * ctx->tune.mac_blocks_per_cu) / (rows / rb);
  if (scratch) nparts = std::min<uint32_t>(nparts, kScratchBytes / ((size_t)rows * 64 * 8));
  if (nparts == 0) nparts = 1;
  const uint32_t min_parts = (cols + kFuseChunkGold - 1) / kFuseChunkGold;
  if (nparts < min_parts) nparts = min_parts;
  const uint32_t max_parts = (cols + kWaves - 1) / kWaves;
  if (nparts > max_parts) nparts = max_parts;
  if (nparts == 0) nparts = 1;
  const uint32_t cpb = (cols + nparts - 1) / nparts;
  nparts = (cols + cpb - 1) / cpb;
  const size_t pbytes = ((size_t)nparts * rows * 64 * 8 + 255) & ~(size_t)255;

  FusedPlan plan;
  if (int st = scratch ? scratch_plan(ctx, pbytes, key->dev_a, &plan)
                       : fused_plan(ctx, ring, s_dev, (size_t)cols * 64 * 8, pbytes, nparts,
                                    rows / rb, 64 * kWaves, kWaves, key->dev_a, defer, &plan))
    return st;
  const uint64_t* tab = (const uint64_t*)ring->dev_tab;
  const uint64_t* A = (const uint64_t*)key->dev_a;
  const uint64_t* S = (const uint64_t*)s_dev;
  uint64_t* P = scratch ? (uint64_t*)scratch : (uint64_t*)((char*)ctx->work + plan.off);
  dim3 grid(plan.s2.nblk + plan.s1.extra + nparts, rows / rb);
  size_t lds_words = (size_t)cpb * kFuseStrideGold + (size_t)kWaves * 64;
  lds_words = std::max(lds_words, (size_t)kWaves * rb * 64);
  lds_words = std::max(lds_words, (size_t)kWaves * 64 * plan.s1.tpb);
  const size_t lds_bytes = lds_words * 8;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!scratch) prof_pair(ctx, 1, &e0, &e1);
#define GH_FUSED_GOLD(R, L)                                                                     \
  hipExtLaunchKernelGGL((k_commit_fused_gold<R, kWaves, L>), grid, dim3(64 * kWaves), lds_bytes, \
                        ctx->stream, e0, e1, 0, FGold{}, tab, A, S, P, rows, cols, cpb, plan.s1,  \
                        plan.s2)
  const bool lx = ctx->tune.quad_lds != 0;
  if (rb == 8 && lx) GH_FUSED_GOLD(8, true);
  else if (rb == 8) GH_FUSED_GOLD(8, false);
  else if (lx) GH_FUSED_GOLD(4, true);
  else GH_FUSED_GOLD(4, false);
#undef GH_FUSED_GOLD
  GH_HIP(hipGetLastError());
  return fused_finish(ctx, ring, plan, nparts, rows, t_dev, defer, scratch);
}

// The no-allocation commitment: the fused kernel with its partial sums in
// `scratch` (this device's address of the scratch array), then the finalize
// from there; both enqueued, nothing allocated.
int launch_commit_noalloc(gh_key* key, const void* s_dev, uint32_t cols, void* t_dev,
                          void* scratch) {
  if (!scratch || key->rows % 4 != 0) return GH_E_INVALID_ARGUMENT;
  if (key->ring->kind == GH_RING_Q32)
    return launch_commit_fused(key, s_dev, cols, t_dev, false, scratch);
  return launch_commit_fused_gold(key, s_dev, cols, t_dev, false, scratch);
}

// t = A[:, :cols] * s; s (cols polys) and t (rows polys) are device pointers.
int launch_commit(gh_key* key, const void* s_dev, uint32_t cols, int domain, void* t_dev,
                  bool defer) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  if (ring->kind == GH_RING_Q32 && ctx->tune.fused && key->rows % 4 == 0 &&
      domain == GH_DOMAIN_COEFF)
    return launch_commit_fused(key, s_dev, cols, t_dev, defer);
  if (ring->kind == GH_RING_GOLDILOCKS && ctx->tune.fused && key->rows % 4 == 0 &&
      domain == GH_DOMAIN_COEFF)
    return launch_commit_fused_gold(key, s_dev, cols, t_dev, defer);
  int fst = ctx_flush(ctx);  // this path finalizes in its own launch
  if (fst) return fst;
  const size_t wb = ring->word_bytes;
  const uint32_t rows = key->rows;
  // block shape: the tuned wave count for the 16-byte q32 kernel, else 16
  const bool mac4 = ring->kind == GH_RING_Q32 && ctx->tune.mac_variant == 1 && rows % 4 == 0;
  const int waves = mac4 ? ctx->tune.mac_waves : kMacWaves;
  // partial count: the tuned blocks per CU, at least one column per wave
  uint32_t nparts = (uint32_t)(ctx->num_cus * ctx->tune.mac_blocks_per_cu);
  uint32_t max_parts = (cols + waves - 1) / waves;
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
    case 16: st = mac_launch<16>(key, s_ntt, cols, partial, nparts, cpb, waves); break;
    case 8: st = mac_launch<8>(key, s_ntt, cols, partial, nparts, cpb, waves); break;
    case 4: st = mac_launch<4>(key, s_ntt, cols, partial, nparts, cpb, waves); break;
    case 2: st = mac_launch<2>(key, s_ntt, cols, partial, nparts, cpb, waves); break;
    default: st = mac_launch<1>(key, s_ntt, cols, partial, nparts, cpb, waves); break;
  }
  if (st) return st;
  hipEvent_t e2 = prof_mark(ctx);
  // a 16-wave block per row fills a CU: beyond one row per CU (amortization's
  // n rows), 4-wave blocks finish the rows in one round
  const bool many = rows > (uint32_t)ctx->num_cus;
  if (ring->kind == GH_RING_Q32) {
    const FQ32 f{(uint32_t)ring->q, ring->qinv32};
    if (many)
      hipLaunchKernelGGL((k_finalize4<FQ32>), dim3(rows), dim3(256), 0, ctx->stream, f,
                         (const uint32_t*)ring->dev_tab, (const uint32_t*)partial, nparts, rows,
                         (int)kTabUntwistMac, (uint32_t*)t_dev);
    else
      hipLaunchKernelGGL((k_finalize<FQ32>), dim3(rows), dim3(1024), 0, ctx->stream, f,
                         (const uint32_t*)ring->dev_tab, (const uint32_t*)partial, nparts, rows,
                         (int)kTabUntwistMac, (uint32_t*)t_dev);
  } else {
    if (many)
      hipLaunchKernelGGL((k_finalize4<FGold>), dim3(rows), dim3(256), 0, ctx->stream, FGold{},
                         (const uint64_t*)ring->dev_tab, (const uint64_t*)partial, nparts, rows,
                         (int)kTabUntwist, (uint64_t*)t_dev);
    else
      hipLaunchKernelGGL((k_finalize<FGold>), dim3(rows), dim3(1024), 0, ctx->stream, FGold{},
                         (const uint64_t*)ring->dev_tab, (const uint64_t*)partial, nparts, rows,
                         (int)kTabUntwist, (uint64_t*)t_dev);
  }
  GH_HIP(hipGetLastError());
  hipEvent_t e3 = prof_mark(ctx);
  prof_span(ctx, 0, e0, e1);
  prof_span(ctx, 1, e1, e2);
  prof_span(ctx, 2, e2, e3);
  return GH_OK;
}

}  // namespace gh
