// Batched commitments against one resident key (gh_commit_batch).
//
// The plain commitment streams the whole key once per witness and is HBM
// bound. A batch shares the key: witnesses are sorted by length (descending)
// and cut into groups of G, and every key column a block loads is multiplied
// against the G witnesses of its group that reach that column. The key is
// then read once per group instead of once per witness, and the work becomes
// VALU bound (the exact 96-bit multiply-accumulate for q < 2^32, the modular
// one for Goldilocks). Three launches:
//   k_batch_fwd_quad   twisted NTT of every witness polynomial of the batch
//                      (four lanes per polynomial, one 64-polynomial segment per
//                      block, segments listed by the host);
//   k_mac_batch_*      grid (column chunk, group, row block); per block RB rows
//                      (one wave per 4 rows), lane = slot; partial sums per
//                      (chunk, item) reduced mod q;
//   k_finalize_batch   per (item, row): sum of its chunks, inverse NTT.
// Goldilocks batches take one launch instead (k_batch_fused_gold, below):
// each block transforms its group's witness columns into LDS, multiplies them
// against the key there, and the last block of an item finalizes it.
// Every item's result equals gh_commit on the same inputs.
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "gh_internal.h"
#include "gh_ntt.h"
#include "gh_common_kernels.h"

namespace gh {

template <class F, bool LX>
__global__ __launch_bounds__(256, 4) void k_batch_fwd_quad(F f, const typename F::W* __restrict__ tab,
                                                        const BatchDesc* __restrict__ desc,
                                                        const BatchSeg* __restrict__ seg,
                                                        uint32_t nseg) {
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
  for (uint32_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const BatchSeg s = seg[sg];
    const BatchDesc d = desc[s.item];
    const uint32_t p = s.first + (threadIdx.x >> 2);
    if (p >= d.len) continue;  // the quad's lanes share p: whole quads drop out
    W x[16];
    const W* sp = reinterpret_cast<const W*>(d.src) + (size_t)p * 64 + t;
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = sp[4 * i];
    if constexpr (LX) {
      quad_fwd64_lds(f, x, tab, ltw, t, qlds + (threadIdx.x >> 2) * kQuadStride<W>);
      quad_store_k1(reinterpret_cast<W*>(d.s_ntt) + (size_t)p * 64, x, t);
      continue;
    }
    quad_fwd64(f, x, tab, ltw, t);
    uint4* dp = reinterpret_cast<uint4*>(reinterpret_cast<W*>(d.s_ntt) + (size_t)p * 64 +
                                         16 * bitrev2(t));
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

// partial layout: [chunk][item][row][slot]
template <int G, int RB>
__global__ __launch_bounds__(16 * RB) void k_mac_batch_q32(
    FQ32 f, const uint32_t* __restrict__ A, const BatchDesc* __restrict__ desc, uint32_t count,
    uint32_t rows, uint32_t chunk_cols) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t ch = blockIdx.x, g0 = blockIdx.y * G;
  const uint32_t row0 = blockIdx.z * RB + 4 * wave;
  const uint32_t c0 = ch * chunk_cols;
  const uint32_t maxlen = desc[g0].len;  // sorted: the group's first item is its longest
  if (c0 >= maxlen || row0 >= rows) return;  // wave-uniform; the last row block may be ragged
  const uint32_t c1 = c0 + chunk_cols < maxlen ? c0 + chunk_cols : maxlen;
  uint32_t len[G];
  const uint32_t* sp[G];
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    len[i] = j < count ? desc[j].len : 0;
    sp[i] = reinterpret_cast<const uint32_t*>(desc[j < count ? j : g0].s_ntt) + lane;
  }
  uint64_t lo[G][4];
  uint32_t hi[G][4];
#pragma unroll
  for (int i = 0; i < G; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) { lo[i][r] = 0; hi[i][r] = 0; }
  const size_t col_stride = (size_t)rows * 64;
  const uint32_t* ap = A + (size_t)row0 * 64 + lane;
  // rows past the key's last (a row count not a multiple of 4) load its last
  // row and are not stored
  uint32_t roff[4];
#pragma unroll
  for (int r = 0; r < 4; r++) roff[r] = (row0 + r < rows ? r : rows - 1 - row0) * 64;
  for (uint32_t c = c0; c < c1; c++) {
    uint32_t a[4];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = ap[(size_t)c * col_stride + roff[r]];
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c < len[i]) {  // uniform: lengths are block-wide values
        const uint32_t s = sp[i][(size_t)c * 64];
        mac96_rows(lo[i], hi[i], a, s);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    if (j < count && c0 < len[i]) {
      uint32_t* pp = reinterpret_cast<uint32_t*>(desc[j].partial);
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (row0 + r < rows)
          pp[((size_t)ch * rows + row0 + r) * 64 + lane] =
              f.reduce96_r2((uint32_t)lo[i][r], (uint32_t)(lo[i][r] >> 32), hi[i][r]);
    }
  }
}

// 16-byte variant for q32: a wave instruction covers 4 rows x 64 slots of a
// column (lane l: row 4w + l/16, slots 4(l%16)..+3), each witness contributes
// one 16-byte load per column, and the next column's loads are issued before
// the current column's multiply-accumulates (one-column software pipeline).
template <int G, int RB>
__global__ __launch_bounds__(16 * RB) void k_mac_batch4_q32(
    FQ32 f, const uint32_t* __restrict__ A, const BatchDesc* __restrict__ desc, uint32_t count,
    uint32_t rows, uint32_t chunk_cols) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int k0 = 4 * (lane & 15);
  const uint32_t ch = blockIdx.x, g0 = blockIdx.y * G;
  const uint32_t row_w = blockIdx.z * RB + 4 * wave;  // this wave's 4 rows
  const uint32_t c0 = ch * chunk_cols;
  const uint32_t maxlen = desc[g0].len;
  if (c0 >= maxlen || row_w >= rows) return;  // wave-uniform; the last row block may be ragged
  const uint32_t c1 = c0 + chunk_cols < maxlen ? c0 + chunk_cols : maxlen;
  uint32_t len[G];
  const u32x4* sp[G];
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    len[i] = j < count ? desc[j].len : 0;
    sp[i] = reinterpret_cast<const u32x4*>(
        reinterpret_cast<const uint32_t*>(desc[j < count ? j : g0].s_ntt) + k0);
  }
  const size_t col_stride4 = (size_t)rows * 16;  // u32x4 per column
  // a lane past the key's last row (a row count not a multiple of 4) loads
  // that row and stores nothing
  const uint32_t row = row_w + (lane >> 4);
  const uint32_t row_ld = row < rows ? row : rows - 1;
  const u32x4* ap = reinterpret_cast<const u32x4*>(A + (size_t)row_ld * 64) + (lane & 15);
  uint64_t lo[G][4];
  uint32_t hi[G][4];
#pragma unroll
  for (int i = 0; i < G; i++)
#pragma unroll
    for (int q = 0; q < 4; q++) { lo[i][q] = 0; hi[i][q] = 0; }
  const u32x4 zero = {0, 0, 0, 0};
  u32x4 an = ap[(size_t)c0 * col_stride4];
  u32x4 sn[G];
#pragma unroll
  for (int i = 0; i < G; i++) sn[i] = c0 < len[i] ? sp[i][(size_t)c0 * 16] : zero;
  for (uint32_t c = c0; c < c1; c++) {
    const u32x4 a = an;
    u32x4 s[G];
#pragma unroll
    for (int i = 0; i < G; i++) s[i] = sn[i];
    if (c + 1 < c1) {
      an = ap[(size_t)(c + 1) * col_stride4];
#pragma unroll
      for (int i = 0; i < G; i++) sn[i] = c + 1 < len[i] ? sp[i][(size_t)(c + 1) * 16] : zero;
    }
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c < len[i]) {
        mac96x4(lo[i], hi[i], a.x, a.y, a.z, a.w, s[i].x, s[i].y, s[i].z, s[i].w);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    if (j < count && c0 < len[i] && row < rows) {
      u32x4 o;
      o.x = f.reduce96_r2((uint32_t)lo[i][0], (uint32_t)(lo[i][0] >> 32), hi[i][0]);
      o.y = f.reduce96_r2((uint32_t)lo[i][1], (uint32_t)(lo[i][1] >> 32), hi[i][1]);
      o.z = f.reduce96_r2((uint32_t)lo[i][2], (uint32_t)(lo[i][2] >> 32), hi[i][2]);
      o.w = f.reduce96_r2((uint32_t)lo[i][3], (uint32_t)(lo[i][3] >> 32), hi[i][3]);
      *reinterpret_cast<u32x4*>(reinterpret_cast<uint32_t*>(desc[j].partial) +
                                ((size_t)ch * rows + row) * 64 + k0) = o;
    }
  }
}

template <int G, int RB>
__global__ __launch_bounds__(16 * RB) void k_mac_batch_gold(
    FGold f, const uint64_t* __restrict__ A, const BatchDesc* __restrict__ desc, uint32_t count,
    uint32_t rows, uint32_t chunk_cols) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t ch = blockIdx.x, g0 = blockIdx.y * G;
  const uint32_t row0 = blockIdx.z * RB + 4 * wave;
  const uint32_t c0 = ch * chunk_cols;
  const uint32_t maxlen = desc[g0].len;
  if (c0 >= maxlen || row0 >= rows) return;  // wave-uniform; the last row block may be ragged
  const uint32_t c1 = c0 + chunk_cols < maxlen ? c0 + chunk_cols : maxlen;
  uint32_t len[G];
  const uint64_t* sp[G];
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    len[i] = j < count ? desc[j].len : 0;
    sp[i] = reinterpret_cast<const uint64_t*>(desc[j < count ? j : g0].s_ntt) + lane;
  }
  // groups of up to 4 accumulate lazily (exact 96-bit partial products, one
  // reduction per chunk); larger groups keep one modular sum per accumulator,
  // which needs fewer registers
  constexpr bool kLazy = G <= 4;
  GoldAcc lz[kLazy ? G : 1][4];
  uint64_t acc[kLazy ? 1 : G][4];
#pragma unroll
  for (int i = 0; i < (kLazy ? 1 : G); i++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[i][r] = 0;
  const size_t col_stride = (size_t)rows * 64;
  const uint64_t* ap = A + (size_t)row0 * 64 + lane;
  // rows past the key's last load its last row and are not stored
  uint32_t roff[4];
#pragma unroll
  for (int r = 0; r < 4; r++) roff[r] = (row0 + r < rows ? r : rows - 1 - row0) * 64;
  // one-column software pipeline: the next column's loads precede this one's math
  uint64_t an[4], sn[G];
#pragma unroll
  for (int r = 0; r < 4; r++) an[r] = ap[(size_t)c0 * col_stride + roff[r]];
#pragma unroll
  for (int i = 0; i < G; i++) sn[i] = c0 < len[i] ? sp[i][(size_t)c0 * 64] : 0;
  for (uint32_t c = c0; c < c1; c++) {
    uint64_t a[4], s[G];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = an[r];
#pragma unroll
    for (int i = 0; i < G; i++) s[i] = sn[i];
    if (c + 1 < c1) {
#pragma unroll
      for (int r = 0; r < 4; r++) an[r] = ap[(size_t)(c + 1) * col_stride + roff[r]];
#pragma unroll
      for (int i = 0; i < G; i++) sn[i] = c + 1 < len[i] ? sp[i][(size_t)(c + 1) * 64] : 0;
    }
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c < len[i]) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          if constexpr (kLazy) gold_mac(lz[i][r], a[r], s[i]);
          else acc[i][r] = f.add(acc[i][r], f.mul(a[r], s[i]));
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    if (j < count && c0 < len[i]) {
      uint64_t* pp = reinterpret_cast<uint64_t*>(desc[j].partial);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (row0 + r >= rows) continue;
        uint64_t v;
        if constexpr (kLazy) v = gold_reduce(f, lz[i][r]);
        else v = acc[i][r];
        pp[((size_t)ch * rows + row0 + r) * 64 + lane] = v;
      }
    }
  }
}

// Grouped Goldilocks MAC with the key and witness columns staged through a
// per-wave LDS ring: each wave owns D stages of 4 KB (its 4 key rows and the
// group's 4 witness columns of one key column) and refills a stage with
// global-to-LDS loads (global_load_lds_dwordx4) as soon as it has read it, so
// D columns are in flight per wave without holding them in registers (the
// lazy sums take 144 of them). A block runs through a run of consecutive
// chunks of one group (the launcher picks `split` runs per group, so every
// block has work and the ring flows across chunk boundaries), storing each
// chunk's partial sums as it completes: the same partial sums as
// k_mac_batch_gold<4, 16>.
// Wave-private stages need no barrier: a wave waits for its own loads
// (vmcnt) before reading a stage and for its reads (lgkmcnt) before refilling it.
// 16 bytes per lane from global memory straight into LDS at lds + 16 * lane
// (global_load_lds_dwordx4; lds must be the same for every lane of the wave)
__device__ __forceinline__ void lds_dma16(const void* g, void* lds) {
  __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
}

template <int D>
__global__ __launch_bounds__(256) void k_mac_batch_gold_ring(
    FGold f, const uint64_t* __restrict__ A, const BatchDesc* __restrict__ desc, uint32_t count,
    uint32_t rows, uint32_t chunk_cols, uint32_t split) {
  constexpr int G = 4;
  __shared__ uint64_t ring[4][D][8][64];  // [wave][stage][key row 0..3, item 0..3][slot]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // block x: group x / split, y = x % split takes the y-th of `split` runs of
  // consecutive chunks of the group's columns (the ring flows through them)
  const uint32_t g0 = (blockIdx.x / split) * G, y = blockIdx.x % split;
  const uint32_t row0 = blockIdx.z * 16 + 4 * wave;  // rows % 16 == 0 (launcher)
  const uint32_t maxlen = desc[g0].len;  // sorted: the group's first item is its longest
  const uint32_t nch = (maxlen + chunk_cols - 1) / chunk_cols;
  const uint32_t per = (nch + split - 1) / split;
  uint32_t ch = y * per;
  if (ch >= nch) return;  // block-uniform
  const uint32_t c0 = ch * chunk_cols;
  const uint32_t ch_end = ch + per < nch ? ch + per : nch;
  const uint32_t c1 = ch_end * chunk_cols < maxlen ? ch_end * chunk_cols : maxlen;
  uint32_t len[G];
  const uint64_t* sbase[G];
  uint64_t* part[G];
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    len[i] = j < count ? desc[j].len : 0;
    sbase[i] = reinterpret_cast<const uint64_t*>(desc[j < count ? j : g0].s_ntt);
    part[i] = reinterpret_cast<uint64_t*>(desc[j < count ? j : g0].partial) +
              (size_t)(row0 * 64 + lane);
  }
  const uint64_t* s_g0 = reinterpret_cast<const uint64_t*>(desc[g0].s_ntt);
  // per-lane sources of the four loads of a stage: lane l moves 16 bytes
  // (words 2l, 2l + 1 of a 128-word span) into the stage at the same offset
  const size_t col_stride = (size_t)rows * 64;
  const uint64_t* key_src = A + (size_t)row0 * 64 + 2 * lane;  // + c * col_stride, + 128
  // witness loads: lanes 0..31 move item 0 (2), lanes 32..63 item 1 (3)
  const bool upper = lane >= 32;
  const int wl = 2 * (lane & 31);
  const uint64_t* wb0 = (upper ? sbase[1] : sbase[0]) + wl;
  const uint64_t* wb1 = (upper ? sbase[3] : sbase[2]) + wl;
  const uint32_t wn0 = upper ? len[1] : len[0], wn1 = upper ? len[3] : len[2];
  const uint64_t* wg = s_g0 + wl;
  auto issue = [&](uint32_t d, uint32_t c) {
    const uint64_t* k0 = key_src + (size_t)c * col_stride;
    lds_dma16(k0, &ring[wave][d][0][0]);
    lds_dma16(k0 + 128, &ring[wave][d][2][0]);
    // items past their length read the group's first (longest) item instead
    const uint64_t* w0 = (c < wn0 ? wb0 : wg) + (size_t)c * 64;
    const uint64_t* w1 = (c < wn1 ? wb1 : wg) + (size_t)c * 64;
    lds_dma16(w0, &ring[wave][d][4][0]);
    lds_dma16(w1, &ring[wave][d][6][0]);
  };
  GoldAcc lz[G][4];
  // fill the ring; columns past the range reload the last one, so that every
  // step has exactly four loads behind the stage it reads
#pragma unroll
  for (int d = 0; d < D; d++) issue(d, c0 + d < c1 ? c0 + d : c1 - 1);
  uint32_t flush_at = (ch + 1) * chunk_cols < c1 ? (ch + 1) * chunk_cols : c1;
  uint32_t d = 0;
  for (uint32_t c = c0; c < c1; c++) {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(4 * (D - 1)) : "memory");
    uint64_t a[4], sv[G];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = ring[wave][d][r][lane];
#pragma unroll
    for (int i = 0; i < G; i++) sv[i] = ring[wave][d][4 + i][lane];
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    issue(d, c + D < c1 ? c + D : c1 - 1);
    d = d + 1 == D ? 0 : d + 1;
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c < len[i]) {
#pragma unroll
        for (int r = 0; r < 4; r++) gold_mac(lz[i][r], a[r], sv[i]);
      }
    }
    if (c + 1 == flush_at) {  // uniform: the chunk's partial sums, then a fresh start
#pragma unroll
      for (int i = 0; i < G; i++) {
        if (g0 + i < count && ch * chunk_cols < len[i]) {
#pragma unroll
          for (int r = 0; r < 4; r++)
            part[i][((size_t)ch * rows + r) * 64] = gold_reduce(f, lz[i][r]);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) lz[i][r] = GoldAcc{};
      }
      ch++;
      flush_at = flush_at + chunk_cols < c1 ? flush_at + chunk_cols : c1;
    }
  }
  // the ring's last loads land before the wave ends
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// The q32 grouped MAC in the same form (groups of 4, 16-row blocks): per wave
// and stage 1 KB of key (its 4 rows, lane l's 16 bytes at 16 l) and the
// group's 4 witness columns (lanes 16 i .. 16 i + 15 move item i's 256 bytes),
// read back as lane l's 4 slots of row 4w + l / 16 and of every item; exact
// 96-bit sums, reduced per chunk as in k_mac_batch4_q32.
template <int G, int D>
__global__ __launch_bounds__(256) void k_mac_batch_q32_ring(
    FQ32 f, const uint32_t* __restrict__ A, const BatchDesc* __restrict__ desc, uint32_t count,
    uint32_t rows, uint32_t chunk_cols, uint32_t split) {
  static_assert(G == 4 || G == 8, "witness loads move four items per instruction");
  constexpr int H = G / 4;  // witness loads per stage
  // [wave][stage][key: 4 rows x 64 | witness: G items x 64]
  __shared__ uint32_t ring[4][D][256 * (1 + H)];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int k0 = 4 * (lane & 15), sub = lane >> 4;
  const uint32_t g0 = (blockIdx.x / split) * G, y = blockIdx.x % split;
  const uint32_t row_w = blockIdx.z * 16 + 4 * wave;  // rows % 16 == 0 (launcher)
  const uint32_t maxlen = desc[g0].len;
  const uint32_t nch = (maxlen + chunk_cols - 1) / chunk_cols;
  const uint32_t per = (nch + split - 1) / split;
  uint32_t ch = y * per;
  if (ch >= nch) return;  // block-uniform
  const uint32_t c0 = ch * chunk_cols;
  const uint32_t ch_end = ch + per < nch ? ch + per : nch;
  const uint32_t c1 = ch_end * chunk_cols < maxlen ? ch_end * chunk_cols : maxlen;
  uint32_t len[G];
  uint32_t* part[G];
#pragma unroll
  for (int i = 0; i < G; i++) {
    const uint32_t j = g0 + i;
    len[i] = j < count ? desc[j].len : 0;
    part[i] = reinterpret_cast<uint32_t*>(desc[j < count ? j : g0].partial) +
              (size_t)(row_w + sub) * 64 + k0;
  }
  // lane l's loads: key words 4 l .. 4 l + 3 of the wave's 4 rows; witness
  // load h: item 4 h + l / 16, slots k0 .. k0 + 3 (the group's first item
  // once past its length)
  const size_t col_stride = (size_t)rows * 64;
  const uint32_t* key_src = A + (size_t)row_w * 64 + 4 * lane;
  const uint32_t* wg = reinterpret_cast<const uint32_t*>(desc[g0].s_ntt) + k0;
  const uint32_t* wmine[H];
  uint32_t my_len[H];
#pragma unroll
  for (int h = 0; h < H; h++) {
    const uint32_t j = g0 + 4 * h + sub;
    my_len[h] = j < count ? desc[j].len : 0;
    wmine[h] = reinterpret_cast<const uint32_t*>(desc[j < count ? j : g0].s_ntt) + k0;
  }
  auto issue = [&](uint32_t d, uint32_t c) {
    lds_dma16(key_src + (size_t)c * col_stride, &ring[wave][d][0]);
#pragma unroll
    for (int h = 0; h < H; h++)
      lds_dma16((c < my_len[h] ? wmine[h] : wg) + (size_t)c * 64, &ring[wave][d][256 * (1 + h)]);
  };
  uint64_t lo[G][4];
  uint32_t hi[G][4];
#pragma unroll
  for (int i = 0; i < G; i++)
#pragma unroll
    for (int q = 0; q < 4; q++) { lo[i][q] = 0; hi[i][q] = 0; }
#pragma unroll
  for (int d = 0; d < D; d++) issue(d, c0 + d < c1 ? c0 + d : c1 - 1);
  uint32_t flush_at = (ch + 1) * chunk_cols < c1 ? (ch + 1) * chunk_cols : c1;
  uint32_t d = 0;
  for (uint32_t c = c0; c < c1; c++) {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"((1 + H) * (D - 1)) : "memory");
    const u32x4 a = *reinterpret_cast<const u32x4*>(&ring[wave][d][4 * lane]);
    u32x4 sv[G];
#pragma unroll
    for (int i = 0; i < G; i++) sv[i] = *reinterpret_cast<const u32x4*>(&ring[wave][d][256 + 64 * i + k0]);
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    issue(d, c + D < c1 ? c + D : c1 - 1);
    d = d + 1 == D ? 0 : d + 1;
#pragma unroll
    for (int i = 0; i < G; i++)
      if (c < len[i]) mac96x4(lo[i], hi[i], a.x, a.y, a.z, a.w, sv[i].x, sv[i].y, sv[i].z, sv[i].w);
    if (c + 1 == flush_at) {  // uniform: the chunk's partial sums, then a fresh start
#pragma unroll
      for (int i = 0; i < G; i++) {
        if (g0 + i < count && ch * chunk_cols < len[i]) {
          u32x4 o;
          o.x = f.reduce96_r2((uint32_t)lo[i][0], (uint32_t)(lo[i][0] >> 32), hi[i][0]);
          o.y = f.reduce96_r2((uint32_t)lo[i][1], (uint32_t)(lo[i][1] >> 32), hi[i][1]);
          o.z = f.reduce96_r2((uint32_t)lo[i][2], (uint32_t)(lo[i][2] >> 32), hi[i][2]);
          o.w = f.reduce96_r2((uint32_t)lo[i][3], (uint32_t)(lo[i][3] >> 32), hi[i][3]);
          *reinterpret_cast<u32x4*>(part[i] + (size_t)ch * rows * 64) = o;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) { lo[i][q] = 0; hi[i][q] = 0; }
      }
      ch++;
      flush_at = flush_at + chunk_cols < c1 ? flush_at + chunk_cols : c1;
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// grid (count, rows): item j's row `row` = INTT(sum over its chunks)
// Waves per (item, row) block of k_finalize_batch: batches have thousands of
// rows, and a 16-wave block would fill a CU on its own (one round of 256
// blocks at a time); 4-wave blocks let several run per CU.
constexpr int kFinBatchWaves = 4;

template <class F>
__global__ __launch_bounds__(64 * kFinBatchWaves) void k_finalize_batch(
    F f, const typename F::W* __restrict__ tab, const BatchDesc* __restrict__ desc, uint32_t rows,
    uint32_t chunk_cols, int post) {
  using W = typename F::W;
  constexpr int G = kFinBatchWaves;
  const int k = threadIdx.x & 63, g = threadIdx.x >> 6;
  const uint32_t j = blockIdx.x, row = blockIdx.y;
  const uint32_t nch = (desc[j].len + chunk_cols - 1) / chunk_cols;
  const W* partial = reinterpret_cast<const W*>(desc[j].partial);
  W acc = 0;
  for (uint32_t c0 = g; c0 < nch; c0 += G * 8) {
    W v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t c = c0 + G * u;
      v[u] = c < nch ? partial[((size_t)c * rows + row) * 64 + k] : W(0);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) acc = f.add(acc, v[u]);
  }
  __shared__ W red[G][64];
  red[g][k] = acc;
  __syncthreads();
  if (g == 0) {
    W s = red[0][k];
#pragma unroll
    for (int i = 1; i < G; i++) s = f.add(s, red[i][k]);
    reinterpret_cast<W*>(desc[j].dst)[(size_t)row * 64 + k] = wave_inv64(f, s, tab, post, k);
  }
}

// ---------------------------------------------------------------------------
// Goldilocks batch in one launch: k_batch_fused_gold.
//
// Grid (column range, group of kBfGroup items, 16-row block). A block owns the
// range [c0, c0 + range_cols) of its group's key columns and walks it in
// sub-chunks of kBfSub columns: wave w transforms item w's witness columns of
// the sub-chunk (four lanes per polynomial) into one of two LDS buffers, a
// barrier, then every wave multiplies its 4 key rows of those columns against
// the group's transformed columns, accumulating exactly and lazily (GoldAcc)
// in registers across the whole range. Two buffers need one barrier per
// sub-chunk: a wave reaching the transform of sub-chunk k + 2 has passed the
// barrier after sub-chunk k + 1's transform, which every wave reaches only
// after its MAC of sub-chunk k.
//
// At the end the block writes one partial sum per (range, item, row), and the
// last block to finish an item finalizes it: each block adds 1 to the item's
// counter in `done` after a device-scope fence; the block that brings it to
// the item's block count (its ranges x row blocks) sums the item's partials
// of every row, inverse-transforms them into t_out and sets the counter back
// to zero, ready for the next launch. No block waits for another.
//
// Measured on 2000 items of kappa 16, lengths log-uniform in [2^8, 2^14]: 12.0
// to 12.9 ms per batch (128-column ranges), against 6.4 to 6.9 ms for the
// transform, grouped MAC and finalize kernels, so batch_fused is 0 by default.
// The exact lazy sums of 4 items x 4 rows hold 144 registers through the
// transform; with them the kernel runs two waves per SIMD, and every
// sub-chunk's transform waits for witness loads that nothing has prefetched
// (groups of 2 without spills: 14.2 to 22.6 ms; one block per CU: 19.9 to
// 25.0 ms). The separate transform kernel runs four waves per SIMD and keeps
// the VALU busy.
constexpr int kBfGroup = 4;     // items per group
constexpr int kBfPolys = 64;    // polynomials per sub-chunk: one per quad of 256 threads
constexpr int kBfStride = 66;   // LDS words per transformed polynomial
constexpr size_t kBfLdsBytes = 2ull * kBfPolys * kBfStride * 8;

// G items per group, SUB = 64 / G key columns per sub-chunk (quad q transforms
// column q % SUB of item q / SUB), MINB blocks per CU for the register budget.
template <int G, int MINB>
__global__ __launch_bounds__(256, MINB) void k_batch_fused_gold(
    FGold f, const uint64_t* __restrict__ tab, const uint64_t* __restrict__ A,
    const BatchDesc* __restrict__ desc, uint32_t count, uint32_t rows, uint32_t range_cols,
    uint32_t* __restrict__ done) {
  constexpr int SUB = kBfPolys / G;
  extern __shared__ uint64_t sw[];  // [2][kBfPolys][kBfStride]
  __shared__ uint32_t fin_mask;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x & 3;
  const uint32_t g0 = blockIdx.y * G;
  const uint32_t c0 = blockIdx.x * range_cols;
  const uint32_t maxlen = desc[g0].len;  // sorted: the group's first item is its longest
  if (c0 >= maxlen) return;              // block-uniform
  const uint32_t c1 = c0 + range_cols < maxlen ? c0 + range_cols : maxlen;
  uint32_t len[G];
#pragma unroll
  for (int i = 0; i < G; i++) len[i] = g0 + i < count ? desc[g0 + i].len : 0;
  // the item this thread's quad transforms (wave-uniform: SUB >= 16)
  const uint32_t quad = threadIdx.x >> 2, my_item = g0 + quad / SUB;
  const uint32_t my_len = my_item < count ? desc[my_item].len : 0;
  const uint64_t* my_src =
      reinterpret_cast<const uint64_t*>(desc[my_item < count ? my_item : g0].src) + t;
  // this wave's 4 key rows; rows past the key's last load its last row and
  // are not stored
  const uint32_t row_w = blockIdx.z * 16 + 4 * wave;
  const bool has_rows = row_w < rows;  // wave-uniform
  const size_t col_stride = (size_t)rows * 64;
  const uint64_t* ap = A + (size_t)(has_rows ? row_w : 0) * 64 + lane;
  uint32_t roff[4];
#pragma unroll
  for (int r = 0; r < 4; r++)
    roff[r] = has_rows ? (row_w + r < rows ? r : rows - 1 - row_w) * 64 : 0;
  GoldAcc acc[G][4];
  uint32_t buf = 0;
  for (uint32_t cs = c0; cs < c1; cs += SUB, buf ^= 1) {
    uint64_t* sb = sw + (size_t)buf * (kBfPolys * kBfStride);
    {  // transform: column cs + quad % SUB of the quad's item
      const uint32_t c = cs + quad % SUB;
      if (c < my_len) {  // whole quads share c
        uint64_t x[16];
#pragma unroll
        for (int i = 0; i < 16; i++) x[i] = my_src[(size_t)c * 64 + 4 * i];
        // the lane twiddles, reloaded (cached) per sub-chunk: registers kept
        // through the MAC would push the accumulators out
        uint64_t ltw[16][1];
#pragma unroll
        for (int k = 0; k < 16; k++) ltw[k][0] = tab[kTabQuad + 16 * t + k];
        quad_fwd64(f, x, tab, ltw, t);
        uint64_t* dp = sb + quad * kBfStride + 16 * bitrev2(t);
#pragma unroll
        for (int v = 0; v < 16; v++) dp[v] = x[v];
      }
    }
    __syncthreads();
    if (has_rows) {
      const uint32_t ce = cs + SUB < c1 ? cs + SUB : c1;
      uint64_t an[4];
#pragma unroll
      for (int r = 0; r < 4; r++) an[r] = ap[(size_t)cs * col_stride + roff[r]];
      for (uint32_t c = cs; c < ce; c++) {
        uint64_t a[4];
#pragma unroll
        for (int r = 0; r < 4; r++) a[r] = an[r];
        if (c + 1 < ce) {
#pragma unroll
          for (int r = 0; r < 4; r++) an[r] = ap[(size_t)(c + 1) * col_stride + roff[r]];
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
          if (c < len[i]) {  // block-uniform
            const uint64_t s = sb[(i * SUB + (c - cs)) * kBfStride + lane];
#pragma unroll
            for (int r = 0; r < 4; r++) gold_mac(acc[i][r], a[r], s);
          }
        }
      }
    }
  }
  // partial sums of this range: partial[range][row][slot] of each item
  if (has_rows) {
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c0 >= len[i]) continue;
      uint64_t* pp = reinterpret_cast<uint64_t*>(desc[g0 + i].partial);
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (row_w + r < rows)
          pp[((size_t)blockIdx.x * rows + row_w + r) * 64 + lane] = gold_reduce(f, acc[i][r]);
    }
  }
  __threadfence();  // the partials are visible device-wide before the count
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < G; i++) {
      if (c0 >= len[i]) continue;
      const uint32_t need = ((len[i] + range_cols - 1) / range_cols) * gridDim.z;
      if (atomicAdd(&done[g0 + i], 1u) == need - 1) m |= 1u << i;
    }
    fin_mask = m;
  }
  __syncthreads();
  const uint32_t m = fin_mask;
  if (!m) return;
  __threadfence();  // every other block's partials of these items are visible
#pragma unroll
  for (int i = 0; i < G; i++) {
    if (!(m & (1u << i))) continue;
    const BatchDesc d = desc[g0 + i];
    const uint32_t nch = (d.len + range_cols - 1) / range_cols;
    const uint64_t* pp = reinterpret_cast<const uint64_t*>(d.partial);
    for (uint32_t row = wave; row < rows; row += 4) {
      uint64_t s = 0;
      for (uint32_t ch = 0; ch < nch; ch++) s = f.add(s, pp[((size_t)ch * rows + row) * 64 + lane]);
      reinterpret_cast<uint64_t*>(d.dst)[(size_t)row * 64 + lane] =
          wave_inv64(f, s, tab, (int)kTabUntwist, lane);
    }
    if (threadIdx.x == 0) done[g0 + i] = 0;
  }
}

int enqueue_batch_fused_gold(gh_key* key, hipStream_t stream, const BatchDesc* desc,
                             uint32_t count, uint32_t maxlen, uint32_t range_cols,
                             uint32_t* done) {
  if (count == 0 || maxlen == 0) return GH_OK;
  if (range_cols == 0 || range_cols % 32) return GH_E_INVALID_ARGUMENT;
  const gh_ring* ring = key->ring;
  dim3 grid((maxlen + range_cols - 1) / range_cols, (count + kBfGroup - 1) / kBfGroup,
            (key->rows + 15) / 16);
  hipLaunchKernelGGL((k_batch_fused_gold<kBfGroup, 2>), grid, dim3(256), kBfLdsBytes, stream,
                     FGold{}, (const uint64_t*)ring->dev_tab, (const uint64_t*)key->dev_a, desc,
                     count, key->rows, range_cols, done);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int enqueue_batch_xform(gh_ring* ring, hipStream_t stream, const BatchDesc* desc,
                        const BatchSeg* seg, uint32_t nseg) {
  if (nseg == 0) return GH_OK;
  const int cus = ring->ctx->num_cus;
  const unsigned grid = nseg < (uint32_t)cus * 16 ? nseg : (unsigned)cus * 16;
  const FQ32 fq{(uint32_t)ring->q, ring->qinv32};
  const bool lx = ring->ctx->tune.quad_lds != 0;
  if (ring->kind == GH_RING_Q32 && lx)
    hipLaunchKernelGGL((k_batch_fwd_quad<FQ32, true>), dim3(grid), dim3(256), 0, stream, fq,
                       (const uint32_t*)ring->dev_tab, desc, seg, nseg);
  else if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_batch_fwd_quad<FQ32, false>), dim3(grid), dim3(256), 0, stream, fq,
                       (const uint32_t*)ring->dev_tab, desc, seg, nseg);
  else if (lx)
    hipLaunchKernelGGL((k_batch_fwd_quad<FGold, true>), dim3(grid), dim3(256), 0, stream, FGold{},
                       (const uint64_t*)ring->dev_tab, desc, seg, nseg);
  else
    hipLaunchKernelGGL((k_batch_fwd_quad<FGold, false>), dim3(grid), dim3(256), 0, stream, FGold{},
                       (const uint64_t*)ring->dev_tab, desc, seg, nseg);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

// The LDS-ring grouped MACs (groups of 4, 16-row blocks), `depth` columns per
// wave: runs per group about eight blocks per block slot of the device (two per
// CU), never more than the longest group's chunks.
static uint32_t ring_split(gh_key* key, uint32_t count, uint32_t nchunks, uint32_t group) {
  const uint64_t groups = (count + group - 1) / group, rblocks = key->rows / 16;
  const uint64_t want = (uint64_t)key->ring->ctx->num_cus * 2 * 8;
  uint64_t split = (want + groups * rblocks - 1) / (groups * rblocks);
  if (split > nchunks) split = nchunks;
  return split < 1 ? 1u : (uint32_t)split;
}

static void mac_ring_launch(gh_key* key, hipStream_t stream, const BatchDesc* desc,
                            uint32_t count, uint32_t chunk_cols, uint32_t nchunks, int depth) {
  gh_ring* ring = key->ring;
  if (ring->kind == GH_RING_Q32) {
    // q32 groups: the tuned width (exact 96-bit sums take 12 registers per
    // item and row set, so 8 items fit)
    const uint32_t g = ring->ctx->tune.q32_ring_group == 8 ? 8 : 4;
    const uint32_t sp = ring_split(key, count, nchunks, g);
    const dim3 grid(((count + g - 1) / g) * sp, 1, key->rows / 16);
    const FQ32 f{(uint32_t)ring->q, ring->qinv32};
    const uint32_t* a = (const uint32_t*)key->dev_a;
    const int dep = depth < 3 ? 2 : (depth == 3 ? 3 : 4);
#define GH_Q32_RING(GG, DD)                                                                   \
  hipLaunchKernelGGL((k_mac_batch_q32_ring<GG, DD>), grid, dim3(256), 0, stream, f, a, desc, \
                     count, key->rows, chunk_cols, sp)
    if (g == 8) {
      if (dep == 2) GH_Q32_RING(8, 2);
      else if (dep == 3) GH_Q32_RING(8, 3);
      else GH_Q32_RING(8, 4);
    } else {
      if (dep == 2) GH_Q32_RING(4, 2);
      else if (dep == 3) GH_Q32_RING(4, 3);
      else GH_Q32_RING(4, 4);
    }
#undef GH_Q32_RING
    return;
  }
  const uint32_t sp = ring_split(key, count, nchunks, 4);
  const dim3 grid(((count + 3) / 4) * sp, 1, key->rows / 16);
  const uint64_t* a = (const uint64_t*)key->dev_a;
  if (depth == 2)
    hipLaunchKernelGGL((k_mac_batch_gold_ring<2>), grid, dim3(256), 0, stream, FGold{}, a, desc,
                       count, key->rows, chunk_cols, sp);
  else if (depth == 3)
    hipLaunchKernelGGL((k_mac_batch_gold_ring<3>), grid, dim3(256), 0, stream, FGold{}, a, desc,
                       count, key->rows, chunk_cols, sp);
  else
    hipLaunchKernelGGL((k_mac_batch_gold_ring<4>), grid, dim3(256), 0, stream, FGold{}, a, desc,
                       count, key->rows, chunk_cols, sp);
}

template <int G, int RB>
static void mac_batch_launch(gh_key* key, hipStream_t stream, const BatchDesc* desc,
                             uint32_t count, uint32_t chunk_cols, uint32_t nchunks, int variant) {
  gh_ring* ring = key->ring;
  if (RB == 16 && (G == 4 || ring->kind == GH_RING_Q32) && ring->ctx->tune.batch_mac_ring >= 2) {
    mac_ring_launch(key, stream, desc, count, chunk_cols, nchunks, ring->ctx->tune.batch_mac_ring);
    return;
  }
  dim3 grid(nchunks, (count + G - 1) / G, (key->rows + RB - 1) / RB);
  if constexpr (G <= 8) {
    if (ring->kind == GH_RING_Q32 && variant == 1) {
      hipLaunchKernelGGL((k_mac_batch4_q32<G, RB>), grid, dim3(16 * RB), 0, stream,
                         FQ32{(uint32_t)ring->q, ring->qinv32}, (const uint32_t*)key->dev_a, desc,
                         count, key->rows, chunk_cols);
      return;
    }
  }
  if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_mac_batch_q32<G, RB>), grid, dim3(16 * RB), 0, stream,
                       FQ32{(uint32_t)ring->q, ring->qinv32}, (const uint32_t*)key->dev_a, desc,
                       count, key->rows, chunk_cols);
  else
    hipLaunchKernelGGL((k_mac_batch_gold<G, RB>), grid, dim3(16 * RB), 0, stream, FGold{},
                       (const uint64_t*)key->dev_a, desc, count, key->rows, chunk_cols);
}

template <int G>
static void mac_batch_rb(gh_key* key, hipStream_t stream, const BatchDesc* desc, uint32_t count,
                         uint32_t chunk_cols, uint32_t nchunks, int variant) {
  const uint32_t rows = key->rows;
  if (rows % 16 == 0)
    mac_batch_launch<G, 16>(key, stream, desc, count, chunk_cols, nchunks, variant);
  else if (rows % 8 == 0)
    mac_batch_launch<G, 8>(key, stream, desc, count, chunk_cols, nchunks, variant);
  else
    mac_batch_launch<G, 4>(key, stream, desc, count, chunk_cols, nchunks, variant);
}

int enqueue_batch_mac(gh_key* key, hipStream_t stream, const BatchDesc* desc, uint32_t count,
                      uint32_t maxlen, uint32_t chunk_cols, int group, int variant) {
  if (count == 0 || maxlen == 0) return GH_OK;
  const uint32_t nchunks = (maxlen + chunk_cols - 1) / chunk_cols;
  if (group == 4) mac_batch_rb<4>(key, stream, desc, count, chunk_cols, nchunks, variant);
  else if (group == 16) mac_batch_rb<16>(key, stream, desc, count, chunk_cols, nchunks, variant);
  else mac_batch_rb<8>(key, stream, desc, count, chunk_cols, nchunks, variant);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int enqueue_batch_finalize(gh_ring* ring, hipStream_t stream, const BatchDesc* desc,
                           uint32_t count, uint32_t rows, uint32_t chunk_cols) {
  if (count == 0) return GH_OK;
  if (ring->kind == GH_RING_Q32)
    hipLaunchKernelGGL((k_finalize_batch<FQ32>), dim3(count, rows), dim3(64 * kFinBatchWaves), 0, stream,
                       FQ32{(uint32_t)ring->q, ring->qinv32}, (const uint32_t*)ring->dev_tab, desc,
                       rows, chunk_cols, (int)kTabUntwistMac);
  else
    hipLaunchKernelGGL((k_finalize_batch<FGold>), dim3(count, rows), dim3(64 * kFinBatchWaves), 0, stream,
                       FGold{}, (const uint64_t*)ring->dev_tab, desc, rows, chunk_cols,
                       (int)kTabUntwist);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

static void batch_layout(gh_key* key, const gh_commit_item* items, uint32_t n, uint32_t cc,
                         char* base, BatchLayout& L) {
  const size_t wb = key->ring->word_bytes;
  const uint32_t rows = key->rows;
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return items[x].len > items[y].len; });
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  uint64_t s_words = 0, p_words = 0;
  size_t nseg = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t len = items[i].len;
    s_words += (uint64_t)len * 64;
    p_words += (uint64_t)((len + cc - 1) / cc) * rows * 64;
    nseg += (len + 63) / 64;
  }
  L.s_bytes = al(s_words * wb);
  L.p_bytes = al(p_words * wb);
  L.d_bytes = al((size_t)n * sizeof(BatchDesc));
  L.g_bytes = al(nseg * sizeof(BatchSeg));
  L.desc.resize(n);
  L.seg.clear();
  L.seg.reserve(nseg);
  uint64_t s_at = 0, p_at = 0;
  for (uint32_t i = 0; i < n; i++) {
    const gh_commit_item& it = items[order[i]];
    L.desc[i] = BatchDesc{(uint64_t)(uintptr_t)(base + s_at * wb), (uint64_t)(uintptr_t)it.witness,
                          (uint64_t)(uintptr_t)it.t_out,
                          (uint64_t)(uintptr_t)(base + L.s_bytes + p_at * wb), it.len, 0};
    for (uint32_t p = 0; p < it.len; p += 64) L.seg.push_back(BatchSeg{i, p});
    s_at += (uint64_t)it.len * 64;
    p_at += (uint64_t)((it.len + cc - 1) / cc) * rows * 64;
  }
}

static bool batch_fused(const gh_ctx* ctx, int ring_kind) {
  return ring_kind == GH_RING_GOLDILOCKS && ctx->tune.batch_fused;
}

// The batch's kernels on `stream`; descriptors already on the device at
// base + s_bytes + p_bytes. `done`: the one-launch form's zeroed counters.
static int batch_enqueue(gh_key* key, hipStream_t stream, const BatchLayout& L, char* base,
                         uint32_t cc, int group, int variant, bool fused, uint32_t* done) {
  const BatchDesc* ddesc = (const BatchDesc*)(base + L.s_bytes + L.p_bytes);
  const BatchSeg* dseg = (const BatchSeg*)(base + L.s_bytes + L.p_bytes + L.d_bytes);
  const uint32_t n = (uint32_t)L.desc.size();
  if (fused) return enqueue_batch_fused_gold(key, stream, ddesc, n, L.desc[0].len, cc, done);
  int st = enqueue_batch_xform(key->ring, stream, ddesc, dseg, (uint32_t)L.seg.size());
  if (st) return st;
  st = enqueue_batch_mac(key, stream, ddesc, n, L.desc[0].len, cc, group, variant);
  if (st) return st;
  return enqueue_batch_finalize(key->ring, stream, ddesc, n, key->rows, cc);
}

// Key columns per block of the grouped MAC: the field's tuned width, halved
// while the grid would hold fewer blocks than a target, so a batch of short
// witnesses still spreads over the device. The target is half the device's
// CUs, down to 32 columns; for the LDS-ring MACs (whose blocks run through
// several chunks when there are many) it is two blocks per CU, down to 16.
// The one-launch form uses the first rule on its column ranges.
static uint32_t batch_chunk(gh_ctx* ctx, const gh_commit_item* items, size_t count,
                            uint32_t rows, int ring_kind) {
  uint32_t maxlen = 0;
  for (size_t i = 0; i < count; i++) maxlen = std::max(maxlen, items[i].len);
  const bool fused = batch_fused(ctx, ring_kind);
  uint32_t group = fused ? (uint32_t)kBfGroup : (uint32_t)ctx->tune.batch_group;
  const bool ring = !fused && rows % 16 == 0 && ctx->tune.batch_mac_ring >= 2 &&
                    (ring_kind == GH_RING_Q32 || group == 4);
  if (ring && ring_kind == GH_RING_Q32) group = ctx->tune.q32_ring_group == 8 ? 8 : 4;
  const uint64_t groups = (count + group - 1) / group;
  const uint64_t rblocks = rows >= 16 ? rows / 16 : 1;
  uint32_t cc = fused ? (uint32_t)ctx->tune.batch_fused_cols : tuned_batch_chunk(ctx->tune, ring_kind);
  const uint32_t floor = ring ? 16u : 32u;
  const uint64_t target = ring ? (uint64_t)ctx->num_cus * 2 : (uint64_t)ctx->num_cus / 2;
  while (cc > floor && (uint64_t)((maxlen + cc - 1) / cc) * groups * rblocks < target)
    cc = std::max(floor, (cc / 2) & ~(floor - 1));
  return cc;
}

// The context's completion counters for a one-launch batch of `count` items:
// zeroed when allocated, and every launch leaves them zero.
static int ctx_batch_counters(gh_ctx* ctx, size_t count, uint32_t** out) {
  if (ctx->batch_done_items < count) {
    if (ctx->batch_done) {
      GH_HIP(hipStreamSynchronize(ctx->stream));
      GH_HIP(hipFree(ctx->batch_done));
      ctx->batch_done = nullptr;
      ctx->batch_done_items = 0;
    }
    const size_t n = std::max<size_t>(count, 4096);
    GH_HIP(hipMalloc(&ctx->batch_done, n * sizeof(uint32_t)));
    GH_HIP(hipMemsetAsync(ctx->batch_done, 0, n * sizeof(uint32_t), ctx->stream));
    ctx->batch_done_items = n;
  }
  *out = ctx->batch_done;
  return GH_OK;
}

// items: device pointers (witness coefficient domain, t_out); lengths checked by the caller.
int launch_commit_batch(gh_key* key, const gh_commit_item* items, size_t count) {
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  if (count == 0) return GH_OK;
  // one item: the plain path
  if (count == 1 || count > 0x7fffffff) {
    for (size_t i = 0; i < count; i++) {
      int st = launch_commit(key, items[i].witness, items[i].len, GH_DOMAIN_COEFF, items[i].t_out);
      if (st) return st;
    }
    return GH_OK;
  }
  const uint32_t cc = batch_chunk(ctx, items, count, key->rows, key->ring->kind);
  BatchLayout sizes;
  batch_layout(key, items, (uint32_t)count, cc, nullptr, sizes);
  void* ws;
  int st = ctx_workspace(ctx, sizes.total(), &ws);
  if (st) return st;
  char* base = (char*)ws;
  BatchLayout& L = ctx->batch_layout;
  batch_layout(key, items, (uint32_t)count, cc, base, L);
  // descriptors and segments go up in one asynchronous copy from pinned
  // staging; the previous batch's copy must have read the staging first
  const size_t dbytes = L.desc.size() * sizeof(BatchDesc), sbytes = L.seg.size() * sizeof(BatchSeg);
  if (ctx->upload_pending) GH_HIP(hipEventSynchronize(ctx->upload_done));
  ctx->upload_pending = false;
  if (!ctx->upload_done) GH_HIP(hipEventCreateWithFlags(&ctx->upload_done, hipEventDisableTiming));
  if (ctx->upload_bytes < L.d_bytes + sbytes) {
    if (ctx->upload_host) GH_HIP(hipHostFree(ctx->upload_host));
    ctx->upload_host = nullptr;
    ctx->upload_bytes = 0;
    const size_t want = std::max<size_t>(L.d_bytes + sbytes, (size_t)1 << 16);
    GH_HIP(hipHostMalloc(&ctx->upload_host, want, hipHostMallocDefault));
    ctx->upload_bytes = want;
  }
  char* h = (char*)ctx->upload_host;
  std::memcpy(h, L.desc.data(), dbytes);
  std::memcpy(h + L.d_bytes, L.seg.data(), sbytes);
  GH_HIP(hipMemcpyAsync(base + L.s_bytes + L.p_bytes, h, L.d_bytes + sbytes,
                        hipMemcpyHostToDevice, ctx->stream));
  GH_HIP(hipEventRecord(ctx->upload_done, ctx->stream));
  ctx->upload_pending = true;
  const bool fused = batch_fused(ctx, ring->kind);
  uint32_t* done = nullptr;
  if (fused) {
    if (int e = ctx_batch_counters(ctx, count, &done)) return e;
  }
  return batch_enqueue(key, ctx->stream, L, base, cc, ctx->tune.batch_group,
                       ctx->tune.mac_variant, fused, done);
}

}  // namespace gh

// ---------------------------------------------------------------------------
// Batch plans: a batch's kernels recorded once into a HIP graph over
// plan-owned device memory, replayed by one graph launch per batch.
// ---------------------------------------------------------------------------

using namespace gh;

struct gh_batch_plan {
  gh_key* key = nullptr;
  void* mem = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t count = 0;
};

static void plan_free(gh_batch_plan* p) {
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->mem) hipFree(p->mem);
  delete p;
}

extern "C" {

GH_API int gh_batch_plan_create(gh_key* key, const gh_commit_item* items, size_t count,
                                gh_batch_plan** out) {
  if (!key || !out || !items || count == 0 || count > 0x7fffffff) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  const uint32_t cols = key->msg_len + key->open_len;
  for (size_t i = 0; i < count; i++) {
    if (!items[i].witness || !items[i].t_out) return GH_E_INVALID_ARGUMENT;
    if (items[i].len == 0 || items[i].len > cols) return GH_E_DIMENSION_MISMATCH;
  }
  gh_ctx* ctx = key->ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  GH_HIP(hipSetDevice(ctx->device));
  if (int fst = ctx_flush(ctx)) return fst;
  gh_batch_plan* p = new (std::nothrow) gh_batch_plan();
  if (!p) return GH_E_OOM;
  p->key = key;
  p->count = count;
  const uint32_t cc = batch_chunk(ctx, items, count, key->rows, key->ring->kind);
  const bool fused = batch_fused(ctx, key->ring->kind);
  BatchLayout L;
  batch_layout(key, items, (uint32_t)count, cc, nullptr, L);
  // the one-launch form's counters after the layout, zeroed here (outside the
  // capture) and left zero by every replay
  const size_t done_bytes = fused ? ((count * sizeof(uint32_t) + 255) & ~(size_t)255) : 0;
  if (hipMalloc(&p->mem, L.total() + done_bytes) != hipSuccess) {
    plan_free(p);
    return GH_E_OOM;
  }
  char* base = (char*)p->mem;
  uint32_t* done = fused ? (uint32_t*)(base + L.total()) : nullptr;
  batch_layout(key, items, (uint32_t)count, cc, base, L);
  int st = GH_OK;
  if (hipMemcpy(base + L.s_bytes + L.p_bytes, L.desc.data(), L.desc.size() * sizeof(BatchDesc),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(base + L.s_bytes + L.p_bytes + L.d_bytes, L.seg.data(),
                L.seg.size() * sizeof(BatchSeg), hipMemcpyHostToDevice) != hipSuccess ||
      (fused && hipMemset(done, 0, done_bytes) != hipSuccess))
    st = GH_E_DEVICE;
  if (st == GH_OK &&
      hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    st = GH_E_DEVICE;
  if (st == GH_OK) {
    const int est = batch_enqueue(key, ctx->stream, L, base, cc, ctx->tune.batch_group,
                                  ctx->tune.mac_variant, fused, done);
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &p->graph);
    st = est ? est : (ce != hipSuccess ? GH_E_DEVICE : GH_OK);
  }
  if (st == GH_OK && hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
    st = GH_E_DEVICE;
  if (st) {
    plan_free(p);
    return st;
  }
  *out = p;
  return GH_OK;
}

GH_API int gh_batch_plan_launch(gh_batch_plan* p) {
  if (!p) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = p->key->ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  GH_HIP(hipSetDevice(ctx->device));
  if (int fst = ctx_flush(ctx)) return fst;
  GH_HIP(hipGraphLaunch(p->exec, ctx->stream));
  return GH_OK;
}

GH_API int gh_batch_plan_destroy(gh_batch_plan* p) {
  if (!p) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = p->key->ring->ctx;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipSetDevice(ctx->device);
    ctx_flush(ctx);
    hipStreamSynchronize(ctx->stream);
  }
  plan_free(p);
  return GH_OK;
}

}  // extern "C"
