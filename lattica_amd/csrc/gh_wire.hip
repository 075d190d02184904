// Canonical wire encodings on the device.
//
// An encoded element is ceil(bits(q - 1) / 8) little-endian bytes
// (PrimeField's CanonicalSerialize, algebra/src/arith/prime.rs:54-58,
// :690-721); polynomials and containers are runs of elements between headers
// the host writes and parses (lattice/types.rs:183-219, :440-490,
// poly/twisted_ntt.rs:647-685, serialize/src/lib.rs:210-245). The kernels move
// the runs: unpack reads bytes into ring words and flags any element >= q
// (the reference's InvalidData), pack writes ring words as bytes. Both walk a
// host-built list of (segment, 4096-element chunk) pairs, one chunk per block
// pass, lane i on element i of the chunk, so reads and writes coalesce. When
// the element width is a multiple of four bytes and the run starts on a
// four-byte boundary (the 32-bit prime below 2^32 and Goldilocks, after
// headers of 4, 8 or 16 bytes), elements move as dwords; otherwise byte by byte.
#include <algorithm>

#include "gh_internal.h"

namespace gh {

namespace {

__device__ __forceinline__ uint64_t wire_load(const uint8_t* p, uint32_t eb, bool dw) {
  if (dw) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(p);
    uint64_t v = d[0];
    if (eb == 8) v |= (uint64_t)d[1] << 32;
    return v;
  }
  uint64_t v = 0;
  for (uint32_t b = 0; b < eb; b++) v |= (uint64_t)p[b] << (8 * b);
  return v;
}

__device__ __forceinline__ void wire_store(uint8_t* p, uint32_t eb, bool dw, uint64_t v) {
  if (dw) {
    uint32_t* d = reinterpret_cast<uint32_t*>(p);
    d[0] = (uint32_t)v;
    if (eb == 8) d[1] = (uint32_t)(v >> 32);
    return;
  }
  for (uint32_t b = 0; b < eb; b++) p[b] = (uint8_t)(v >> (8 * b));
}

template <class W>
__global__ __launch_bounds__(256) void k_wire_unpack(uint64_t q, uint32_t eb,
                                                     const WireSeg* __restrict__ segs,
                                                     const WireChunk* __restrict__ chunks,
                                                     uint32_t nchunks, uint32_t* __restrict__ flags) {
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const WireChunk ch = chunks[c];
    const WireSeg s = segs[ch.seg];
    const uint64_t end = std::min<uint64_t>(ch.first + kWireChunk, s.count);
    const uint8_t* src = reinterpret_cast<const uint8_t*>(s.bytes);
    W* dst = reinterpret_cast<W*>(s.words);
    const bool dw = (eb & 3) == 0 && (s.bytes & 3) == 0;
    bool bad = false;
    for (uint64_t i = ch.first + threadIdx.x; i < end; i += 256) {
      const uint64_t v = wire_load(src + i * eb, eb, dw);
      bad |= v >= q;
      if (dst) dst[i] = (W)v;
    }
    if (bad) atomicOr(&flags[s.flag], 1u);
  }
}

template <class W>
__global__ __launch_bounds__(256) void k_wire_pack(uint32_t eb, const WireSeg* __restrict__ segs,
                                                   const WireChunk* __restrict__ chunks,
                                                   uint32_t nchunks) {
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const WireChunk ch = chunks[c];
    const WireSeg s = segs[ch.seg];
    const uint64_t end = std::min<uint64_t>(ch.first + kWireChunk, s.count);
    const W* src = reinterpret_cast<const W*>(s.words);
    uint8_t* dst = reinterpret_cast<uint8_t*>(s.bytes);
    const bool dw = (eb & 3) == 0 && (s.bytes & 3) == 0;
    for (uint64_t i = ch.first + threadIdx.x; i < end; i += 256)
      wire_store(dst + i * eb, eb, dw, (uint64_t)src[i]);
  }
}

// dst[row][col] = key column (col + col_offset), row `row`: the inverse of the
// key's [col][row][slot] streaming layout, one wave per polynomial.
template <class W>
__global__ __launch_bounds__(256) void k_key_rowmajor(const W* __restrict__ key_a,
                                                      W* __restrict__ dst, uint32_t rows,
                                                      uint32_t cols, uint32_t col_offset) {
  const size_t polys = (size_t)rows * cols;
  const int lane = threadIdx.x & 63;
  size_t p = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const size_t stride = ((size_t)gridDim.x * blockDim.x) >> 6;
  for (; p < polys; p += stride) {
    const uint32_t row = (uint32_t)(p / cols), col = (uint32_t)(p % cols);
    dst[p * 64 + lane] = key_a[(((size_t)(col + col_offset)) * rows + row) * 64 + lane];
  }
}

unsigned wire_grid(gh_ctx* ctx, uint32_t nchunks) {
  return (unsigned)std::max<uint32_t>(1, std::min<uint32_t>(nchunks, (uint32_t)ctx->num_cus * 8));
}

}  // namespace

void wire_chunks(const std::vector<WireSeg>& segs, std::vector<WireChunk>* out) {
  out->clear();
  for (uint32_t i = 0; i < segs.size(); i++)
    for (uint64_t f = 0; f < segs[i].count; f += kWireChunk) out->push_back(WireChunk{f, i, 0});
}

int enqueue_wire_unpack(gh_ring* ring, hipStream_t stream, const WireSeg* segs,
                        const WireChunk* chunks, uint32_t nchunks, uint32_t* flags) {
  if (nchunks == 0) return GH_OK;
  const unsigned grid = wire_grid(ring->ctx, nchunks);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_wire_unpack<uint32_t>, dim3(grid), dim3(256), 0, stream, ring->q,
                       ring->wire_bytes, segs, chunks, nchunks, flags);
  else
    hipLaunchKernelGGL(k_wire_unpack<uint64_t>, dim3(grid), dim3(256), 0, stream, ring->q,
                       ring->wire_bytes, segs, chunks, nchunks, flags);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int enqueue_wire_pack(gh_ring* ring, hipStream_t stream, const WireSeg* segs,
                      const WireChunk* chunks, uint32_t nchunks) {
  if (nchunks == 0) return GH_OK;
  const unsigned grid = wire_grid(ring->ctx, nchunks);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_wire_pack<uint32_t>, dim3(grid), dim3(256), 0, stream, ring->wire_bytes,
                       segs, chunks, nchunks);
  else
    hipLaunchKernelGGL(k_wire_pack<uint64_t>, dim3(grid), dim3(256), 0, stream, ring->wire_bytes,
                       segs, chunks, nchunks);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

int launch_key_rowmajor(gh_ring* ring, const void* key_a, uint32_t rows, uint32_t cols,
                        uint32_t col_offset, void* dst) {
  const size_t polys = (size_t)rows * cols;
  if (polys == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  const unsigned grid =
      (unsigned)std::min<size_t>((polys + 3) / 4, (size_t)ctx->num_cus * 16);
  if (ring->word_bytes == 4)
    hipLaunchKernelGGL(k_key_rowmajor<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t*)key_a, (uint32_t*)dst, rows, cols, col_offset);
  else
    hipLaunchKernelGGL(k_key_rowmajor<uint64_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint64_t*)key_a, (uint64_t*)dst, rows, cols, col_offset);
  GH_HIP(hipGetLastError());
  return GH_OK;
}

}  // namespace gh
