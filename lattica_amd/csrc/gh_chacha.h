// ChaCha20 block function (RFC 8439 section 2.3) for the JL matrices: the
// reference draws part p's matrix from the keystream with key = seed and
// nonce = LE u32(p) || 0^8 (labrador/src/jl.rs:85-113).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gh {

__device__ __forceinline__ uint32_t rotl(uint32_t x, int n) { return __builtin_rotateleft32(x, n); }

#define GH_QR(a, b, c, d)                 \
  a += b; d ^= a; d = rotl(d, 16);        \
  c += d; b ^= c; b = rotl(b, 12);        \
  a += b; d ^= a; d = rotl(d, 8);         \
  c += d; b ^= c; b = rotl(b, 7);

struct ChaKey {
  uint32_t k[8];
  uint32_t n0;  // nonce word 0 = part index; words 1, 2 are zero
};

__device__ __forceinline__ void chacha_block(const ChaKey& key, uint32_t counter, uint32_t (&out)[16]) {
  uint32_t x0 = 0x61707865, x1 = 0x3320646e, x2 = 0x79622d32, x3 = 0x6b206574;
  uint32_t x4 = key.k[0], x5 = key.k[1], x6 = key.k[2], x7 = key.k[3];
  uint32_t x8 = key.k[4], x9 = key.k[5], x10 = key.k[6], x11 = key.k[7];
  uint32_t x12 = counter, x13 = key.n0, x14 = 0, x15 = 0;
#pragma unroll
  for (int i = 0; i < 10; i++) {
    GH_QR(x0, x4, x8, x12) GH_QR(x1, x5, x9, x13) GH_QR(x2, x6, x10, x14) GH_QR(x3, x7, x11, x15)
    GH_QR(x0, x5, x10, x15) GH_QR(x1, x6, x11, x12) GH_QR(x2, x7, x8, x13) GH_QR(x3, x4, x9, x14)
  }
  out[0] = x0 + 0x61707865; out[1] = x1 + 0x3320646e; out[2] = x2 + 0x79622d32; out[3] = x3 + 0x6b206574;
  out[4] = x4 + key.k[0]; out[5] = x5 + key.k[1]; out[6] = x6 + key.k[2]; out[7] = x7 + key.k[3];
  out[8] = x8 + key.k[4]; out[9] = x9 + key.k[5]; out[10] = x10 + key.k[6]; out[11] = x11 + key.k[7];
  out[12] = x12 + counter; out[13] = x13 + key.n0; out[14] = x14; out[15] = x15;
}

// key words from the 32-byte seed (little-endian)
inline ChaKey chacha_key(const uint8_t seed[32]) {
  ChaKey key;
  for (int i = 0; i < 8; i++)
    key.k[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
               ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
  key.n0 = 0;
  return key;
}

}  // namespace gh
