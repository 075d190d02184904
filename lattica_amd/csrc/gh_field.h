// Device arithmetic over Z_q for the two modulus families the GPU path runs:
//
//   FQ32  — any odd prime q < 2^32. Montgomery radix 2^32 with the signed-
//           difference REDC: for t = a*b (64-bit) and m = lo(t)*q^-1 mod 2^32,
//           (t - m*q)/2^32 = hi(t) - hi(m*q) exactly (the low words cancel), a
//           value in (-q, q) fixed by one conditional +q. This works for
//           q > 2^31 (the 32-bit q = 2^32 - 639 has no lazy headroom) and
//           needs three 32-bit multiplies when the multiplier is a fixed
//           twiddle with its w*q^-1 precomputed.
//   FGold — the Goldilocks prime p = 2^64 - 2^32 + 1 with the special-form
//           reduction 2^64 = 2^32 - 1, 2^96 = -1 (mod p); no Montgomery form.
//
// Values crossing kernels are canonical residues in [0, q). The reference does
// its arithmetic in Montgomery form with R = 2^64 (algebra/src/arith/prime.rs:
// 141-226); canonical outputs are identical because the ring operations are
// exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gh {

// native 16-byte vector (accepted by the nontemporal load builtin)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint64_t kGoldilocks = 0xffffffff00000001ull;
constexpr uint64_t kGoldEps = 0xffffffffull;  // 2^64 mod p

struct FQ32 {
  using W = uint32_t;
  static constexpr int kTwWords = 2;  // twiddle entry: (w*2^32 mod q, that * q^-1 mod 2^32)
  static constexpr bool kPow2Roots = false;
  uint32_t q;
  uint32_t qinv;  // q^-1 mod 2^32

  // r - s mod q for r, s < 2^32 with r - s in (-q, q): the borrow of the
  // subtraction (v_sub_co_u32) selects the +q. Written with the overflow
  // builtin so the compiler keeps the compare in 32 bits (an ordinary
  // `r < s` after __umulhi operands becomes a 64-bit compare plus moves).
  __device__ __forceinline__ W sub_fix(W r, W s) const {
    W d;
    const bool borrow = __builtin_sub_overflow(r, s, &d);
    return d + (borrow ? q : 0u);
  }
  __device__ __forceinline__ W add(W a, W b) const { return sub_fix(a, q - b); }
  __device__ __forceinline__ W sub(W a, W b) const { return sub_fix(a, b); }
  // a * b * 2^-32 mod q, for a < 2^32 and b < q
  __device__ __forceinline__ W mont(W a, W b) const {
    W lo = a * b;
    W hi = __umulhi(a, b);
    W m = lo * qinv;
    return sub_fix(hi, __umulhi(m, q));
  }
  // a * v mod q where tw = (v*2^32 mod q, (v*2^32 mod q) * q^-1 mod 2^32)
  __device__ __forceinline__ W mul_tw(W a, const W* tw) const {
    W m = a * tw[1];
    return sub_fix(__umulhi(a, tw[0]), __umulhi(m, q));
  }
  // REDC of a 64-bit value t < q * 2^32: t * 2^-32 mod q
  __device__ __forceinline__ W redc64(uint64_t t) const {
    W lo = (W)t, hi = (W)(t >> 32);
    W m = lo * qinv;
    return sub_fix(hi, __umulhi(m, q));
  }
  // (h2*2^64 + h1*2^32 + h0) * 2^-64 mod q, for h2 < q
  __device__ __forceinline__ W reduce96_r2(W h0, W h1, W h2) const {
    W y = redc64((uint64_t)h0);
    W z = redc64((uint64_t)h1 + y);
    return add(h2, z);
  }
};

struct FGold {
  using W = uint64_t;
  static constexpr int kTwWords = 1;  // twiddles are canonical values
  // The forward transform's roots are powers of two: omega = psi^2 = 2^39 for
  // the plan's psi = 7^((p-1)/128) = 2^43 - 2^91 (the plan checks it), and
  // 2^96 = -1, 2^192 = 1. Multiplying by them is shifting (mul_pow2).
  static constexpr bool kPow2Roots = true;
  static constexpr int kOmegaLog2 = 39;

  __device__ __forceinline__ W add(W a, W b) const {
    W s = a + b;
    if (s < a) return s + kGoldEps;  // wrapped: a+b-2^64 < p - 2^32, + eps stays < p
    return s >= kGoldilocks ? s - kGoldilocks : s;
  }
  __device__ __forceinline__ W sub(W a, W b) const {
    W r = a - b;
    return a < b ? r - kGoldEps : r;  // a-b+2^64 - eps = a-b+p
  }
  // (hi*2^64 + lo) mod p, canonical
  __device__ __forceinline__ W reduce128(W lo, W hi) const {
    W hh = hi >> 32, hl = hi & kGoldEps;
    W t0 = lo - hh;
    if (lo < hh) t0 -= kGoldEps;
    W t1 = (hl << 32) - hl;
    W t2 = t0 + t1;
    if (t2 < t1) t2 += kGoldEps;
    return t2 >= kGoldilocks ? t2 - kGoldilocks : t2;
  }
  __device__ __forceinline__ W mul(W a, W b) const { return reduce128(a * b, __umul64hi(a, b)); }
  // x * 2^e mod p for 0 <= e < 192 (e a compile-time constant after
  // unrolling): 2^96 = -1 folds the upper half into a negation, and for
  // 64 <= e < 96, 2^64 = 2^32 - 1 gives x 2^(e-32) - x 2^(e-64).
  __device__ __forceinline__ W shl_reduce(W x, int s) const {  // 0 <= s < 64
    return s == 0 ? x : reduce128(x << s, x >> (64 - s));
  }
  __device__ __forceinline__ W mul_pow2(W x, int e) const {
    const bool neg = e >= 96;
    if (neg) e -= 96;
    const W y = e < 64 ? shl_reduce(x, e) : sub(shl_reduce(x, e - 32), shl_reduce(x, e - 64));
    return neg ? (y == 0 ? 0 : kGoldilocks - y) : y;
  }
  __device__ __forceinline__ W mont(W a, W b) const { return mul(a, b); }
  __device__ __forceinline__ W mul_tw(W a, const W* tw) const { return mul(a, tw[0]); }
};

// Twiddle table layout shared by host (plan) and device (kernels), in entries
// of F::kTwWords words each. One table per ring, resident in HBM.
enum TabOffset : int {
  kTabTwist = 0,          // psi^i                       (twisted forward)
  kTabFwd = 64,           // forward DIT stage twiddles, stage s at kTabFwd + 2^s - 1
  kTabInv = 128,          // inverse DIF stage twiddles, stage s at kTabInv + 64 - (64 >> s)
  kTabUntwist = 192,      // psi^-i * d^-1                (twisted inverse)
  kTabUntwistMul = 256,   // psi^-i * d^-1 * R            (after a Montgomery pointwise product)
  kTabUnscale = 320,      // d^-1                         (cyclic inverse)
  kTabUntwistMac = 384,   // psi^-i * d^-1 * R^2          (after a reduce96_r2 accumulation)
  kTabQuad = 448,         // psi^(t*(2k+1)), t = 0..3, k = 0..15 at kTabQuad + 16t + k (quad NTT)
  kTabUntwistMacHalf = 512,  // psi^-i * d^-1 * R^2 * 2^-1  (accumulated sums halved: garbage h)
  kTabQuadInv = 576,      // psi^-(t*(2k+1)), t = 0..3, k = 0..15 at kTabQuadInv + 16t + k (quad inverse)
  kTabEntries = 640,
};

}  // namespace gh
