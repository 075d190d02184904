// Register-resident length-64 transforms for one polynomial per thread, and a
// wave-cooperative variant with one slot per lane.
//
// Thread-per-polynomial: all 64 coefficients live in VGPRs; every index below
// is a compile-time constant after unrolling, so the bit-reversal permutations
// cost nothing and every twiddle index is wave-uniform (the table pointer is a
// kernel argument), which lets the compiler fetch twiddles with scalar loads
// into SGPRs — one twiddle set per wavefront.
//
// Output order matches the reference's twisted NTT (poly/ntt.rs:75-87): slot k
// holds a(psi^(2k+1)) in natural k order, because the forward path is the same
// twist + bit-reverse + DIT sequence (arith/ntt.rs:292-312). The inverse is the
// DIF + bit-reverse sequence of arith/ntt.rs:314-335 with d^-1 folded into the
// untwist table.
#pragma once

#include "gh_field.h"

namespace gh {

__host__ __device__ constexpr int bitrev6(int i) {
  return ((i & 1) << 5) | ((i & 2) << 3) | ((i & 4) << 1) | ((i & 8) >> 1) | ((i & 16) >> 3) |
         ((i & 32) >> 5);
}

// Forward transform in place. TWIST: multiply by psi^i first (negacyclic);
// otherwise the plain cyclic NTT of the NTTRing seam (arith/ntt.rs:341-384).
template <bool TWIST, class F>
__device__ __forceinline__ void fwd64(const F& f, typename F::W (&x)[64],
                                      const typename F::W* __restrict__ tab) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  W y[64];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    W v = x[i];
    if (TWIST && i != 0) v = f.mul_tw(v, tab + (kTabTwist + i) * TW);
    y[bitrev6(i)] = v;
  }
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int half = 1 << s, len = 2 << s;
#pragma unroll
    for (int st = 0; st < 64; st += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        W u = y[st + j];
        W t = y[st + half + j];
        if (j != 0) t = f.mul_tw(t, tab + (kTabFwd + half - 1 + j) * TW);
        y[st + j] = f.add(u, t);
        y[st + half + j] = f.sub(u, t);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 64; i++) x[i] = y[i];
}

// Inverse transform in place; the final per-slot multiplier comes from the
// table at `post` (untwist, untwist*R, d^-1, ...).
template <class F>
__device__ __forceinline__ void inv64(const F& f, typename F::W (&x)[64],
                                      const typename F::W* __restrict__ tab, int post) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int len = 64 >> s, half = len / 2, off = 64 - (64 >> s);
#pragma unroll
    for (int st = 0; st < 64; st += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        W e = x[st + j], o = x[st + half + j];
        x[st + j] = f.add(e, o);
        W dlt = f.sub(e, o);
        if (j != 0) dlt = f.mul_tw(dlt, tab + (kTabInv + off + j) * TW);
        x[st + half + j] = dlt;
      }
    }
  }
  W y[64];
#pragma unroll
  for (int i = 0; i < 64; i++) y[i] = f.mul_tw(x[bitrev6(i)], tab + (post + i) * TW);
#pragma unroll
  for (int i = 0; i < 64; i++) x[i] = y[i];
}

// 64-bit safe cross-lane exchange (ds_bpermute works on 32-bit words).
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int src_lane) {
  return (uint32_t)__shfl((int)v, src_lane, 64);
}
__device__ __forceinline__ uint64_t lane_read(uint64_t v, int src_lane) {
  uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src_lane, 64);
  uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src_lane, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Wave-cooperative inverse: lane k holds slot k of one polynomial (natural
// order); returns coefficient k. Same DIF stages as inv64, with the butterfly
// partner fetched across lanes. All seven per-lane twiddles are gathered
// before the first stage so their load latencies overlap (a j = 0 twiddle is
// the table's 1, so every upper lane multiplies unconditionally).
template <class F>
__device__ __forceinline__ typename F::W wave_inv64(const F& f, typename F::W x,
                                                    const typename F::W* __restrict__ tab,
                                                    int post, int lane) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  W tw[7][TW];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int half = 32 >> s, off = 64 - (64 >> s);
    const W* p = tab + (kTabInv + off + (lane & (half - 1))) * TW;
#pragma unroll
    for (int w = 0; w < TW; w++) tw[s][w] = p[w];
  }
#pragma unroll
  for (int w = 0; w < TW; w++) tw[6][w] = tab[(post + lane) * TW + w];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int half = 32 >> s;
    W other = lane_read(x, lane ^ half);
    if ((lane & half) == 0) x = f.add(x, other);
    else x = f.mul_tw(f.sub(other, x), tw[s]);
  }
  W y = lane_read(x, bitrev6(lane));
  return f.mul_tw(y, tw[6]);
}

// ---------------------------------------------------------------------------
// Quad transform: one polynomial on four adjacent lanes (t = lane & 3).
//
// With i = t + 4*i2 and k = k2 + 16*k1,
//   X_k = sum_i a_i psi^(i(2k+1))
//       = sum_t iota^(t*k1) * psi^(t(2k2+1)) * sum_i2 a_(t+4i2) (psi^4)^(i2(2k2+1)),
// iota = psi^32. Each lane runs the inner 16-point twisted NTT on its stride-4
// subsequence (uniform twiddles: the 64-point tables already hold psi^(4i) and
// the 16-point stage roots), multiplies by its lane twiddle psi^(t(2k2+1)) and
// the 4-point DFT over t finishes across the quad with two DPP exchanges.
// Four lanes per polynomial cut the per-polynomial latency about threefold,
// which matters when a call has only tens of thousands of polynomials.
// ---------------------------------------------------------------------------

__host__ __device__ constexpr int bitrev4(int i) {
  return ((i & 1) << 3) | ((i & 2) << 1) | ((i & 4) >> 1) | ((i & 8) >> 3);
}
__host__ __device__ constexpr int bitrev2(int i) { return ((i & 1) << 1) | ((i & 2) >> 1); }

template <int CTRL>
__device__ __forceinline__ uint32_t quad_dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ uint64_t quad_dpp(uint64_t v) {
  uint32_t lo = quad_dpp<CTRL>((uint32_t)v), hi = quad_dpp<CTRL>((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
constexpr int kDppXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]

// The lanes' part of the quad transform. In: x[i2] = a[t + 4*i2]. Out:
// y[k2] = psi^(t(2k2+1)) * sum_i2 a_(t+4i2) (psi^4)^(i2(2k2+1)), the inner
// 16-point twisted NTT times the lane twiddle.
// ltw: this lane's psi^(t(2k2+1)) table entries, k2 = 0..15.
template <class F>
__device__ __forceinline__ void quad_lane16(const F& f, const typename F::W (&x)[16],
                                            typename F::W (&y)[16],
                                            const typename F::W* __restrict__ tab,
                                            const typename F::W (&ltw)[16][F::kTwWords]) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    W v = x[i];
    if (i != 0) {
      if constexpr (F::kPow2Roots)  // psi^(4i) = omega^(2i)
        v = f.mul_pow2(v, (2 * F::kOmegaLog2 * i) % 192);
      else
        v = f.mul_tw(v, tab + (kTabTwist + 4 * i) * TW);
    }
    y[bitrev4(i)] = v;
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int half = 1 << s, len = 2 << s;
#pragma unroll
    for (int st = 0; st < 16; st += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        W u = y[st + j];
        W v = y[st + half + j];
        if (j != 0) {
          if constexpr (F::kPow2Roots)  // (omega^(64 / len))^j
            v = f.mul_pow2(v, (F::kOmegaLog2 * (64 / len) * j) % 192);
          else
            v = f.mul_tw(v, tab + (kTabFwd + half - 1 + j) * TW);
        }
        y[st + j] = f.add(u, v);
        y[st + half + j] = f.sub(u, v);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 16; k++) y[k] = f.mul_tw(y[k], ltw[k]);
}

// v * iota, iota = psi^32
template <class F>
__device__ __forceinline__ typename F::W mul_iota(const F& f, typename F::W v,
                                                  const typename F::W* __restrict__ tab) {
  if constexpr (F::kPow2Roots)  // psi^32 = omega^16
    return f.mul_pow2(v, (16 * F::kOmegaLog2) % 192);
  else
    return f.mul_tw(v, tab + (kTabTwist + 32) * F::kTwWords);
}

// In: x[i2] = a[t + 4*i2]. Out: x[k2] = X[16*bitrev2(t) + k2].
template <class F>
__device__ __forceinline__ void quad_fwd64(const F& f, typename F::W (&x)[16],
                                           const typename F::W* __restrict__ tab,
                                           const typename F::W (&ltw)[16][F::kTwWords], int t) {
  using W = typename F::W;
  W y[16];
  quad_lane16(f, x, y, tab, ltw);
  // 4-point DFT over t: E/O sums across lanes t, t^2 ...
#pragma unroll
  for (int k = 0; k < 16; k++) {
    W o = quad_dpp<kDppXor2>(y[k]);
    y[k] = (t & 2) ? f.sub(o, y[k]) : f.add(y[k], o);
  }
  // ... lane 3 holds C1 - C3 and takes the factor iota ...
  if (t == 3) {
#pragma unroll
    for (int k = 0; k < 16; k++) y[k] = mul_iota(f, y[k], tab);
  }
  // ... and t, t^1 combine: lanes 0..3 end with Y_0, Y_2, Y_1, Y_3.
#pragma unroll
  for (int k = 0; k < 16; k++) {
    W o = quad_dpp<kDppXor1>(y[k]);
    x[k] = (t & 1) ? f.sub(o, y[k]) : f.add(y[k], o);
  }
}

// 16 bytes of words
template <class W> struct Vec16;
template <> struct Vec16<uint32_t> {
  typedef uint32_t T __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};
template <> struct Vec16<uint64_t> {
  typedef uint64_t T __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};

// quad_fwd64 with the 4-point DFT over t done through LDS instead of lane
// exchanges: each lane writes its 16 values y_t[k2] to the polynomial's
// 64-word area sc ([t][k2], 16-byte aligned), reads back the four lanes'
// values for k2 = 4t .. 4t+3 and finishes those four outputs of each k1 on its
// own (8 additions and one multiplication by iota per k2, where the exchange
// form spends 32 selected additions and 16 multiplications per lane).
// Out: x[4*k1 + u] = X[16*k1 + 4*t + u]. The caller may store the result over
// sc: the wave's LDS operations complete in order.
template <class F>
__device__ __forceinline__ void quad_fwd64_lds(const F& f, typename F::W (&x)[16],
                                               const typename F::W* __restrict__ tab,
                                               const typename F::W (&ltw)[16][F::kTwWords],
                                               int t, typename F::W* sc) {
  using W = typename F::W;
  using V = typename Vec16<W>::T;
  constexpr int N = Vec16<W>::N;
  W y[16];
  quad_lane16(f, x, y, tab, ltw);
#pragma unroll
  for (int v = 0; v < 16 / N; v++) {
    V o;
#pragma unroll
    for (int e = 0; e < N; e++) o[e] = y[N * v + e];
    *reinterpret_cast<V*>(sc + 16 * t + N * v) = o;
  }
  // the quad's four lanes are in this wave: the reads below follow the
  // writes above in the wave's LDS order; keep the compiler from moving them
  __asm__ volatile("" ::: "memory");
  W w[4][4];  // w[t'][u] = y_t'[4t + u]
#pragma unroll
  for (int tp = 0; tp < 4; tp++) {
#pragma unroll
    for (int v = 0; v < 4 / N; v++) {
      const V o = *reinterpret_cast<const V*>(sc + 16 * tp + 4 * t + N * v);
#pragma unroll
      for (int e = 0; e < N; e++) w[tp][N * v + e] = o[e];
    }
  }
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int u = 0; u < 4; u++) {
    // X[k2 + 16 k1] = sum_t iota^(t k1) y_t[k2]
    const W s0 = f.add(w[0][u], w[2][u]), s1 = f.sub(w[0][u], w[2][u]);
    const W s2 = f.add(w[1][u], w[3][u]), s3 = mul_iota(f, f.sub(w[1][u], w[3][u]), tab);
    x[u] = f.add(s0, s2);
    x[8 + u] = f.sub(s0, s2);
    x[4 + u] = f.add(s1, s3);
    x[12 + u] = f.sub(s1, s3);
  }
}

template <class F>
__device__ __forceinline__ void quad_inv_lane16(const F& f, typename F::W (&x)[16],
                                                const typename F::W* __restrict__ tab, int post,
                                                int t);

// Inverse of quad_fwd64 with the per-coefficient factor at `post`, for a
// negacyclic product: in, lane t holds x[k2] = C[16*bitrev2(t) + k2]; out,
// x[i2] = c[t + 4*i2] * post[4*i2], where c_i = sum_k C_k psi^(-i(2k+1)).
// With i = t + 4*i2 and k = k2 + 16*k1,
//   psi^(-i(2k+1)) = psi^(-t(2k2+1)) * iota^(-t*k1) * (psi^-8)^(i2*k2) * psi^(-4*i2),
// so the 4-point inverse DFT over k1 runs across the quad first (the
// forward's two exchanges in reverse order, iota^-1 on lane 3), each lane
// multiplies by psi^(-t(2k2+1)) (kTabQuadInv), runs the 16-point inverse DFT
// over k2 (the 64-point DIF stages of length 16..2 share its roots) and
// multiplies by post[4*i2] (psi^(-4*i2) d^-1, times R where the pointwise
// product left R^-1).
template <class F>
__device__ __forceinline__ void quad_inv64(const F& f, typename F::W (&x)[16],
                                           const typename F::W* __restrict__ tab, int post,
                                           int t) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  // 4-point inverse DFT over k1: t, t^1 combine ...
#pragma unroll
  for (int k = 0; k < 16; k++) {
    W o = quad_dpp<kDppXor1>(x[k]);
    x[k] = (t & 1) ? f.sub(o, x[k]) : f.add(x[k], o);
  }
  // ... lane 3 holds C1 - C3 and takes the factor iota^-1 ...
  if (t == 3) {
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if constexpr (F::kPow2Roots)  // psi^-32 = omega^-16 = 2^(-16 * kOmegaLog2)
        x[k] = f.mul_pow2(x[k], (192 - (16 * F::kOmegaLog2) % 192) % 192);
      else
        x[k] = f.mul_tw(x[k], tab + (kTabInv + 16) * TW);  // stage 0 root iomega^16
    }
  }
  // ... and t, t^2 combine: lane t ends with D_t.
#pragma unroll
  for (int k = 0; k < 16; k++) {
    W o = quad_dpp<kDppXor2>(x[k]);
    x[k] = (t & 2) ? f.sub(o, x[k]) : f.add(x[k], o);
  }
  quad_inv_lane16(f, x, tab, post, t);
}

// The lanes' part of the quad inverse. In: x[k2] = D_t[k2] =
// sum_k1 iota^(-t*k1) C[k2 + 16*k1]. Out: x[i2] = c[t + 4*i2] * post[4*i2].
template <class F>
__device__ __forceinline__ void quad_inv_lane16(const F& f, typename F::W (&x)[16],
                                                const typename F::W* __restrict__ tab, int post,
                                                int t) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  W itw[16][TW];
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int w = 0; w < TW; w++) itw[k][w] = tab[(kTabQuadInv + 16 * t + k) * TW + w];
  quad_inv_lane16(f, x, tab, post, itw);
}
// the same with this lane's psi^(-t(2k2+1)) entries preloaded (itw[k2])
template <class F>
__device__ __forceinline__ void quad_inv_lane16(const F& f, typename F::W (&x)[16],
                                                const typename F::W* __restrict__ tab, int post,
                                                const typename F::W (&itw)[16][F::kTwWords]) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = f.mul_tw(x[k], itw[k]);
  // 16-point DIF with the 64-point tables' stages of length 16, 8, 4, 2
#pragma unroll
  for (int s = 2; s < 6; s++) {
    const int len = 64 >> s, half = len / 2, off = 64 - (64 >> s);
#pragma unroll
    for (int st = 0; st < 16; st += len) {
#pragma unroll
      for (int j = 0; j < half; j++) {
        W e = x[st + j], o = x[st + half + j];
        x[st + j] = f.add(e, o);
        W dlt = f.sub(e, o);
        if (j != 0) {
          if constexpr (F::kPow2Roots)  // (omega^-(64 / len))^j
            dlt = f.mul_pow2(dlt, (192 - (F::kOmegaLog2 * (64 / len) * j) % 192) % 192);
          else
            dlt = f.mul_tw(dlt, tab + (kTabInv + off + j) * TW);
        }
        x[st + half + j] = dlt;
      }
    }
  }
  W y[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    if constexpr (F::kPow2Roots) {
      // Goldilocks (R = 1): psi^(-4i) d^-1 = omega^(-2i) 2^-6 = 2^(-2i*kOmegaLog2 - 6),
      // one more halving for the half table: a shift, not a product
      if (post == kTabUntwist || post == kTabUntwistMul || post == kTabUntwistMac ||
          post == kTabUntwistMacHalf) {
        const int e = 2 * F::kOmegaLog2 * i + (post == kTabUntwistMacHalf ? 7 : 6);
        y[i] = f.mul_pow2(x[bitrev4(i)], (192 * 8 - e) % 192);
        continue;
      }
    }
    y[i] = f.mul_tw(x[bitrev4(i)], tab + (post + 4 * i) * TW);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = y[i];
}

// v * iota^-1
template <class F>
__device__ __forceinline__ typename F::W mul_iota_inv(const F& f, typename F::W v,
                                                      const typename F::W* __restrict__ tab) {
  if constexpr (F::kPow2Roots)  // psi^-32 = omega^-16 = 2^(-16 * kOmegaLog2)
    return f.mul_pow2(v, (192 - (16 * F::kOmegaLog2) % 192) % 192);
  else
    return f.mul_tw(v, tab + (kTabInv + 16) * F::kTwWords);  // stage 0 root iomega^16
}

// quad_inv64 with its 4-point inverse DFT over k1 done in each lane and the
// transpose through LDS (the inverse of quad_fwd64_lds). In: x[4*k1 + u] =
// C[16*k1 + 4*t + u], the layout quad_fwd64_lds leaves. Each lane finishes
// D_t'[k2] for its k2 = 4t .. 4t+3 and all t', writes them to the
// polynomial's 64-word area sc ([t'][k2], 16-byte aligned) and reads back its
// own D_t[0..15]. Out: as quad_inv64.
template <class F, class LaneTw>
__device__ __forceinline__ void quad_inv64_lds(const F& f, typename F::W (&x)[16],
                                               const typename F::W* __restrict__ tab, int post,
                                               int t, typename F::W* sc, const LaneTw& itw) {
  using W = typename F::W;
  using V = typename Vec16<W>::T;
  constexpr int N = Vec16<W>::N;
  W d[4][4];  // d[t'][u] = D_t'[4t + u]
#pragma unroll
  for (int u = 0; u < 4; u++) {
    // D_t'[k2] = sum_k1 iota^(-t' k1) C[k2 + 16 k1]
    const W s0 = f.add(x[u], x[8 + u]), s1 = f.sub(x[u], x[8 + u]);
    const W s2 = f.add(x[4 + u], x[12 + u]);
    const W s3 = mul_iota_inv(f, f.sub(x[4 + u], x[12 + u]), tab);
    d[0][u] = f.add(s0, s2);
    d[2][u] = f.sub(s0, s2);
    d[1][u] = f.add(s1, s3);
    d[3][u] = f.sub(s1, s3);
  }
#pragma unroll
  for (int tp = 0; tp < 4; tp++) {
#pragma unroll
    for (int v = 0; v < 4 / N; v++) {
      V o;
#pragma unroll
      for (int e = 0; e < N; e++) o[e] = d[tp][N * v + e];
      *reinterpret_cast<V*>(sc + 16 * tp + 4 * t + N * v) = o;
    }
  }
  // the quad's four lanes are in this wave: the reads below follow the
  // writes above in the wave's LDS order; keep the compiler from moving them
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int v = 0; v < 16 / N; v++) {
    const V o = *reinterpret_cast<const V*>(sc + 16 * t + N * v);
#pragma unroll
    for (int e = 0; e < N; e++) x[N * v + e] = o[e];
  }
  __asm__ volatile("" ::: "memory");
  quad_inv_lane16(f, x, tab, post, itw);
}
// with the lane twiddles read from the table
template <class F>
__device__ __forceinline__ void quad_inv64_lds(const F& f, typename F::W (&x)[16],
                                               const typename F::W* __restrict__ tab, int post,
                                               int t, typename F::W* sc) {
  quad_inv64_lds(f, x, tab, post, t, sc, t);
}

// LDS area per polynomial of the quad transforms' 4-point stage (64 words,
// padded to spread a wave's 16 polynomials over the banks; 16-byte aligned)
template <class W>
constexpr int kQuadStride = sizeof(W) == 4 ? 68 : 66;

// words 16*k1 + 4*t .. +3 of a polynomial, k1 = 0..3, from or to x[4*k1 ..]
template <class W>
__device__ __forceinline__ void quad_store_k1(W* poly, const W (&x)[16], int t) {
  using V = typename Vec16<W>::T;
  constexpr int N = Vec16<W>::N;
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
    for (int v = 0; v < 4 / N; v++) {
      V o;
#pragma unroll
      for (int e = 0; e < N; e++) o[e] = x[4 * k1 + N * v + e];
      *reinterpret_cast<V*>(poly + 16 * k1 + 4 * t + N * v) = o;
    }
}
template <class W>
__device__ __forceinline__ void quad_load_k1(const W* poly, W (&x)[16], int t) {
  using V = typename Vec16<W>::T;
  constexpr int N = Vec16<W>::N;
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
    for (int v = 0; v < 4 / N; v++) {
      const V o = *reinterpret_cast<const V*>(poly + 16 * k1 + 4 * t + N * v);
#pragma unroll
      for (int e = 0; e < N; e++) x[4 * k1 + N * v + e] = o[e];
    }
}

// Wave-cooperative forward twisted transform (lane k holds coefficient k).
template <bool TWIST, class F>
__device__ __forceinline__ typename F::W wave_fwd64(const F& f, typename F::W x,
                                                    const typename F::W* __restrict__ tab,
                                                    int lane) {
  using W = typename F::W;
  constexpr int TW = F::kTwWords;
  if (TWIST && lane) x = f.mul_tw(x, tab + (kTabTwist + lane) * TW);
  x = lane_read(x, bitrev6(lane));
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int half = 1 << s;
    W other = lane_read(x, lane ^ half);
    const int j = lane & (half - 1);
    if ((lane & half) == 0) {
      W t = j ? f.mul_tw(other, tab + (kTabFwd + half - 1 + j) * TW) : other;
      x = f.add(x, t);
    } else {
      W t = j ? f.mul_tw(x, tab + (kTabFwd + half - 1 + j) * TW) : x;
      x = f.sub(other, t);
    }
  }
  return x;
}

}  // namespace gh
