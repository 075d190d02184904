// Host-side ring plan: root selection and the device twiddle table.
//
// Root selection restates the reference exactly, because NTT-domain values are
// only interchangeable with the reference's when psi is the same element:
//   find_generator   algebra/src/arith/prime.rs:645-687 (Goldilocks pinned to 7,
//                    else the smallest g whose (q-1)/p powers are all != 1)
//   root_of_unity    algebra/src/arith/prime.rs:567-585 (g^((q-1)/n), order check)
//   NttPlan::build   algebra/src/arith/ntt.rs:124-153 (stage twiddles w_len^j)
//   TwistedNttPlan   algebra/src/poly/ntt.rs:140-153 (psi = root_of_unity(2d))
#include "gh_internal.h"

#include <cstring>
#include <vector>

namespace gh {

using u128 = unsigned __int128;

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
  uint64_t r = 1 % q;
  a %= q;
  while (e) {
    if (e & 1) r = mulmod(r, a, q);
    a = mulmod(a, a, q);
    e >>= 1;
  }
  return r;
}

static bool is_prime_u64(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : small) {
    if (n == p) return true;
    if (n % p == 0) return false;
  }
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; s++; }
  for (uint64_t a : small) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; r++) {
      x = mulmod(x, x, n);
      if (x == n - 1) { comp = false; break; }
    }
    if (comp) return false;
  }
  return true;
}

static int find_generator(uint64_t q, uint64_t* g_out) {
  if (q == kGoldilocks) { *g_out = 7; return GH_OK; }
  uint64_t factors[64];
  int nf = 0;
  uint64_t n = q - 1;
  for (uint64_t d = 2; d <= n / d; d++) {
    if (n % d == 0) {
      factors[nf++] = d;
      while (n % d == 0) n /= d;
    }
  }
  if (n > 1) factors[nf++] = n;
  for (uint64_t g = 2; g < q; g++) {
    bool ok = true;
    for (int i = 0; i < nf && ok; i++) ok = powmod(g, (q - 1) / factors[i], q) != 1;
    if (ok) { *g_out = g; return GH_OK; }
  }
  return GH_E_UNSUPPORTED_SIZE;
}

static int root_of_unity(uint64_t q, uint64_t g, uint64_t n, uint64_t* out) {
  if (n == 0 || (q - 1) % n != 0) return GH_E_UNSUPPORTED_SIZE;
  uint64_t r = powmod(g, (q - 1) / n, q);
  if (powmod(r, n, q) != 1) return GH_E_UNSUPPORTED_SIZE;
  if (n > 1 && powmod(r, n / 2, q) == 1) return GH_E_UNSUPPORTED_SIZE;
  *out = r;
  return GH_OK;
}

// Fills ring->info and ring->host_tab. Error codes mirror NttError
// (arith/ntt.rs:40-49) in the order the reference raises them.
int plan_ring(uint64_t q, uint32_t d, gh_ring* ring) {
  if (d == 0 || (d & (d - 1)) != 0) return GH_E_LENGTH_NOT_POW2;
  if (q < 3 || (q & 1) == 0 || !is_prime_u64(q)) return GH_E_INVALID_ARGUMENT;
  uint64_t g;
  int st = find_generator(q, &g);
  if (st != GH_OK) return st;
  uint64_t omega, psi;
  if (root_of_unity(q, g, d, &omega) != GH_OK) return GH_E_UNSUPPORTED_SIZE;
  if (root_of_unity(q, g, 2ull * d, &psi) != GH_OK) return GH_E_UNSUPPORTED_TWIST;
  // The HIP kernels are written for d = 64 and for q < 2^32 or Goldilocks.
  if (d != 64) return GH_E_UNSUPPORTED_SIZE;
  int kind;
  if (q < (1ull << 32)) kind = GH_RING_Q32;
  else if (q == kGoldilocks) kind = GH_RING_GOLDILOCKS;
  else return GH_E_UNSUPPORTED_MODULUS;
  // the Goldilocks forward transform multiplies by its roots as shifts
  // (FGold::kOmegaLog2): they must be this psi's
  if (kind == GH_RING_GOLDILOCKS && omega != powmod(2, FGold::kOmegaLog2, q))
    return GH_E_UNSUPPORTED_TWIST;

  ring->q = q;
  ring->d = d;
  ring->kind = kind;
  ring->word_bytes = kind == GH_RING_Q32 ? 4 : 8;
  // PrimeField::SERIALIZED_BYTES (arith/prime.rs:54-58)
  uint32_t bits = 0;
  for (uint64_t v = q - 1; v; v >>= 1) bits++;
  ring->wire_bytes = (bits + 7) / 8;
  ring->generator = g;
  ring->psi = psi;
  ring->omega = omega;
  uint64_t d_inv = powmod(d, q - 2, q);
  ring->d_inv = d_inv;
  uint64_t ipsi = powmod(psi, q - 2, q), iomega = powmod(omega, q - 2, q);

  // canonical multipliers for every table entry
  std::vector<uint64_t> v(kTabEntries, 0);
  for (uint32_t i = 0; i < 64; i++) v[kTabTwist + i] = powmod(psi, i, q);
  for (int s = 0; s < 6; s++) {
    uint32_t half = 1u << s, len = 2u << s;
    uint64_t w = powmod(omega, 64 / len, q);
    for (uint32_t j = 0; j < half; j++) v[kTabFwd + half - 1 + j] = powmod(w, j, q);
  }
  for (int s = 0; s < 6; s++) {
    uint32_t len = 64u >> s, half = len / 2, off = 64 - (64 >> s);
    uint64_t w = powmod(iomega, 64 / len, q);
    for (uint32_t j = 0; j < half; j++) v[kTabInv + off + j] = powmod(w, j, q);
  }
  // Montgomery radix of the device pointwise product / accumulation
  uint64_t R = kind == GH_RING_Q32 ? (1ull << 32) % q : 1;
  uint64_t R2 = mulmod(R, R, q);
  const uint64_t inv2 = q / 2 + 1;  // q odd
  for (uint32_t i = 0; i < 64; i++) {
    uint64_t u = mulmod(powmod(ipsi, i, q), d_inv, q);
    v[kTabUntwist + i] = u;
    v[kTabUntwistMul + i] = mulmod(u, R, q);
    v[kTabUnscale + i] = d_inv;
    v[kTabUntwistMac + i] = mulmod(u, R2, q);
    v[kTabUntwistMacHalf + i] = mulmod(mulmod(u, R2, q), inv2, q);
  }
  for (uint32_t t = 0; t < 4; t++)
    for (uint32_t k = 0; k < 16; k++) {
      v[kTabQuad + 16 * t + k] = powmod(psi, t * (2 * k + 1), q);
      v[kTabQuadInv + 16 * t + k] = powmod(ipsi, t * (2 * k + 1), q);
    }
  if (kind == GH_RING_Q32) {
    uint32_t q32 = (uint32_t)q;
    uint32_t inv = 1;
    for (int i = 0; i < 5; i++) inv *= 2 - q32 * inv;  // q^-1 mod 2^32
    ring->qinv32 = inv;
    ring->host_tab.resize(kTabEntries * 2);
    uint32_t* t = reinterpret_cast<uint32_t*>(ring->host_tab.data());
    for (int i = 0; i < kTabEntries; i++) {
      uint32_t wm = (uint32_t)mulmod(v[i], R, q);
      t[2 * i] = wm;
      t[2 * i + 1] = wm * inv;
    }
  } else {
    ring->qinv32 = 0;
    ring->host_tab.resize(kTabEntries * 2);
    std::memcpy(ring->host_tab.data(), v.data(), kTabEntries * sizeof(uint64_t));
  }
  return GH_OK;
}

}  // namespace gh
