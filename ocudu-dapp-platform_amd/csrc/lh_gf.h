// GF(2) polynomial arithmetic modulo a CRC generator, for the device CRC kernels:
// multiplication by x, general multiplication, x^n from a table of x^(2^k), and a shift by
// whole bytes from a base-16 digit table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

__device__ __forceinline__ uint32_t gf_mulx(uint32_t v, uint32_t poly, int order)
{
  v <<= 1;
  return (v >> order) & 1u ? v ^ poly : v;
}
__device__ __forceinline__ uint32_t gf_mulmod(uint32_t a, uint32_t b, uint32_t poly, int order)
{
  uint32_t res = 0;
  for (int i = order - 1; i >= 0; --i) {
    res = gf_mulx(res, poly, order);
    if ((b >> i) & 1u) {
      res ^= a;
    }
  }
  return res;
}
// x^n mod G from the table of x^(2^k) mod G.
__device__ __forceinline__ uint32_t gf_xpow(uint32_t n, const uint32_t* pow2, uint32_t poly, int order)
{
  uint32_t res = 1;
  for (int k = 0; n != 0; ++k, n >>= 1) {
    if (n & 1u) {
      res = gf_mulmod(res, pow2[k], poly, order);
    }
  }
  return res;
}

// rem x^(8m) mod G (order 24): rem moved past m bytes, with win[16 i + d] = x^(8 d 16^i) mod G
// (one factor per nonzero base-16 digit of m; m < 16^6).
__device__ __forceinline__ uint32_t gf_shift_bytes(uint32_t rem, uint32_t m, const uint32_t* win, uint32_t poly)
{
  uint32_t mult = win[m & 15u];
  m >>= 4;
  for (int i = 1; m != 0; ++i, m >>= 4) {
    const uint32_t d = m & 15u;
    if (d != 0) {
      mult = gf_mulmod(mult, win[16 * i + d], poly, 24);
    }
  }
  return gf_mulmod(rem, mult, poly, 24);
}

} // namespace lh
