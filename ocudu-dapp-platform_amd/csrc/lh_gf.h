// GF(2) polynomial arithmetic modulo a CRC generator, for the device CRC kernels:
// multiplication by x, general multiplication, and x^n from a table of x^(2^k).
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

} // namespace lh
