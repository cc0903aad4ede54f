// Field membership and rank of one RE of an OFDM symbol under the UL-SCH multiplexing of
// TS38.212 6.2.7, in closed form from the symbol's spacings (demux_sym).  Shared by the
// device demultiplexer and its host-side check.
//
// With cnt_F(x) = |F before RE x|:
//   x in R  <=> x % dr == 0, x / dr < nr        cnt_R = min(nr, ceil(x / dr))
//   x in A  <=> x % da == 0, x / da < na        cnt_A = min(na, ceil(x / da))
//   A and R share the multiples of lcm(da, dr) below lim_ar: cnt_AR = ceil(min(x, lim_ar) / lcm)
//   T = not A, not R: cnt_T = x - cnt_A - cnt_R + cnt_AR
//   x in C1 <=> x in T, cnt_T % d1 == 0, cnt_T / d1 < n1;  cnt_C1 = min(n1, ceil(cnt_T / d1))
//   U = not A, not C1: cnt_U = x - cnt_A - cnt_C1
//   x in C2 <=> x in U, cnt_U % d2 == 0, cnt_U / d2 < n2;  cnt_C2 = min(n2, ceil(cnt_U / d2))
//   x in S  <=> x in R, (x / dr) % ds == 0, x / dr / ds < ns;  cnt_S = min(ns, ceil(x / (ds dr)))
//   UL-SCH: not A, C1, C2, rank x - cnt_A - cnt_C1 - cnt_C2.
// These are the ranks re_set_select (ulsch_demultiplex_impl.cpp:55-78) walks to: every
// d-th element of its base set, from the first.
#pragma once
#include <hip/hip_runtime.h>

#include "lh_internal.h"

namespace lh {

constexpr uint32_t kDemuxNone = 0xffffffffu;

// Rank of the RE in each field's soft-bit output (kDemuxNone: not in it); `punctured`:
// a HARQ-ACK RE of a 1- or 2-bit payload, whose UL-SCH / CSI Part 2 soft bits read zero.
struct demux_re {
  uint32_t sch, ack, c1, c2;
  bool punctured;
};

__host__ __device__ inline uint32_t demux_ceil_div(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }
__host__ __device__ inline uint32_t demux_min(uint32_t a, uint32_t b) { return a < b ? a : b; }

__host__ __device__ inline demux_re demux_classify(const demux_sym& s, uint32_t x)
{
  bool in_r = false, in_a = false, in_s = false;
  uint32_t c_a = 0, c_r = 0, c_ar = 0, c_s = 0;
  if (s.nr != 0) {
    const uint32_t j = x / s.dr;
    in_r = j * s.dr == x && j < s.nr;
    c_r = demux_min(s.nr, demux_ceil_div(x, s.dr));
    if (s.ns != 0) {
      const uint32_t k = j / s.ds;
      in_s = in_r && k * s.ds == j && k < s.ns;
      c_s = demux_min(s.ns, demux_ceil_div(x, (uint32_t)s.ds * s.dr));
    }
  }
  if (s.na != 0) {
    const uint32_t j = x / s.da;
    in_a = j * s.da == x && j < s.na;
    c_a = demux_min(s.na, demux_ceil_div(x, s.da));
    if (s.nr != 0) {
      c_ar = demux_ceil_div(demux_min(x, s.lim_ar), s.lcm_ar);
    }
  }
  bool in_c1 = false, in_c2 = false;
  uint32_t c_c1 = 0, c_c2 = 0;
  if (s.n1 != 0) {
    const uint32_t c_t = x - c_a - c_r + c_ar;
    const uint32_t j = c_t / s.d1;
    in_c1 = !in_a && !in_r && j * s.d1 == c_t && j < s.n1;
    c_c1 = demux_min(s.n1, demux_ceil_div(c_t, s.d1));
  }
  if (s.n2 != 0) {
    const uint32_t c_u = x - c_a - c_c1;
    const uint32_t j = c_u / s.d2;
    in_c2 = !in_a && !in_c1 && j * s.d2 == c_u && j < s.n2;
    c_c2 = demux_min(s.n2, demux_ceil_div(c_u, s.d2));
  }
  demux_re f;
  f.ack = (in_a || in_s) ? c_a + c_s : kDemuxNone;
  f.c1 = in_c1 ? c_c1 : kDemuxNone;
  f.c2 = in_c2 ? c_c2 : kDemuxNone;
  f.sch = (!in_a && !in_c1 && !in_c2) ? x - c_a - c_c1 - c_c2 : kDemuxNone;
  f.punctured = in_s;
  return f;
}

} // namespace lh
