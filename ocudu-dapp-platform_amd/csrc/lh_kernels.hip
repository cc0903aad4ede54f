// HIP kernels of the PUSCH LDPC receive chain for gfx950 (MI355X).
//
// Decoder: a persistent workgroup pulls work items of G codeblocks; one thread per lifted
// check row r (r < Z), or per row pair (r, r + Z/2) in the two 16-bit lanes of a
// register.  A thread owns, for every layer, the check-to-variable messages of its rows,
// so the layered min-sum update needs no cross-lane exchange: the cyclic shift of each
// edge becomes the LDS address of the variable-node soft bit that the row reads and
// writes.  Soft bits live in LDS as int8; the four core layers keep their packed
// check-to-variable messages in VGPRs, the extension layers a compressed word per row in
// LDS.  The lifted edge table of each graph is read with scalar loads (uniform per
// workgroup).  Barriers separate layers (strict layer order of the reference).
#include <hip/hip_runtime.h>

#include <atomic>

#include "bg_tables.inc"
#include "lh_gf.h"
#include "lh_internal.h"

namespace lh {

// Row starts widened to 32-bit words for runtime indices in device code: the scalar unit
// loads whole dwords only, so the 16-bit table would be read through the vector memory
// path (an L1 round trip on every extension layer's critical path).
template <int N>
struct row_start_words {
  uint32_t v[N];
};
template <int N>
constexpr row_start_words<N> widen_row_starts(const uint16_t (&a)[N])
{
  row_start_words<N> w{};
  for (int i = 0; i < N; ++i) {
    w.v[i] = a[i];
  }
  return w;
}
__constant__ const row_start_words<47> kRowStart1 = widen_row_starts(lh_bg1_row_start);
__constant__ const row_start_words<43> kRowStart2 = widen_row_starts(lh_bg2_row_start);

template <int BG>
struct bgt;

template <>
struct bgt<1> {
  static constexpr int K = 22, M = 46, NNZ = 316, DEG_EXT = 10, NSHORT = 66;
  __host__ __device__ static constexpr int rs(int i) { return lh_bg1_row_start[i]; }
  __device__ static int rs_word(int i) { return (int)kRowStart1.v[i]; }
};

template <>
struct bgt<2> {
  static constexpr int K = 10, M = 42, NNZ = 197, DEG_EXT = 6, NSHORT = 50;
  __host__ __device__ static constexpr int rs(int i) { return lh_bg2_row_start[i]; }
  __device__ static int rs_word(int i) { return (int)kRowStart2.v[i]; }
};

// --- int8 log_likelihood_ratio arithmetic, values held in int -----------------------
// (lib/phy/upper/log_likelihood_ratio.cpp:22-70).  Infinite LLR: |x| > 120.
__device__ __forceinline__ int inf_mask(int x) { return -(int)((unsigned)(x + 120) > 240u); }
__device__ __forceinline__ int llr_sat(int x) { return min(max(x, -120), 120); }
__device__ __forceinline__ int sel(int m, int a, int b) { return (a & m) | (b & ~m); }
// Saturated sum with the special cases in the reference's order.
__device__ __forceinline__ int llr_add(int a, int b)
{
  int s = llr_sat(a + b);
  s = sel(inf_mask(b), b, s);
  s = sel(inf_mask(a), a, s);
  return a == -b ? 0 : s;
}
__device__ __forceinline__ int med3i(int a, int lo, int hi) { return min(max(a, lo), hi); }
// x * x for |x| < 2^11 with the full-rate 24-bit multiplier.
__device__ __forceinline__ uint32_t sq24(int x)
{
  uint32_t d;
  asm("v_mul_i32_i24 %0, %1, %1" : "=v"(d) : "v"(x));
  return d;
}
__device__ __forceinline__ uint32_t med3u_asm(uint32_t a, uint32_t b, uint32_t c)
{
  uint32_t d;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// sqrt of a perfect square n^2 <= 14400 (rounded, exact).
__device__ __forceinline__ int isqrt_sq(uint32_t q) { return (int)(__fsqrt_rn((float)q) + 0.5f); }
// round(0.8 * m) for 0 <= m <= 127 (0.8m never ends in .5).
__device__ __forceinline__ int scale08(int m) { return (4 * m + 2) / 5; }
// Learned normalisation and offset of one (iteration, layer): ab = A | B << 16 with
// A = round(256 alpha) in [0, 511] and B = round(256 beta) a signed 16-bit value;
// magnitude min(120, max(0, floor((A m - B + 128) / 256))) = round(alpha m - beta)
// clamped to the finite LLR range.  A = 205, B = -1 gives round(0.8 m) exactly.
__device__ __forceinline__ int learned_mag(int m, uint32_t ab)
{
  return med3i(((int)(ab & 0xffffu) * m - ((int)ab >> 16) + 128) >> 8, 0, 120);
}

// Soft bits in LDS hold a finite LLR in [-120, 120] or the infinity +-121 (any input
// beyond +-120, i.e. the reference's +-127, becomes +-121 on load): the promotion sum of
// the layer update is then a single clamp to [-121, 121].
//
// v2c from a soft bit and the previous check-to-variable message: saturated difference
// for a finite soft bit; an infinite soft bit passes through as +-(242..361), which never
// becomes a minimum and promotes back to +-121 whatever the new message.
__device__ __forceinline__ int v2c_of(int sv, int c)
{
  const int ex = sv - med3i(sv, -120, 120); // 0, or +-1 for an infinite soft bit
  return med3i(sv - c, -120, 120) + 241 * ex;
}

// Soft bits are addressed by their LDS byte address (one add per edge, no generic-pointer
// arithmetic).
typedef __attribute__((address_space(3))) int8_t lds_i8;
__device__ __forceinline__ lds_i8* lds_at(uint32_t a) { return reinterpret_cast<lds_i8*>((uintptr_t)a); }
__device__ __forceinline__ uint32_t lds_addr(const int8_t* p)
{
  return (uint32_t)(uintptr_t)(const lds_i8*)p;
}
// A global-memory word pointer (address space 1): loads through it are global loads,
// which a wait for LDS results does not cover (a generic pointer's flat load counts on
// both the LDS and the memory counter, so every LDS wait would also wait for it).
typedef __attribute__((address_space(1))) const uint32_t gbl_u32;
__device__ __forceinline__ gbl_u32* gbl_at(const uint32_t* p)
{
  return reinterpret_cast<gbl_u32*>((uintptr_t)p);
}
// Eight loaded values kept in registers up to this point: the loads before it are all
// issued before the first wait for their results.
template <typename T>
__device__ __forceinline__ void hold8(T (&w)[8])
{
  asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]), "+v"(w[4]), "+v"(w[5]), "+v"(w[6]), "+v"(w[7]));
}

// The N uniform edge words of a row with a runtime degree, loaded together (merged scalar
// loads, one wait) before any of them is used: loaded inside the per-edge degree test, each
// word would be loaded and waited for on its own, and a scalar-load wait also waits for the
// LDS reads in flight.  Rows shorter than their degree class read past their last edge into
// the table's padding.  Core rows (fixed degree) read the table directly and leave the
// schedule of their loads to the compiler.
template <int N>
__device__ __forceinline__ void load_edge_words(const uint32_t* __restrict__ gt, uint32_t (&w)[N])
{
#pragma unroll
  for (int e = 0; e < N; ++e) {
    w[e] = gt[e];
  }
#pragma unroll
  for (int e = 0; e < N; ++e) {
    asm volatile("" : "+s"(w[e]));
  }
}

// One layer update for check row r (pass 0: gather the deg soft bits; pass 1: v2c and
// two-min over composite keys; pass 2: c2v, promoted soft-bit update).  DEG is the
// compile-time degree (core rows) or the maximum degree with a runtime `deg`.
// Check-to-variable state, two forms:
//  * PACKED: the DEG messages themselves, int8, four per word (core layers, in VGPRs);
//  * otherwise one word: bits 0-6 scaled min1, 7-13 scaled min2, 14-18 argmin edge,
//    19-31 the message signs (extension layers, degree <= 13, in LDS).
// A zero state reproduces the reference's first visit (v2c = soft).
template <int DEG, bool RT, bool PACKED, bool LEARNED>
__device__ __forceinline__ void layer_update(int8_t* __restrict__ soft, int goff, const uint32_t* __restrict__ gt,
                                             int deg, unsigned r, unsigned Z, uint32_t* __restrict__ state,
                                             uint32_t ab)
{
  constexpr int NW = (DEG + 3) / 4;
  const uint32_t gbase = lds_addr(soft) + (uint32_t)goff;
  uint32_t addr[DEG];
  int sv[DEG];
  uint32_t gw[RT ? 2 * DEG : 1];
  if constexpr (RT) {
    load_edge_words(gt, gw);
  }
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (!RT || e < deg) {
      uint32_t sh, cz;
      if constexpr (RT) {
        sh = gw[2 * e];
        cz = gw[2 * e + 1];
      } else {
        sh = gt[2 * e];
        cz = gt[2 * e + 1];
      }
      unsigned j = r + sh;
      j = min(j, j - Z);
      addr[e] = cz + j + gbase; // one three-input add
      sv[e] = *lds_at(addr[e]);
    }
  }
  int s1o = 0, s2o = 0, amo = 0;
  uint32_t sgo = 0;
  if (!PACKED) {
    const uint32_t st = state[0];
    s1o = st & 127;
    s2o = (st >> 7) & 127;
    amo = (st >> 14) & 31;
    sgo = st >> 19;
  }
  // Composite keys v2c^2 << 5 | edge: the minimum key gives min1 and the first edge
  // reaching it (the reference's strict '<'); both inits are the reference's LLR_MAX.
  uint32_t m1 = (120u * 120u) << 5, m2 = ((120u * 120u) << 5) | 31u;
  int sp = 0;
  int v[DEG];
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (!RT || e < deg) {
      int c;
      if (PACKED) {
        c = (int)(int8_t)(state[e >> 2] >> (8 * (e & 3)));
      } else {
        const int mag = (e == amo) ? s2o : s1o;
        const int neg = -(int)((sgo >> e) & 1u);
        c = (mag ^ neg) - neg;
      }
      const int x = v2c_of(sv[e], c);
      v[e] = x;
      const uint32_t key = (sq24(x) << 5) | (uint32_t)e;
      m2 = med3u_asm(m1, m2, key); // second smallest of {m1 <= m2, key}
      m1 = min(m1, key);
      sp ^= x;
    }
  }
  const int amin = (int)(m1 & 31u);
  int s1, s2;
  if constexpr (LEARNED) {
    s1 = learned_mag(isqrt_sq(m1 >> 5), ab);
    s2 = learned_mag(isqrt_sq(m2 >> 5), ab);
  } else {
    s1 = scale08(isqrt_sq(m1 >> 5));
    s2 = scale08(isqrt_sq(m2 >> 5));
  }
  // Materialise the two scaled magnitudes once (keeps the per-edge selects cheap).
  asm volatile("" : "+v"(s1), "+v"(s2));
  uint32_t ns = 0;
  uint32_t w[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    w[k] = 0;
  }
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (!RT || e < deg) {
      const int mag = (e == amin) ? s2 : s1;
      const int neg = (sp ^ v[e]) >> 31; // 0 or -1: sign product without edge e
      const int c = (mag ^ neg) - neg;
      if (PACKED) {
        w[e >> 2] |= (uint32_t)(c & 0xff) << (8 * (e & 3));
      } else {
        ns |= (uint32_t)(neg & 1) << e;
      }
      *lds_at(addr[e]) = (int8_t)med3i(c + v[e], -121, 121); // promotion sum
    }
  }
  if (PACKED) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      state[k] = w[k];
    }
  } else {
    state[0] = (uint32_t)s1 | ((uint32_t)s2 << 7) | ((uint32_t)amin << 14) | (ns << 19);
  }
}

// Extension-layer update for check row r (one row per thread) with the messages as int8
// in LDS: the message of edge e at msg[e * Z] (zero: the reference's first visit).  The
// arithmetic is layer_update's; only where the messages live differs (no compressed word
// to take apart and put back together).
template <int DEG, bool LEARNED>
__device__ __forceinline__ void layer_update_ext8(int8_t* __restrict__ soft, int goff, const uint32_t* __restrict__ gt,
                                                  int deg, unsigned r, unsigned Z, int8_t* __restrict__ msg,
                                                  uint32_t ab)
{
  const uint32_t gbase = lds_addr(soft) + (uint32_t)goff;
  const uint32_t mbase = lds_addr(msg);
  uint32_t addr[DEG];
  int sv[DEG], cv[DEG];
  uint32_t gw[2 * DEG];
  load_edge_words(gt, gw);
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (e < deg) {
      unsigned j = r + gw[2 * e];
      j = min(j, j - Z);
      addr[e] = gw[2 * e + 1] + j + gbase;
      sv[e] = *lds_at(addr[e]);
      cv[e] = *lds_at(mbase + (uint32_t)e * Z);
    }
  }
  uint32_t m1 = (120u * 120u) << 5, m2 = ((120u * 120u) << 5) | 31u;
  int sp = 0;
  int v[DEG];
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (e < deg) {
      const int x = v2c_of(sv[e], cv[e]);
      v[e] = x;
      const uint32_t key = (sq24(x) << 5) | (uint32_t)e;
      m2 = med3u_asm(m1, m2, key);
      m1 = min(m1, key);
      sp ^= x;
    }
  }
  const int amin = (int)(m1 & 31u);
  int s1, s2;
  if constexpr (LEARNED) {
    s1 = learned_mag(isqrt_sq(m1 >> 5), ab);
    s2 = learned_mag(isqrt_sq(m2 >> 5), ab);
  } else {
    s1 = scale08(isqrt_sq(m1 >> 5));
    s2 = scale08(isqrt_sq(m2 >> 5));
  }
  asm volatile("" : "+v"(s1), "+v"(s2));
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (e < deg) {
      const int mag = (e == amin) ? s2 : s1;
      const int neg = (sp ^ v[e]) >> 31;
      const int c = (mag ^ neg) - neg;
      *lds_at(mbase + (uint32_t)e * Z) = (int8_t)c;
      *lds_at(addr[e]) = (int8_t)med3i(c + v[e], -121, 121); // promotion sum
    }
  }
}

// --- Two check rows per thread (even Z) ---------------------------------------------
// Thread p < h = Z/2 owns check rows p and p + h.  Row p + h reads column
// (p + h + shift) mod Z, which is row p's column +-h.  In this row-pair layout the rows
// j and j + h of a column are adjacent bytes (row j at byte 2 (j mod h) + (j >= h) of
// the column), so the two soft bits of an edge sit at a and a ^ 1 and one address
// serves both rows; their values travel in the two 16-bit lanes of a register and every
// packed 16-bit instruction (v_pk_*) updates both rows.  Half the threads per codeblock
// do 1.4x the instructions of one row each: the mode pays where a compute unit's
// resident codeblocks are limited by waves rather than by LDS.

// LDS byte of soft-bit index q = column * Z + j (j = q mod Z): q itself in the one-row
// layout, the row-pair byte with PAIR.
template <bool PAIR>
__device__ __forceinline__ int soft_pos(int q, int j, int Z)
{
  return !PAIR ? q : (j < Z / 2 ? q + j : q + j - Z + 1);
}
typedef short i16x2 __attribute__((ext_vector_type(2)));
// LDS bytes addressed by their 32-bit LDS address: the row-pair partner is then address
// ^ 1 with no base to add again.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ i16x2 as_i16x2(uint32_t w) { return __builtin_bit_cast(i16x2, w); }
__device__ __forceinline__ u16x2 as_u16x2(uint32_t w) { return __builtin_bit_cast(u16x2, w); }

// LDS byte of row (r + shift) mod Z of a column for thread r (and a ^ 1 for row r + h).
// With u = r + (shift mod h), f = (shift >= h), S = 2 (shift mod h) + f and
// T = Z - 1 + 2f, the byte is min(2r + S, 2r + S - T) in 16-bit unsigned arithmetic (the
// second wraps exactly when the row stays in its half; both stay below 2Z otherwise).
// The row-pair edge table holds two words per edge: col * Z | S << 16, and the 16-bit
// pair S | (S - T) << 16, so one packed add forms both candidates.  r2x2 = 2r in both
// halves.
__device__ __forceinline__ int pair_addr(uint32_t g0, uint32_t g1, uint32_t r2x2)
{
  const u16x2 t = as_u16x2(r2x2) + as_u16x2(g1);
  return (int)((g0 & 0xffffu) + min((uint32_t)t.x, (uint32_t)t.y));
}
template <typename V>
__device__ __forceinline__ uint32_t as_u32(V v)
{
  return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ i16x2 clamp2(i16x2 x, short lo, short hi)
{
  return __builtin_elementwise_min(__builtin_elementwise_max(x, (i16x2)lo), (i16x2)hi);
}
// Infinity offsets of two soft bits s in [-121, 121]: sat16(273 s) - 273 s, i.e. -266 for
// +121, 265 for -121 and 0 for a finite soft bit (273 * 120 still fits 16 bits): the
// clamp form of v_pk_mad_i16, then a wrapping v_pk_mad_i16 by -273, with wait states
// around them (the compiler's hazard tracking does not look inside the asm).
__device__ __forceinline__ i16x2 inf_offset2(i16x2 s, uint32_t k273, uint32_t km273)
{
  uint32_t d;
  asm("s_nop 1\n\tv_pk_mad_i16 %0, %1, %2, 0 clamp\n\ts_nop 1\n\tv_pk_mad_i16 %0, %1, %3, %0\n\ts_nop 1"
      : "=&v"(d)
      : "v"(as_u32(s)), "v"(k273), "v"(km273));
  return as_i16x2(d);
}

// Scaled magnitudes of both rows (lanes of `m`, each <= 120).
template <bool LEARNED>
__device__ __forceinline__ uint32_t scale_pair(uint32_t m, uint32_t ab)
{
  const int lo = (int)(m & 0xffffu), hi = (int)(m >> 16);
  if constexpr (LEARNED) {
    return (uint32_t)learned_mag(lo, ab) | ((uint32_t)learned_mag(hi, ab) << 16);
  } else {
    return (uint32_t)scale08(lo) | ((uint32_t)scale08(hi) << 16);
  }
}

// layer_update for the row pair (r, r + h): the same arithmetic per row, with keys
// |v2c| << 5 | edge (ordered like v2c^2 << 5 | edge).  State forms:
//  * PACKED: word k holds the int8 messages of edges 2k and 2k+1 as bytes
//    [2k row r, 2k+1 row r, 2k row r+h, 2k+1 row r+h];
//  * otherwise two words, row r then row r+h, each in layer_update's one-word form.
template <int DEG, bool RT, bool PACKED, bool LEARNED>
__device__ __forceinline__ void layer_update_pair(int8_t* __restrict__ soft, int goff,
                                                  const uint32_t* __restrict__ gt, int deg, unsigned r,
                                                  uint32_t* __restrict__ state, uint32_t ab)
{
  constexpr int NW = (DEG + 1) / 2;
  // Opaque (1, 1): min(x, 1) then stays one v_pk_min_u16 (with a literal 1 the compiler
  // turns it into per-lane compares and selects whose lane masks crowd the SGPRs).
  uint32_t one_w = 0x00010001u;
  asm volatile("" : "+v"(one_w));
  const u16x2 one2 = as_u16x2(one_w);
  const uint32_t k273 = 0x01110111u, km273 = 0xfeeffeefu; // (273, 273), (-273, -273)
  // Opaque (32, 32): the key |v2c| * 32 + edge stays one v_pk_mad_u16.
  uint32_t k32_w = 0x00200020u;
  asm volatile("" : "+v"(k32_w));
  const u16x2 k32 = as_u16x2(k32_w);
  // LDS address of the group's soft bits (even: the partner byte is address ^ 1).
  const uint32_t gbase = lds_addr(soft) + (uint32_t)goff;
  const uint32_t r2x2 = 2u * r * 0x10001u;
  uint32_t a1[DEG];
  i16x2 sv[DEG];
  uint32_t gw[RT ? 2 * DEG : 1];
  if constexpr (RT) {
    load_edge_words(gt, gw);
  }
  // Four edges' eight reads in flight before one wait.
#pragma unroll
  for (int e0 = 0; e0 < DEG; e0 += 4) {
    int raw[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k;
      raw[2 * k] = raw[2 * k + 1] = 0;
      if (e < DEG && (!RT || e < deg)) {
        if constexpr (RT) {
          a1[e] = (uint32_t)pair_addr(gw[2 * e], gw[2 * e + 1], r2x2) + gbase;
        } else {
          a1[e] = (uint32_t)pair_addr(gt[2 * e], gt[2 * e + 1], r2x2) + gbase;
        }
        raw[2 * k] = *lds_at(a1[e]);
        raw[2 * k + 1] = *lds_at(a1[e] ^ 1u);
      }
    }
    hold8(raw);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + k;
      if (e < DEG && (!RT || e < deg)) {
        i16x2 x;
        x.x = (short)raw[2 * k];
        x.y = (short)raw[2 * k + 1];
        sv[e] = x;
      }
    }
  }
  // Previous messages of the compressed form: magnitudes, argmin edges, sign bits.
  u16x2 S2o = 0, Do = 0, AMo = 0;
  uint32_t SGo = 0;
  if (!PACKED) {
    const uint32_t lo = state[0], hi = state[1];
    const u16x2 s1 = as_u16x2((lo & 127u) | ((hi & 127u) << 16));
    S2o = as_u16x2(((lo >> 7) & 127u) | (((hi >> 7) & 127u) << 16));
    Do = s1 - S2o;
    AMo = as_u16x2(((lo >> 14) & 31u) | (((hi >> 14) & 31u) << 16));
    SGo = (lo >> 19) | ((hi >> 19) << 16);
  }
  u16x2 m1 = (u16x2)(unsigned short)(120u << 5), m2 = (u16x2)(unsigned short)((120u << 5) | 31u);
  uint32_t sp = 0;
  i16x2 v[DEG];
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (!RT || e < deg) {
      i16x2 c;
      if (PACKED) {
        const i16x2 w = as_i16x2(state[e >> 1]);
        c = (e & 1) ? (w >> (short)8) : ((i16x2)(w << (short)8) >> (short)8);
      } else {
        const u16x2 sel = __builtin_elementwise_min(AMo ^ (u16x2)(unsigned short)e, one2);
        const i16x2 mag = __builtin_bit_cast(i16x2, (u16x2)(S2o + sel * Do));
        const i16x2 neg = as_i16x2(SGo << (15 - e)) >> (short)15;
        c = (mag ^ neg) - neg;
      }
      const i16x2 s = sv[e];
      // v2c: the saturated difference, an infinite soft bit pushed to +-(267..386).
      const i16x2 x = clamp2(s - c, -120, 120) - inf_offset2(s, k273, km273);
      v[e] = x;
      const u16x2 ax = __builtin_bit_cast(u16x2, __builtin_elementwise_max(x, -x));
      const u16x2 key = ax * k32 + (u16x2)(unsigned short)e;
      m2 = __builtin_elementwise_min(m2, __builtin_elementwise_max(m1, key));
      m1 = __builtin_elementwise_min(m1, key);
      sp ^= as_u32(x);
    }
  }
  const uint32_t m1w = as_u32(m1), m2w = as_u32(m2);
  const u16x2 amin = as_u16x2(m1w & 0x001f001fu);
  const uint32_t s1 = scale_pair<LEARNED>((m1w >> 5) & 0x07ff07ffu, ab);
  const uint32_t s2 = scale_pair<LEARNED>((m2w >> 5) & 0x07ff07ffu, ab);
  // Row signs folded into the magnitudes: message e = sign(v2c_e) * (+-s1 or +-s2).
  const i16x2 rs = as_i16x2(sp) >> (short)15;
  const i16x2 S1 = (as_i16x2(s1) ^ rs) - rs;
  const i16x2 S2 = (as_i16x2(s2) ^ rs) - rs;
  const i16x2 D = S1 - S2;
  uint32_t w[NW];
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    w[k] = 0;
  }
  uint32_t ns = 0;
  i16x2 ce = 0;
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (!RT || e < deg) {
      const u16x2 sel = __builtin_elementwise_min(amin ^ (u16x2)(unsigned short)e, one2);
      const i16x2 mag = S2 + __builtin_bit_cast(i16x2, sel) * D;
      const i16x2 neg = v[e] >> (short)15;
      const i16x2 c = (mag ^ neg) - neg;
      if (PACKED) {
        if (e & 1) {
          w[e >> 1] = __builtin_amdgcn_perm(as_u32(c), as_u32(ce), 0x06020400u);
        } else if (e + 1 == DEG) {
          w[e >> 1] = __builtin_amdgcn_perm(0u, as_u32(c), 0x06020400u);
        } else {
          ce = c;
        }
      } else {
        ns |= ((as_u32(c) >> 15) & 0x00010001u) << e;
      }
      const i16x2 y = clamp2(c + v[e], -121, 121); // promotion sum
      *lds_at(a1[e]) = (int8_t)y.x;
      *lds_at(a1[e] ^ 1u) = (int8_t)y.y;
    }
  }
  if (PACKED) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      state[k] = w[k];
    }
  } else {
    const uint32_t mg = (s1 & 0x007f007fu) | ((s2 & 0x007f007fu) << 7);
    const uint32_t am = as_u32(amin);
    state[0] = (mg & 0x3fffu) | ((am & 31u) << 14) | ((ns & 0x1fffu) << 19);
    state[1] = ((mg >> 16) & 0x3fffu) | (((am >> 16) & 31u) << 14) | (((ns >> 16) & 0x1fffu) << 19);
  }
}

template <int BG, int ROW, bool LEARNED, bool PAIR>
__device__ __forceinline__ void core_layer(int8_t* soft, int goff, const uint32_t* gt, unsigned r, unsigned Z,
                                           uint32_t* cpk, const uint32_t* abt)
{
  constexpr int E0 = bgt<BG>::rs(ROW);
  constexpr int DEG = bgt<BG>::rs(ROW + 1) - E0;
  if constexpr (PAIR) {
    layer_update_pair<DEG, false, true, LEARNED>(soft, goff, gt + 2 * E0, DEG, r, cpk, LEARNED ? abt[ROW] : 0u);
  } else {
    layer_update<DEG, false, true, LEARNED>(soft, goff, gt + 2 * E0, DEG, r, Z, cpk, LEARNED ? abt[ROW] : 0u);
  }
}

// Wave64 + workgroup reductions (workgroups of up to 16 waves).  `red` holds 32 words.
__device__ __forceinline__ void block_reduce_xor_or(uint32_t& x, uint32_t& z, uint32_t* red)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    x ^= (uint32_t)__shfl_xor((int)x, o);
    z |= (uint32_t)__shfl_xor((int)z, o);
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = x;
    red[16 + w] = z;
  }
  __syncthreads();
  x = 0;
  z = 0;
  for (int i = 0; i < nw; ++i) {
    x ^= red[i];
    z |= red[16 + i];
  }
}

__device__ __forceinline__ int block_reduce_max(int x, int* red)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    x = max(x, __shfl_xor(x, o));
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = x;
  }
  __syncthreads();
  x = red[0];
  for (int i = 1; i < nw; ++i) {
    x = max(x, red[i]);
  }
  return x;
}

// Parity of every lifted check row of layers < L and the zero-soft-bit veto
// (check_syndrome, ldpc_decoder_impl.cpp:329-387).  Nonzero when it fails.
template <int BG>
__device__ uint32_t syndrome_fail(const int8_t* soft, int goff, const uint32_t* gt, unsigned r, unsigned Z, int L)
{
  uint32_t bad = 0;
  for (int l = 0; l < L; ++l) {
    const int e0 = bgt<BG>::rs_word(l), e1 = bgt<BG>::rs_word(l + 1);
    uint32_t p = 0;
    for (int e = e0; e < e1; ++e) {
      unsigned j = r + gt[2 * e];
      j = min(j, j - Z);
      const int s = soft[(int)(gt[2 * e + 1] + j) + goff];
      bad |= (s == 0);
      p ^= (s <= 0);
    }
    bad |= p;
  }
  return bad;
}

// The same for the row pair (r, r + Z/2) of thread r in the row-pair layout.
template <int BG>
__device__ uint32_t syndrome_fail_pair(const int8_t* soft, int goff, const uint32_t* gt, unsigned r, int L)
{
  uint32_t bad = 0;
  for (int l = 0; l < L; ++l) {
    const int e0 = bgt<BG>::rs_word(l), e1 = bgt<BG>::rs_word(l + 1);
    uint32_t p = 0;
    for (int e = e0; e < e1; ++e) {
      const int a = pair_addr(gt[2 * e], gt[2 * e + 1], 2u * r * 0x10001u) + goff;
      const int s0 = soft[a], s1 = soft[a ^ 1];
      bad |= (uint32_t)(s0 == 0) | (uint32_t)(s1 == 0);
      p ^= (uint32_t)(s0 <= 0) | ((uint32_t)(s1 <= 0) << 1);
    }
    bad |= p;
  }
  return bad;
}

// Hard decisions of soft[0, nbits), MSB-first; bits of the last byte beyond nbits keep
// their previous value.  all_ones: ones instead (no-CRC failure path).
// PAIR: soft bits in the row-pair layout of lifting size Z; stride * 8 is a multiple of
// Z (the row of a thread's first bit is the same in every pass).
// Bit 7 of each byte of w: set where the byte, as int8, is <= 0 (sign bit, or zero: no
// bit of the low seven set; the per-byte sums 0x7f + x <= 0xfe never carry).
__device__ __forceinline__ uint32_t le0_mask(uint32_t w)
{
  const uint32_t nz = ((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w;
  return (w | ~nz) & 0x80808080u;
}

template <bool PAIR>
__device__ void write_hard_bits(const int8_t* soft, uint8_t* out, int nbits, bool all_ones, int first, int stride,
                                int Z)
{
  if (!all_ones && (nbits & 7) == 0 && (!PAIR || (Z & 15) == 0)) {
    // Whole output bytes from wide LDS reads (soft is 16-byte aligned).
    if (!PAIR) {
      // Bits 8b .. 8b + 7 are the eight bytes at 8b: one 8-byte read per output byte.
      const uint2* s8 = reinterpret_cast<const uint2*>(soft);
      for (int b = first; b < (nbits >> 3); b += stride) {
        const uint2 w = s8[b];
        const uint32_t m0 = le0_mask(w.x), m1 = le0_mask(w.y);
        // Bit k (k = 0 .. 7, soft bit 8b + k) goes to bit 7 - k of the byte.
        const uint32_t v = ((m0 >> 7) & 1u) | ((m0 >> 14) & 2u) | ((m0 >> 21) & 4u) | ((m0 >> 28) & 8u) |
                           ((m1 >> 3) & 16u) | ((m1 >> 10) & 32u) | ((m1 >> 17) & 64u) | ((m1 >> 24) & 128u);
        out[b] = (uint8_t)(__builtin_bitreverse32(v) >> 24);
      }
    } else {
      // Row pairs: rows j .. j + 7 of half 0 and of half 1 (j a multiple of 8 below Z/2) of
      // a column are the even and the odd bytes of 16 bytes at column * Z + 2j: one
      // 16-byte read gives output bytes (column Z + j) / 8 and (column Z + Z/2 + j) / 8.
      const int h = Z >> 1, hb = h >> 3;
      const int tasks = (nbits / Z) * hb;
      for (int t = first; t < tasks; t += stride) {
        const int c = t / hb, j = (t - c * hb) * 8;
        const uint4 w = *reinterpret_cast<const uint4*>(soft + c * Z + 2 * j);
        const uint32_t m[4] = {le0_mask(w.x), le0_mask(w.y), le0_mask(w.z), le0_mask(w.w)};
        uint32_t v0 = 0, v1 = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          // Dword q: rows 2q (byte 0), 2q + 1 (byte 2) of half 0, the same rows of half 1
          // in bytes 1 and 3.
          v0 |= (((m[q] >> 7) & 1u) | ((m[q] >> 22) & 2u)) << (2 * q);
          v1 |= (((m[q] >> 15) & 1u) | ((m[q] >> 30) & 2u)) << (2 * q);
        }
        out[(c * Z + j) >> 3] = (uint8_t)(__builtin_bitreverse32(v0) >> 24);
        out[(c * Z + h + j) >> 3] = (uint8_t)(__builtin_bitreverse32(v1) >> 24);
      }
    }
    return;
  }
  const int nbytes = (nbits + 7) >> 3;
  const int j0 = (8 * first) % Z;
  for (int b = first; b < nbytes; b += stride) {
    uint32_t byte = 0;
    int j = j0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = 8 * b + k;
      if (i < nbits) {
        byte |= (uint32_t)(all_ones || soft[soft_pos<PAIR>(i, j, Z)] <= 0) << (7 - k);
      }
      j = j + 1 == Z ? 0 : j + 1;
    }
    const int valid = nbits - 8 * b;
    if (valid < 8) {
      byte |= out[b] & (0xffu >> valid);
    }
    out[b] = (uint8_t)byte;
  }
}

// Rate-dematched soft bit at position p of the N-long buffer (behaviour of
// ldpc_rate_dematcher_impl.cpp:110-189 in gather form): fillers are +127; position p
// with non-filler rank q receives the deinterleaved inputs i = q - start (mod P),
// + P, + 2P, ... < E, each added with saturation onto 0 (new data) or the stored value.
__device__ __forceinline__ int dematch_at(const rm_params& rp, const int8_t* __restrict__ in, int old, uint32_t p,
                                          bool new_data)
{
  if (p >= rp.info && p < rp.nsys) {
    return 127;
  }
  int val = new_data ? 0 : old;
  if (p >= rp.Ncb) {
    return val;
  }
  const uint32_t q = p < rp.info ? p : p - (rp.nsys - rp.info);
  for (uint32_t i = q >= rp.start ? q - rp.start : q + rp.P - rp.start; i < rp.E; i += rp.P) {
    // e[i] = in[(i mod Kq) * Qm + i div Kq] (deinterleave_bits_Qm, :192-202); the
    // quotient is below Qm, the float estimate is corrected to the exact value.
    int qt = (int)((float)i * rp.rcp_kq);
    int rm = (int)i - qt * (int)rp.Kq;
    if (rm < 0) {
      --qt;
      rm += (int)rp.Kq;
    } else if (rm >= (int)rp.Kq) {
      ++qt;
      rm -= (int)rp.Kq;
    }
    val = llr_add(val, in[rm * (int)rp.Qm + qt]);
  }
  return val;
}

// Fused dematching of positions [0, n_in): four consecutive positions per thread, one
// dword HARQ access when aligned; writes the clamped soft bits of positions < cap.
// Soft bits of lifting size Z, NT threads per codeblock; PAIR: row-pair layout (4 NT is
// then a multiple of Z: the four positions of a thread start at the same row in every
// pass).
template <bool PAIR>
__device__ __forceinline__ void fused_positions(const rm_params& rp, const int8_t* __restrict__ src,
                                                int8_t* __restrict__ harq, bool new_data, bool wr, int n_in, int cap,
                                                int8_t* __restrict__ soft2, unsigned r, unsigned NT, int Z, int& last)
{
  const bool al4 = ((reinterpret_cast<uintptr_t>(harq)) & 3) == 0;
  const int j0 = PAIR ? (4 * (int)r) % Z : 0;
  int p_from = 4 * (int)r; // first position of the general loop below
  if (new_data && rp.E <= rp.P) {
    // New data in one pass over the circular buffer (E <= P): position p receives at
    // most the LLR e[i] = src[(i mod Kq) Qm + i div Kq], i its rank from k0 among the
    // non-filler positions (the saturated sum onto 0 is the LLR itself); fillers are
    // +127 and every other position 0.  rcp_kq is biased so that the float quotient
    // truncates to i div Kq exactly (i, Kq < 2^19).
    const uint32_t F = rp.nsys - rp.info;
    // Positions from `tail` on hold 0: past the last LLR and the fillers, or past Ncb.
    uint32_t tail = rp.Ncb;
    if (rp.start + rp.E <= rp.P) {
      const uint32_t ql = rp.start + rp.E - 1;
      tail = max(ql < rp.info ? ql + 1 : ql + F + 1, rp.nsys);
    }
    const int n_main = min(n_in, (int)((tail + 3) & ~3u));
    const int step = 4 * (int)NT;
    // Groups of four positions per thread, kGatherBatch groups at a time: their gathers
    // all go out before the first is used (one memory latency per batch instead of one
    // per group).  Branch-free: a position that takes no LLR, or a group past n_main,
    // gathers a valid byte that is not used.
    constexpr int kGatherBatch = 4;
    for (int pb = 4 * (int)r; pb < n_main; pb += kGatherBatch * step) {
      int x[kGatherBatch][4];
      uint32_t kind[kGatherBatch]; // bit k: filler, bit 4+k: LLR
#pragma unroll
      for (int b = 0; b < kGatherBatch; ++b) {
        const int p0 = pb + b * step;
        kind[b] = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint32_t p = (uint32_t)(p0 + k);
          const uint32_t q = p < rp.info ? p : p - F;
          const uint32_t i = q >= rp.start ? q - rp.start : q + rp.P - rp.start;
          const bool fill = p - rp.info < F;
          const bool llr = !fill && p < rp.Ncb && i < rp.E && p0 < n_main;
          const uint32_t qt = (uint32_t)((float)i * rp.rcp_kq);
          const uint32_t at = llr ? (i - qt * rp.Kq) * rp.Qm + qt : 0u;
          x[b][k] = src[at];
          kind[b] |= ((uint32_t)fill << k) | ((uint32_t)llr << (4 + k));
        }
      }
#pragma unroll
      for (int b = 0; b < kGatherBatch; ++b) {
        const int p0 = pb + b * step;
        if (p0 < n_main) {
          uint32_t outw = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int p = p0 + k;
            const int v = ((kind[b] >> k) & 1u) ? 127 : (((kind[b] >> (4 + k)) & 1u) ? x[b][k] : 0);
            if (p < n_in) {
              outw |= (uint32_t)(uint8_t)v << (8 * k);
              last = v != 0 ? p : last;
              if (p < cap) {
                int j = j0 + k;
                j = j >= Z ? j - Z : j;
                soft2[soft_pos<PAIR>(p, j, Z)] = (int8_t)min(max(v, -64), 64);
              }
            }
          }
          if (wr) {
            if (al4 && p0 + 4 <= n_in) {
              *reinterpret_cast<uint32_t*>(harq + p0) = outw;
            } else {
              for (int k = 0; k < 4 && p0 + k < n_in; ++k) {
                harq[p0 + k] = (int8_t)(outw >> (8 * k));
              }
            }
          }
        }
      }
    }
    // The zero tail: soft bits below cap, then the HARQ buffer in 16-byte stores from
    // its first 16-byte boundary (bytes before and after it one by one).
    const int zend = min(n_in, cap);
    const int jt0 = PAIR ? (n_main + 4 * (int)r) % Z : 0;
    for (int p0 = n_main + 4 * (int)r; p0 < zend; p0 += 4 * (int)NT) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (p0 + k < zend) {
          int j = jt0 + k;
          j = j >= Z ? j - Z : j;
          soft2[soft_pos<PAIR>(p0 + k, j, Z)] = 0;
        }
      }
    }
    if (wr && n_main < n_in) {
      const int lead = (int)((16u - (uint32_t)((reinterpret_cast<uintptr_t>(harq) + (uint32_t)n_main) & 15u)) & 15u);
      const int a16 = min(n_main + lead, n_in);
      const int nb = (n_in - a16) >> 4;
      for (int p = n_main + (int)r; p < a16; p += (int)NT) {
        harq[p] = 0;
      }
      uint4* h4 = reinterpret_cast<uint4*>(harq + a16);
      for (int b = (int)r; b < nb; b += (int)NT) {
        h4[b] = make_uint4(0u, 0u, 0u, 0u);
      }
      for (int p = a16 + 16 * nb + (int)r; p < n_in; p += (int)NT) {
        harq[p] = 0;
      }
    }
    p_from = n_in; // nothing left for the general loop
  }
  for (int p0 = p_from; p0 < n_in; p0 += 4 * (int)NT) {
    uint32_t oldw = 0;
    if (!new_data) {
      if (al4 && p0 + 4 <= n_in) {
        oldw = *reinterpret_cast<const uint32_t*>(harq + p0);
      } else {
        for (int k = 0; k < 4 && p0 + k < n_in; ++k) {
          oldw |= (uint32_t)(uint8_t)harq[p0 + k] << (8 * k);
        }
      }
    }
    uint32_t outw = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int p = p0 + k;
      if (p < n_in) {
        const int x = dematch_at(rp, src, (int)(int8_t)(oldw >> (8 * k)), (uint32_t)p, new_data);
        outw |= (uint32_t)(uint8_t)x << (8 * k);
        last = x != 0 ? p : last;
        if (p < cap) {
          int j = j0 + k;
          j = j >= Z ? j - Z : j;
          soft2[soft_pos<PAIR>(p, j, Z)] = (int8_t)min(max(x, -64), 64);
        }
      }
    }
    if (wr) {
      if (al4 && p0 + 4 <= n_in) {
        *reinterpret_cast<uint32_t*>(harq + p0) = outw;
      } else {
        for (int k = 0; k < 4 && p0 + k < n_in; ++k) {
          harq[p0 + k] = (int8_t)(outw >> (8 * k));
        }
      }
    }
  }
}

// --- Scatter load of new data (every position receives at most one LLR) ------------
// For new data with E <= P the dematched buffer is a permutation of the codeword LLRs
// into zeros and +127 fillers: LLR e[i] = cw[(i mod Kq) Qm + i div Kq] lands on rank
// q = start + i (mod P) and position p = q (q < info) or q + F.  The codeword is read in
// aligned 16-byte blocks and each byte is written to its position in natural order, in
// the LDS of the soft bits; the HARQ buffer is then written from there in 8-byte units,
// and the soft bits are clamped to +-64 (and put in the row-pair order) in place.

// Four int8 values clamped to [-64, 64] in packed 16-bit lanes.
__device__ __forceinline__ uint32_t clamp64x4(uint32_t w)
{
  const i16x2 lo = as_i16x2(w << 8) >> (short)8; // bytes 0, 2
  const i16x2 hi = as_i16x2(w) >> (short)8;      // bytes 1, 3
  const i16x2 a = clamp2(lo, -64, 64), b = clamp2(hi, -64, 64);
  return __builtin_amdgcn_perm(as_u32(b), as_u32(a), 0x06020400u);
}

// Natural position of rank q (q < P): past the filler run when q >= info.
__device__ __forceinline__ uint32_t rank_pos(uint32_t q, uint32_t info, uint32_t F)
{
  return q < info ? q : q + F;
}

// Codeword bytes c0 .. c0 + 15 of one 16-byte block (c0 = 16 b - lead) to their natural
// positions.  LQ = log2(Qm) when Qm divides c0 and the ranks do not wrap (start + E <= P):
// byte m is then LLR i = (m mod Qm) Kq + c0 / Qm + m / Qm.  LQ < 0: any Qm, lead and
// wrap, with an exact division per byte.
template <int LQ>
__device__ __forceinline__ void scatter_block(const uint4 w, int c0, const rm_params& rp, uint32_t nat, bool whole)
{
  const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
  const uint32_t F = rp.nsys - rp.info;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int c = c0 + m;
    if (whole || (unsigned)c < rp.E) {
      uint32_t q;
      if constexpr (LQ >= 0) {
        constexpr int QM = 1 << LQ;
        q = (uint32_t)(c0 >> LQ) + ((uint32_t)(m & (QM - 1)) * rp.Kq + (uint32_t)(m >> LQ) + rp.start);
      } else {
        const int qm = (int)rp.Qm;
        int k = (int)((float)c * (1.0f / (float)qm));
        int jq = c - k * qm;
        if (jq < 0) {
          --k;
          jq += qm;
        } else if (jq >= qm) {
          ++k;
          jq -= qm;
        }
        q = (uint32_t)jq * rp.Kq + (uint32_t)k + rp.start;
        q = min(q, q - rp.P); // wrap (q < 2P)
      }
      *lds_at(nat + rank_pos(q, rp.info, F)) = (int8_t)(ww[m >> 2] >> (8 * (m & 3)));
    }
  }
}

// The scatter load of one codeblock (group), positions [0, n_in) of the HARQ buffer and
// [0, cap) of the soft bits (cap <= n_in, cap a multiple of Z, Z a multiple of 8).  All
// threads of the workgroup take part (barriers); `work` groups load.  PASS_COLS columns
// per row-pair pass: each thread holds two words per round of the pass.
template <bool PAIR>
__device__ __forceinline__ void scatter_load(const rm_params& rp, const int8_t* __restrict__ cw,
                                             int8_t* __restrict__ harq, bool wr, bool work, int8_t* soft,
                                             unsigned r, unsigned NT, int Z, int cap, int n_in, int& last)
{
  int8_t* const nat = soft + 2 * Z; // 16-byte aligned (Z % 8 == 0)
  const uint32_t nat_a = lds_addr(nat);
  // 1. Zeros, and +127 over the filler positions [info, nsys).
  if (work) {
    uint2* n8 = reinterpret_cast<uint2*>(nat);
    const int units = cap >> 3;
    for (int d = (int)r; d < units; d += (int)NT) {
      const uint32_t p0 = 8u * (uint32_t)d;
      uint32_t v[2] = {0u, 0u};
      if (p0 + 8 > rp.info && p0 < rp.nsys) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const uint32_t p = p0 + (uint32_t)k;
          v[k >> 2] |= (p >= rp.info && p < rp.nsys) ? (0x7fu << (8 * (k & 3))) : 0u;
        }
      }
      n8[d] = make_uint2(v[0], v[1]);
    }
  }
  __syncthreads();
  // 2. LLRs to their positions.
  if (work) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(cw);
    const uint4* blk = reinterpret_cast<const uint4*>(a & ~(uintptr_t)15);
    const int lead = (int)(a & 15);
    const int nblk = (lead + (int)rp.E + 15) >> 4;
    const int qm = (int)rp.Qm;
    const bool pow2 = (qm & (qm - 1)) == 0 && (lead & (qm - 1)) == 0 && rp.start + rp.E <= rp.P;
    for (int b = (int)r; b < nblk; b += (int)NT) {
      const uint4 w = blk[b];
      const int c0 = 16 * b - lead;
      const bool whole = c0 >= 0 && c0 + 16 <= (int)rp.E;
      if (pow2) {
        switch (qm) {
        case 1: scatter_block<0>(w, c0, rp, nat_a, whole); break;
        case 2: scatter_block<1>(w, c0, rp, nat_a, whole); break;
        case 4: scatter_block<2>(w, c0, rp, nat_a, whole); break;
        default: scatter_block<3>(w, c0, rp, nat_a, whole); break;
        }
      } else {
        scatter_block<-1>(w, c0, rp, nat_a, whole);
      }
    }
  }
  __syncthreads();
  // 3. HARQ buffer in 8-byte units (zeros from cap on), the last nonzero position, and in
  // the one-row layout the clamp in place (each unit is read and written by one thread).
  if (work) {
    uint2* n8 = reinterpret_cast<uint2*>(nat);
    const int units = (n_in + 7) >> 3;
    const int cap_units = cap >> 3;
    const uintptr_t ha = reinterpret_cast<uintptr_t>(harq);
    for (int d = (int)r; d < units; d += (int)NT) {
      uint2 v = make_uint2(0u, 0u);
      if (d < cap_units) {
        v = n8[d];
        if (v.y != 0u) {
          last = max(last, 8 * d + 4 + ((31 - __builtin_clz(v.y)) >> 3));
        } else if (v.x != 0u) {
          last = max(last, 8 * d + ((31 - __builtin_clz(v.x)) >> 3));
        }
        if (!PAIR) {
          n8[d] = make_uint2(clamp64x4(v.x), clamp64x4(v.y));
        }
      }
      if (wr) {
        const int p = 8 * d;
        if (p + 8 <= n_in && ((ha + (uint32_t)p) & 7) == 0) {
          *reinterpret_cast<uint2*>(harq + p) = v;
        } else if (p + 8 <= n_in && ((ha + (uint32_t)p) & 3) == 0) {
          *reinterpret_cast<uint32_t*>(harq + p) = v.x;
          *reinterpret_cast<uint32_t*>(harq + p + 4) = v.y;
        } else {
          for (int k = 0; k < 8 && p + k < n_in; ++k) {
            harq[p + k] = (int8_t)((k < 4 ? v.x : v.y) >> (8 * (k & 3)));
          }
        }
      }
    }
  }
  if constexpr (PAIR) {
    // 4. Row-pair order, in place: unit t of column c holds rows 4t .. 4t + 3 of both
    // halves, bytes [A0 B0 A1 B1 A2 B2 A3 B3] at c Z + 8t from A = rows 4t.., B = rows
    // h + 4t..  The NT = Z/2 threads cover four columns per round (thread r: column
    // r / (Z/8) of the four, unit r mod (Z/8)).  A pass reads its columns, waits, then
    // writes them (a pass covers 4 kRounds whole columns, so passes share no column).
    constexpr int kRounds = 6;
    const int h = Z >> 1;
    const int upc = Z >> 3; // units per column
    const int ncol = cap / Z;
    const int cr = (int)r / upc;
    const int o = cr * Z + 4 * ((int)r - cr * upc); // byte of A in column 0 of the round
    const uint32_t* n4 = reinterpret_cast<const uint32_t*>(nat);
    for (int c0 = 0; c0 < ncol; c0 += 4 * kRounds) {
      uint32_t A[kRounds], B[kRounds];
#pragma unroll
      for (int k = 0; k < kRounds; ++k) {
        A[k] = B[k] = 0u;
        if (work && c0 + 4 * k + cr < ncol) {
          const int a = (c0 + 4 * k) * Z + o;
          A[k] = n4[a >> 2];
          B[k] = n4[(a + h) >> 2];
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kRounds; ++k) {
        if (work && c0 + 4 * k + cr < ncol) {
          const int a = (c0 + 4 * k) * Z + o;
          const uint32_t lo = clamp64x4(__builtin_amdgcn_perm(B[k], A[k], 0x05010400u));
          const uint32_t hi = clamp64x4(__builtin_amdgcn_perm(B[k], A[k], 0x07030602u));
          // Output bytes c Z + 8t .. (from A at c Z + 4t): offset a + 4t = 2a - c Z.
          *reinterpret_cast<uint2*>(nat + 2 * a - (c0 + 4 * k + cr) * Z) = make_uint2(lo, hi);
        }
      }
    }
  }
}

// Per-group reductions of the persistent decoder: LDS atomics, with a wave-level
// shuffle reduction first when the whole wave belongs to one group.
__device__ __forceinline__ bool wave_in_one_group(int g)
{
  return __shfl(g, 0) == __shfl(g, 63);
}

__device__ __forceinline__ void group_xor_or(uint32_t x, uint32_t z, int g, bool active, uint32_t* xa, uint32_t* za)
{
  if (wave_in_one_group(g)) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      x ^= (uint32_t)__shfl_xor((int)x, o);
      z |= (uint32_t)__shfl_xor((int)z, o);
    }
    if ((threadIdx.x & 63) == 0 && active) {
      atomicXor(&xa[g], x);
      atomicOr(&za[g], z);
    }
  } else if (active) {
    if (x) {
      atomicXor(&xa[g], x);
    }
    if (z) {
      atomicOr(&za[g], z);
    }
  }
}

__device__ __forceinline__ void group_max(int v, int g, bool active, int* ma)
{
  if (wave_in_one_group(g)) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      v = max(v, __shfl_xor(v, o));
    }
    if ((threadIdx.x & 63) == 0 && active) {
      atomicMax(&ma[g], v);
    }
  } else if (active) {
    atomicMax(&ma[g], v);
  }
}

// Per-group control words in LDS (structure of arrays, kMaxGroups each).
struct grp_ctl {
  int last[kMaxGroups];        // last nonzero soft bit
  uint32_t crc[2][kMaxGroups]; // CRC remainder accumulators (iteration parity)
  uint32_t zf[2][kMaxGroups];  // zero-soft-bit / syndrome flags (iteration parity)
  int ndone;                   // groups finished
  int lmax_item;               // largest layer count over the groups
};

constexpr int kCtlBytes = (int)sizeof(grp_ctl);
// The last 16 bytes of the dynamic LDS hold the workgroup's claimed item: with no static
// allocation the soft bits start at LDS address 0, so a soft-bit index is the LDS byte
// address itself (and a ^ 1 its row-pair partner, with no offset to fold).
constexpr int kQueueBytes = 16;

// Decodes one work item (G codeblocks of one (BG, Z)), behaviour of
// ldpc_decoder_impl::decode (lib/phy/upper/channel_coding/ldpc/ldpc_decoder_impl.cpp:41-145)
// with the arithmetic of ldpc_decoder_generic.cpp:12-104 for every codeblock.  With
// kFlagFused the group first rate-dematches its codeblock (writing the HARQ buffer when
// asked) instead of reading a dematched buffer.  With n threads per codeblock (Z, or Z/2
// with PAIR), thread t works on row t mod n (and t mod n + Z/2 with PAIR) of group t / n;
// the threads beyond G*n only take part in barriers.  Both modes use the same LDS
// amounts (only the order of the soft bits in a column differs).
// Inlined into each kernel instance: as a call its register state would spill.
// Returns whether this thread's group wrote a result <= 0 (a codeblock the decoder did not
// pass).
template <int BG, bool LEARNED, bool PAIR, bool EXT8>
__device__ __forceinline__ bool decode_item(const dec_item& item, const dec_job* __restrict__ jobs,
                                            int8_t* __restrict__ rm_base, const int8_t* __restrict__ cw_base,
                                            uint8_t* __restrict__ bits_base, int32_t* __restrict__ results,
                                            const uint32_t* __restrict__ crc_x, const uint32_t* __restrict__ gtab,
                                            const uint32_t* __restrict__ scale_tab, int fb_iters, uint8_t* lds)
{
  using T = bgt<BG>;
  constexpr int K = T::K;
  const int G = item.ngroups;
  const int lmax = item.lmax;
  const int t = (int)threadIdx.x;
  // Item-uniform values in scalar registers: every job of an item has the same (BG, Z)
  // and graph, so the lifted edge table is read with scalar loads.
  const int Z = __builtin_amdgcn_readfirstlane((int)jobs[item.first_job].Z);
  // Threads per codeblock.
  const int NT = PAIR ? Z / 2 : Z;
  const int g = t / NT;
  const bool active = g < G;
  const int gg = active ? g : 0;
  const unsigned r = (unsigned)(t - g * NT);
  dec_job jb = jobs[item.first_job + gg];
  if (fb_iters > 0) {
    // Fallback pass over the jobs of a fused PUSCH pass: the plain decoder on the HARQ
    // buffer that pass wrote, for the codeblocks it did not decode.
    jb.flags = (uint8_t)((jb.flags & kFlagCrcAfter) | kFlagSkipPassed);
    jb.max_iter = (uint8_t)fb_iters;
    jb.in_off = jb.rm.rm_off;
    jb.n_in = jb.rm.N;
  }
  // Fallback pass: codeblocks that an earlier pass decoded are left alone.
  const bool skip = active && (jb.flags & kFlagSkipPassed) && results[jb.out_idx] > 0;
  const bool work = active && !skip; // groups that load (and decode unless `passed`)
  // CRC passed in an earlier transmission (pusch_decoder_impl.cpp:412-431): the new LLRs
  // are combined into the HARQ buffer, the codeblock is not decoded again and its bits
  // are left as they are.
  const bool passed = work && (jb.flags & kFlagPassed);

  // LDS: per-group soft bits | extension-layer state | codeword stage, then control.
  const int soft_stride = ((K + lmax) * Z + 15) & ~15;
  // Extension-layer messages: one compressed word per check row and layer, or with EXT8
  // (one row per thread) an int8 message per edge and row (layer l, edge e of its row
  // list: byte (rs(l) - rs(4) + e) Z + r).
  const int ext_stride = EXT8 ? (((T::rs_word(lmax) - T::rs(4)) * Z + 15) & ~15) : 4 * (lmax - 4) * Z;
  const int stage_stride = (int)item.stage;
  int8_t* const soft0 = reinterpret_cast<int8_t*>(lds); // uniform base of the soft bits
  const int goff = gg * soft_stride;                      // this group's soft bits
  int8_t* soft = soft0 + goff;
  uint8_t* const ldsd = lds;
  uint32_t* ext = reinterpret_cast<uint32_t*>(ldsd + G * soft_stride + gg * ext_stride);
  int8_t* ext8 = reinterpret_cast<int8_t*>(ext);
  int8_t* stage = reinterpret_cast<int8_t*>(ldsd + G * soft_stride + G * ext_stride) + gg * stage_stride;
  grp_ctl* ctl = reinterpret_cast<grp_ctl*>(ldsd + G * soft_stride + G * ext_stride + G * stage_stride);

  if (t < G) {
    ctl->last[t] = -1;
    ctl->crc[0][t] = ctl->crc[1][t] = 0;
    ctl->zf[0][t] = ctl->zf[1][t] = 0;
  }
  if (t == 0) {
    ctl->ndone = 0;
    ctl->lmax_item = 4;
  }
  __syncthreads();

  // Soft-bit load (load_soft_bits, ldpc_decoder_impl.cpp:147-188): two punctured nodes at
  // zero, full nodes clamped to [-64, 64], a partial last node unclamped (infinity as
  // +-121); track the last nonzero input (:67-71).
  const int n_in = (int)jb.n_in;
  const int full = (n_in / Z) * Z;
  const int cap = (K + lmax - 2) * Z;
  int last = -1;
  if (item.scatter != 0 && fb_iters == 0) {
    // New data, at most one LLR per position: scatter load (uniform per item).
    scatter_load<PAIR>(jb.rm, cw_base + jb.rm.cw_off, rm_base + jb.rm.rm_off, (jb.flags & kFlagHarqWrite) != 0, work,
                       soft, r, (unsigned)NT, Z, cap, n_in, last);
  } else if (jb.flags & kFlagFused) {
    const rm_params rp = jb.rm;
    const int8_t* in = cw_base + rp.cw_off;
    int8_t* harq = rm_base + rp.rm_off;
    const bool new_data = jb.flags & kFlagNewData;
    const bool wr = jb.flags & kFlagHarqWrite;
    // Stage the E codeword LLRs in LDS with aligned 16-byte loads (a 16-byte block that
    // holds a byte of the codeblock lies in that byte's page).
    const bool staged = stage_stride > 0;
    if (staged && work) {
      const uintptr_t a = reinterpret_cast<uintptr_t>(in);
      const uint4* blk = reinterpret_cast<const uint4*>(a & ~(uintptr_t)15);
      const int lead = (int)(a & 15);
      const int nblk = (lead + (int)rp.E + 15) >> 4;
      // The stage of a group starts at a multiple of 4 bytes: with a 16-byte aligned
      // codeblock, whole blocks go in as four dwords.
      const bool whole = lead == 0;
      for (int b = (int)r; b < nblk; b += NT) {
        const uint4 w = blk[b];
        if (whole && 16 * b + 16 <= (int)rp.E) {
          uint32_t* st4 = reinterpret_cast<uint32_t*>(stage + 16 * b);
          st4[0] = w.x;
          st4[1] = w.y;
          st4[2] = w.z;
          st4[3] = w.w;
          continue;
        }
        const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const int idx = 16 * b + k - lead;
          if (idx >= 0 && idx < (int)rp.E) {
            stage[idx] = (int8_t)(ww[k >> 2] >> (8 * (k & 3)));
          }
        }
      }
    }
    if (staged) {
      __syncthreads();
    }
    if (work) {
      if (staged) {
        fused_positions<PAIR>(rp, stage, harq, new_data, wr, n_in, cap, soft + 2 * Z, r, (unsigned)NT, Z, last);
      } else {
        fused_positions<PAIR>(rp, in, harq, new_data, wr, n_in, cap, soft + 2 * Z, r, (unsigned)NT, Z, last);
      }
    }
  } else if (work) {
    // Aligned 4-byte loads of the dematched soft bits.
    const int8_t* in = rm_base + jb.in_off;
    const uintptr_t a = reinterpret_cast<uintptr_t>(in);
    const uint32_t* blk = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
    const int lead = (int)(a & 3);
    const int nblk = (lead + n_in + 3) >> 2;
    // Row of the first position of a word (the same in every pass: the stride 4 NT is a
    // multiple of Z).
    const int jb = PAIR ? ((4 * (int)r - lead) % Z + Z) % Z : 0;
    // kLoadBatch words per thread in flight together.
    constexpr int kLoadBatch = 8;
    for (int bb = (int)r; bb < nblk; bb += kLoadBatch * NT) {
      uint32_t wv[kLoadBatch];
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u) {
        const int b = bb + u * NT;
        wv[u] = b < nblk ? blk[b] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u) {
        const int b = bb + u * NT;
        if (b < nblk) {
          const uint32_t w = wv[u];
          int j = jb;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int p = 4 * b + k - lead;
            const int x = (int)(int8_t)(w >> (8 * k));
            if (p >= 0 && p < n_in) {
              last = x != 0 ? p : last;
              if (p < cap) {
                soft[soft_pos<PAIR>(2 * Z + p, j, Z)] =
                    (int8_t)(p < full ? min(max(x, -64), 64) : med3i(x, -121, 121));
              }
            }
            if (PAIR) {
              j = j + 1 == Z ? 0 : j + 1;
            }
          }
        }
      }
    }
  }
  if (work) {
    for (int i = (int)r; i < 2 * Z; i += NT) {
      soft[i] = 0;
    }
    int j = PAIR ? (n_in + (int)r) % Z : 0;
    for (int i = n_in + (int)r; i < cap; i += NT) {
      soft[soft_pos<PAIR>(2 * Z + i, j, Z)] = 0;
      if (PAIR) {
        j = j + NT >= Z ? j + NT - Z : j + NT;
      }
    }
  }
  group_max(last, g, work, ctl->last);
  __syncthreads();
  const int input_size = ctl->last[gg] + 1;

  // Per-group decode state.
  const int msg = K * Z;
  uint8_t* out = bits_base + jb.bits_off;
  const bool use_crc = jb.crc_idx != 0;
  const bool es_crc = use_crc && !(jb.flags & kFlagCrcAfter);
  const bool es_syn = !use_crc && (jb.flags & kFlagEsSyndrome);
  const int max_iter = jb.max_iter;
  int result = 0;
  bool done = !active;
  int L = 4;
  if (active) {
    if (skip) {
      done = true;
    } else if (passed) {
      done = true;
      result = (int)jb.preset;
    } else if (input_size < msg && !(jb.flags & kFlagForce)) {
      // Without a CRC calculator the output is set to ones (ldpc_decoder_impl.cpp:75-81).
      if (jb.crc_idx == 0 || (jb.flags & kFlagCrcAfter)) {
        write_hard_bits<PAIR>(soft, out, msg, true, (int)r, NT, Z);
      }
      done = true;
      result = 0;
    } else {
      int cbl = max(input_size + 2 * Z, msg + 4 * Z);
      cbl = (cbl + Z - 1) / Z * Z;
      L = cbl / Z - K;
      if (L > lmax) { // host bound violated: refuse rather than overrun LDS
        done = true;
        result = -1;
      }
    }
  }
  if (active && !done) {
    if (PAIR) {
      // One state word per check row: thread r owns words 2r, 2r + 1 of each layer.
      for (int l = 4; l < L; ++l) {
        ext[(l - 4) * Z + 2 * (int)r] = 0;
        ext[(l - 4) * Z + 2 * (int)r + 1] = 0;
      }
    } else if (!EXT8) {
      for (int l = 4; l < L; ++l) {
        ext[(l - 4) * Z + (int)r] = 0;
      }
    } else {
      // The int8 messages of layers 4 .. L - 1, zeroed as words (the region is 16-byte
      // aligned and padded to 16 bytes).
      const int nw = ((T::rs_word(L) - T::rs(4)) * Z + 3) >> 2;
      for (int i = (int)r; i < nw; i += NT) {
        ext[i] = 0;
      }
    }
  }
  if (active && r == 0) {
    if (done) {
      atomicAdd(&ctl->ndone, 1);
    } else {
      atomicMax(&ctl->lmax_item, L);
    }
  }

  // Core-layer check-to-variable messages, int8 packed four per word (VGPRs).
  // Words per core row: four int8 messages per word, or two edges of the row pair.
  constexpr int EPW = PAIR ? 2 : 4;
  constexpr int CW0 = (bgt<BG>::rs(1) - bgt<BG>::rs(0) + EPW - 1) / EPW;
  constexpr int CW1 = (bgt<BG>::rs(2) - bgt<BG>::rs(1) + EPW - 1) / EPW;
  constexpr int CW2 = (bgt<BG>::rs(3) - bgt<BG>::rs(2) + EPW - 1) / EPW;
  constexpr int CW3 = (bgt<BG>::rs(4) - bgt<BG>::rs(3) + EPW - 1) / EPW;
  uint32_t c0[CW0], c1[CW1], c2[CW2], c3[CW3];
#pragma unroll
  for (int k = 0; k < CW0; ++k) c0[k] = 0;
#pragma unroll
  for (int k = 0; k < CW1; ++k) c1[k] = 0;
#pragma unroll
  for (int k = 0; k < CW2; ++k) c2[k] = 0;
#pragma unroll
  for (int k = 0; k < CW3; ++k) c3[k] = 0;

  // CRC weights: hard bit i of the message contributes x^(len-1-i+order) mod G; thread r
  // reads the weights of its bits n*Z + rr (n < K) from the table when it checks the CRC
  // (index clamped at 0 for bits beyond len, whose weights are masked off).
  const int len = msg - (int)jb.F;
  const int order = jb.crc_idx == 3 ? 16 : 24;
  const uint32_t* __restrict__ xbase = crc_x + (use_crc ? (jb.crc_idx - 1) * kCrcXZLen : 0) + kCrcPad;
  const int xi0 = (len - 1 + order) - (int)r; // weight index of bit r
  __syncthreads();
  const int Litem = ctl->lmax_item;

  for (int it = 0;; ++it) {
    const bool run = !done;
    const int par = it & 1;
    // The edge table is re-read every iteration: values derived from it stay in the
    // layer that uses them instead of being hoisted out of the loop into registers.
    // Two words per edge, ready for the address arithmetic: the row-pair table, or the
    // one-row table (shift, column * Z) behind it.
    int gto = (PAIR ? kGraphTableLen : 3 * kGraphTableLen) + 2 * (int)jb.gt_off;
    asm volatile("" : "+v"(gto));
    const uint32_t* __restrict__ gti = gtab + __builtin_amdgcn_readfirstlane(gto);
    // Learned tables: one row of M (A, B) words per iteration, the last row repeated.
    const uint32_t* abt = LEARNED ? scale_tab + min(it, kLearnedIters - 1) * T::M : nullptr;
    if (run) {
      core_layer<BG, 0, LEARNED, PAIR>(soft0, goff, gti, r, (unsigned)Z, c0, abt);
    }
    __syncthreads();
    // Uniform exit: every group finished.  The count is read after this barrier, which
    // orders the previous check's updates; a finished launch reaches it with layer 0
    // skipped by every group.  The check therefore needs no closing barrier of its own.
    if (ctl->ndone >= G) {
      break;
    }
    if (run) {
      core_layer<BG, 1, LEARNED, PAIR>(soft0, goff, gti, r, (unsigned)Z, c1, abt);
    }
    __syncthreads();
    if (run) {
      core_layer<BG, 2, LEARNED, PAIR>(soft0, goff, gti, r, (unsigned)Z, c2, abt);
    }
    __syncthreads();
    if (run) {
      core_layer<BG, 3, LEARNED, PAIR>(soft0, goff, gti, r, (unsigned)Z, c3, abt);
    }
    __syncthreads();
    // Row starts of the extension layers: the next layer's end is read one layer ahead,
    // off the critical path of the layer chain.
    int e_cur = T::rs(4), e_nxt = T::rs(5);
    for (int l = 4; l < Litem; ++l) {
      const int e_after = T::rs_word(min(l + 2, T::M));
      if (run && l < L) {
        const int e0 = e_cur, dg = e_nxt - e_cur;
        const uint32_t ab = LEARNED ? abt[l] : 0u;
        uint32_t* st = &ext[(l - 4) * Z + (PAIR ? 2 * (int)r : (int)r)]; // compressed words
        int8_t* msg = ext8 + (e0 - T::rs(4)) * Z + (int)r;                 // EXT8
        // Degree classes (uniform branch): the unrolled update does the work of its
        // class's largest degree, so small rows take a small instance.
        if constexpr (PAIR) {
          if (BG == 1) {
            if (dg <= 5) {
              layer_update_pair<5, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, st, ab);
            } else if (dg <= 7) {
              layer_update_pair<7, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, st, ab);
            } else {
              layer_update_pair<10, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, st, ab);
            }
          } else {
            if (dg <= 4) {
              layer_update_pair<4, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, st, ab);
            } else {
              layer_update_pair<6, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, st, ab);
            }
          }
        } else if constexpr (EXT8) {
          if (BG == 1) {
            if (dg <= 5) {
              layer_update_ext8<5, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, msg, ab);
            } else if (dg <= 7) {
              layer_update_ext8<7, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, msg, ab);
            } else {
              layer_update_ext8<10, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, msg, ab);
            }
          } else {
            if (dg <= 4) {
              layer_update_ext8<4, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, msg, ab);
            } else {
              layer_update_ext8<6, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, msg, ab);
            }
          }
        } else if (BG == 1) {
          if (dg <= 5) {
            layer_update<5, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, st, ab);
          } else if (dg <= 7) {
            layer_update<7, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, st, ab);
          } else {
            layer_update<10, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, st, ab);
          }
        } else {
          if (dg <= 4) {
            layer_update<4, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, st, ab);
          } else {
            layer_update<6, true, false, LEARNED>(soft0, goff, gti + 2 * e0, dg, r, (unsigned)Z, st, ab);
          }
        }
      }
      __syncthreads();
      e_cur = e_nxt;
      e_nxt = e_after;
    }

    // Per-iteration check: CRC (ldpc_decoder_impl.cpp:112-119), syndrome (:120-127), or
    // the final syndrome (+ CRC for PUSCH without early stop) after the last iteration.
    const bool last_it = it + 1 >= max_iter;
    const bool check = run && (es_crc || es_syn || last_it);
    uint32_t x = 0, z = 0;
    if (check) {
      const bool want_crc = use_crc && (es_crc || last_it);
      // Check rows of this thread: r, and r + Z/2 with row pairs; row rr is byte rr of a
      // column, or byte 2r + half with row pairs.  The weight of bit n Z + rr is entry
      // xi - n Z of the zero-padded table, zero for filler bits (no bounds test).
#pragma unroll 1
      for (int half = 0; half < (PAIR ? 2 : 1); ++half) {
        const int rr = (int)r + half * NT;
        // Opaque base: the weight addresses are formed at each check instead of being
        // hoisted out of the iteration loop into registers.
        const uint32_t* xt0 = xbase + (xi0 - half * NT);
        asm volatile("" : "+v"(xt0));
        gbl_u32* xt = gbl_at(xt0);
        const int8_t* sc = soft0 + goff + (PAIR ? 2 * (int)r + half : rr);
        uint32_t acc = 0, zz = 0;
        if (want_crc) {
          // The weights in batches of eight loads, all in flight before the batch's one
          // wait (scheduled one by one, each load would wait for the one before).
#pragma unroll
          for (int n0 = 0; n0 < K; n0 += 8) {
            uint32_t wv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              wv[k] = n0 + k < K ? xt[-(n0 + k) * Z] : 0u;
            }
            int sv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              sv[k] = n0 + k < K ? sc[(n0 + k) * Z] : 1;
            }
            hold8(wv);
            hold8(sv);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              if (n0 + k < K) {
                acc ^= sv[k] <= 0 ? wv[k] : 0u; // weight of a hard one
                zz |= (uint32_t)(sv[k] == 0);
              }
            }
          }
        } else {
#pragma unroll
          for (int n = 0; n < K; ++n) {
            zz |= (uint32_t)(sc[n * Z] == 0);
          }
        }
        x ^= acc;
        z |= zz;
      }
      if (!es_crc) {
        if (PAIR) {
          z |= syndrome_fail_pair<BG>(soft0, goff, gti, r, L);
        } else {
          z |= syndrome_fail<BG>(soft0, goff, gti, r, (unsigned)Z, L);
        }
      }
    }
    group_xor_or(x, z, g, check, ctl->crc[par], ctl->zf[par]);
    __syncthreads();
    if (run) {
      const uint32_t xr = ctl->crc[par][g], zr = ctl->zf[par][g];
      bool fin = false;
      if (es_crc) {
        if (zr == 0 && xr == 0) {
          result = it + 1;
          fin = true;
        }
      } else if (es_syn) {
        if (zr == 0) {
          result = it + 1;
          fin = true;
        }
      } else if (last_it) {
        result = (zr == 0 && (!use_crc || xr == 0)) ? max_iter : 0;
        fin = true;
      }
      if (!fin && last_it) {
        result = 0;
        fin = true;
      }
      if (fin) {
        done = true;
        if (r == 0) {
          atomicAdd(&ctl->ndone, 1);
        }
      }
    }
    // Clear the other parity's accumulators for the next iteration (its check follows four
    // layer barriers later; this iteration's are read before the next layer-0 barrier).
    if (t < G) {
      ctl->crc[par ^ 1][t] = 0;
      ctl->zf[par ^ 1][t] = 0;
    }
  }

  if (active && !skip) {
    if (result >= 0 && !passed && !(input_size < msg && !(jb.flags & kFlagForce))) {
      write_hard_bits<PAIR>(soft, out, msg, false, (int)r, NT, Z);
    }
    if (r == 0) {
      results[jb.out_idx] = result;
    }
  }
  __syncthreads(); // the next item reuses the LDS
  return active && !skip && result <= 0;
}

template <bool SIGNAL>
__device__ __forceinline__ void tb_finalize_signal(uint32_t tb, uint32_t n_tbs, const lh_tb_desc_dev* __restrict__ tbs,
                                                   const uint8_t* __restrict__ cb_bits,
                                                   const int32_t* __restrict__ cb_res, uint8_t* __restrict__ tb_out,
                                                   int32_t* __restrict__ status, const uint32_t* __restrict__ lut24a,
                                                   const uint32_t* __restrict__ pow2_24a, tb_signal sig,
                                                   uint32_t* red, uint32_t* lut, int& all_ok);

// Fused TB finalize (the serving runtime; its arguments at fap, in device memory): the
// item's codeblocks are counted off their transport blocks (one thread per group; its
// device-scope fence first makes the workgroup's bits and results, complete at the item's
// closing barrier, visible to every workgroup); the transport blocks whose last codeblock
// this workgroup counted are then finalized here, after a fence that makes the other
// workgroups' codeblocks visible.  Scratch: the start of the dynamic LDS, free between
// items.
__device__ __forceinline__ void fused_tb_complete(const dec_item& item, const dec_job* __restrict__ jobs,
                                                  const fused_tb_args* __restrict__ fap, const uint8_t* __restrict__ bits,
                                                  const int32_t* __restrict__ results, uint8_t* lds)
{
  const fused_tb_args fa = *fap;
  uint32_t* lut = reinterpret_cast<uint32_t*>(lds);
  uint32_t* red = lut + 352;
  int& all_ok = *reinterpret_cast<int*>(red + 32);
  uint32_t& n_done = red[33];
  uint32_t* done = red + 34;
  const int t = (int)threadIdx.x;
  if (t == 0) {
    n_done = 0;
  }
  __syncthreads();
  if (t < (int)item.ngroups) {
    const uint32_t tb = fa.cb_tb[jobs[item.first_job + t].out_idx];
    __threadfence();
    if (atomicSub(&fa.tb_left[tb], 1u) == 1u) {
      done[atomicAdd(&n_done, 1u)] = tb;
    }
  }
  __syncthreads();
  const uint32_t n = n_done;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t tb = done[k];
    __threadfence();
    tb_finalize_signal<true>(tb, fa.n_tbs, fa.tbs, bits, results, fa.tb_out, fa.status, fa.lut24a, fa.lut24a + 256,
                             fa.sig, red, lut, all_ok);
    __syncthreads(); // the scratch serves the next transport block
  }
}

// Persistent decoder.  The host sorts the items largest first.  Workgroup w takes item w
// (the first round needs no claim), then claims items of its sub-queue w mod nq through
// that sub-queue's counter: the claims of a launch spread over nq device-scope counters
// in separate cache lines instead of queueing on one, and each sub-queue still balances
// uneven items (and early-stopping codeblocks) dynamically.  The claimed index goes to
// the workgroup through a word at the end of the dynamic LDS (qslot).
// One persistent decoder per row mode: the row-pair instance holds more state per thread
// and runs at three waves per SIMD (168 VGPRs), the one-row instance at four.
template <bool LEARNED, bool PAIR, bool EXT8>
__global__ void __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(PAIR ? 3 : 4)))
k_ldpc_decode(launch_header* __restrict__ hdr, const dec_item* __restrict__ items, const dec_job* __restrict__ jobs,
              int8_t* __restrict__ rm_base, const int8_t* __restrict__ cw_base, uint8_t* __restrict__ bits_base,
              int32_t* __restrict__ results, const uint32_t* __restrict__ crc_x, const uint32_t* __restrict__ gtab,
              const uint32_t* __restrict__ scale_tab, int fb_iters, unsigned qslot,
              const fused_tb_args* __restrict__ fa)
{
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t& claimed = *reinterpret_cast<uint32_t*>(lds + qslot);
  const uint32_t n_items = hdr->n_items, nq = hdr->nq;
  const uint32_t W = gridDim.x, w = blockIdx.x, q = nq ? w % nq : 0u;
  uint32_t* ctr = reinterpret_cast<uint32_t*>(hdr) + hdr->ctr_off + q * kCounterStride;
  for (uint32_t id = w; id < n_items;) {
    const dec_item item = items[hdr->first_item + id];
    // LEARNED with fb_iters > 0: the learned decoder, then at once in this workgroup the
    // plain decoder (fb_iters iterations, on the HARQ buffer the learned pass wrote) over
    // the item's codeblocks whose result is <= 0 -- what a separate fallback launch over
    // the same descriptors does, without waiting for the launch's other items.
    bool failed;
    if (item.bg == 1) {
      failed = decode_item<1, LEARNED, PAIR, EXT8>(item, jobs, rm_base, cw_base, bits_base, results, crc_x, gtab,
                                                   scale_tab, LEARNED ? 0 : fb_iters, lds);
    } else {
      failed = decode_item<2, LEARNED, PAIR, EXT8>(item, jobs, rm_base, cw_base, bits_base, results, crc_x, gtab,
                                                   LEARNED ? scale_tab + kLearnedIters * 46 : nullptr,
                                                   LEARNED ? 0 : fb_iters, lds);
    }
    if constexpr (LEARNED) {
      // The results the fallback's skip test reads were written before decode_item's
      // closing barrier; the OR is itself a barrier of the whole workgroup.
      if (fb_iters > 0 && __syncthreads_or(failed ? 1 : 0) != 0) {
        if (item.bg == 1) {
          decode_item<1, false, PAIR, EXT8>(item, jobs, rm_base, cw_base, bits_base, results, crc_x, gtab, nullptr,
                                            fb_iters, lds);
        } else {
          decode_item<2, false, PAIR, EXT8>(item, jobs, rm_base, cw_base, bits_base, results, crc_x, gtab, nullptr,
                                            fb_iters, lds);
        }
      }
    }
    if (fa != nullptr) {
      fused_tb_complete(item, jobs, fa, bits_base, results, lds);
    }
    if (nq == 0) {
      break;
    }
    if (threadIdx.x == 0) {
      claimed = atomicAdd(ctr, 1u);
    }
    __syncthreads();
    id = W + q + nq * claimed;
    __syncthreads();
  }
}

// Standalone rate dematcher (ldpc_rate_dematcher::rate_dematch).  Grid: x = position
// chunks, y = codeblock.
__global__ void __launch_bounds__(256) k_rate_dematch(const rm_params* __restrict__ jobs,
                                                      const int8_t* __restrict__ cw, int8_t* __restrict__ rm)
{
  const rm_params rp = jobs[blockIdx.y];
  const int8_t* in = cw + rp.cw_off;
  int8_t* out = rm + rp.rm_off;
  const bool new_data = rp.flags & kFlagNewData;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < rp.N; p += gridDim.x * blockDim.x) {
    const int old = new_data ? 0 : out[p];
    out[p] = (int8_t)dematch_at(rp, in, old, p, new_data);
  }
}

// CRC remainders of packed bit buffers, one workgroup per buffer.
__global__ void __launch_bounds__(256) k_crc(const uint64_t* __restrict__ offs, const uint32_t* __restrict__ nbits,
                                             const uint8_t* __restrict__ bits, const uint32_t* __restrict__ xtab,
                                             int order, uint32_t* __restrict__ out)
{
  __shared__ uint32_t red[32];
  const uint8_t* b = bits + offs[blockIdx.x];
  const int len = (int)nbits[blockIdx.x];
  uint32_t x = 0, z = 0;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    if ((b[i >> 3] >> (7 - (i & 7))) & 1) {
      x ^= xtab[len - 1 - i + order];
    }
  }
  block_reduce_xor_or(x, z, red);
  if (threadIdx.x == 0) {
    out[blockIdx.x] = x;
  }
}

// Byte of the concatenated TB data whose first bit is bit j of codeblock c (MSB-first
// bytes, codeblocks `stride` bytes apart, D data bits each); the bits past D come from the
// next codeblock.  Positions past the TB's last codeblock read its CRC bits, which the
// caller discards.
__device__ __forceinline__ uint32_t tb_byte(const uint8_t* __restrict__ first, uint32_t stride, uint32_t D,
                                            uint32_t c, uint32_t j)
{
  const uint8_t* src = first + (size_t)c * stride;
  if (j + 8u <= D) {
    // The eight bits lie in one codeblock: bits j..j+7 of its MSB-first bytes (the byte
    // after is inside the codeblock's message, whose CRC follows the data).
    const uint32_t q = j >> 3, sh = j & 7u;
    return ((((uint32_t)src[q] << 8) | src[q + 1]) >> (8u - sh)) & 0xffu;
  }
  uint32_t v = 0;
  for (int k = 0; k < 8; ++k) {
    const uint8_t* sk = first + (size_t)c * stride;
    v |= (uint32_t)((sk[j >> 3] >> (7 - (j & 7))) & 1u) << (7 - k);
    if (++j == D) {
      j = 0;
      ++c;
    }
  }
  return v;
}

// TB reassembly + TB CRC24A of transport block d, one workgroup (pusch_decoder_impl.cpp:612-781).
// WIDE: the TB bytes go out as aligned 4-byte words (the TB offset is a multiple of 4),
// which matters when they are written across the bus into host memory.
template <bool WIDE>
__device__ __forceinline__ void tb_finalize_one(const lh_tb_desc_dev& d, uint32_t tb, const uint8_t* __restrict__ cb_bits,
                                                const int32_t* __restrict__ cb_res, uint8_t* __restrict__ tb_out,
                                                int32_t* __restrict__ status, const uint32_t* __restrict__ lut24a,
                                                const uint32_t* __restrict__ pow2_24a, uint32_t* red, uint32_t* lut,
                                                int& all_ok)
{
  const int t = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int i = t; i < 256 + 16 * kTbShiftDigits; i += nt) {
    lut[i] = i < 256 ? lut24a[i] : pow2_24a[32 + i - 256];
  }
  if (t == 0) {
    all_ok = 1;
  }
  __syncthreads();
  for (uint32_t c = (uint32_t)t; c < d.nof_cbs; c += (uint32_t)nt) {
    if (cb_res[d.first_cb + c] <= 0) {
      all_ok = 0;
    }
  }
  __syncthreads();
  uint8_t* out = tb_out + d.tb_offset;
  const uint32_t nbytes = d.tbs_bits / 8;
  const uint8_t* first = cb_bits + d.cb_bits_offset;
  if (d.nof_cbs == 1) {
    // One codeblock: its CRC is the TB CRC; copy the data on success.
    if (all_ok) {
      if (WIDE) {
        uint32_t* out4 = reinterpret_cast<uint32_t*>(out);
        for (uint32_t w = (uint32_t)t; 4 * w < nbytes; w += (uint32_t)nt) {
          if (4 * w + 4 <= nbytes) {
            out4[w] = (uint32_t)first[4 * w] | ((uint32_t)first[4 * w + 1] << 8) | ((uint32_t)first[4 * w + 2] << 16) |
                      ((uint32_t)first[4 * w + 3] << 24);
          } else {
            for (uint32_t b = 4 * w; b < nbytes; ++b) {
              out[b] = first[b];
            }
          }
        }
      } else {
        for (uint32_t b = (uint32_t)t; b < nbytes; b += (uint32_t)nt) {
          out[b] = first[b];
        }
      }
    }
    if (t == 0) {
      status[tb] = all_ok ? 1 : 0;
    }
    return;
  }
  if (!all_ok) {
    if (t == 0) {
      status[tb] = 0;
    }
    return;
  }
  // Concatenation: TB bit i comes from codeblock i / D, bit i mod D.  Pass 1 assembles the
  // TB with consecutive threads on consecutive words, so each wave store is one contiguous
  // run (which matters when the output lies across the bus in host memory).  Pass 2 folds
  // each thread's contiguous chunk of TB bytes into its CRC24A remainder (table driven;
  // its codeblock reads hit the cache lines pass 1 brought in); the remainders are moved
  // to their place by x^(8 * bytes after the chunk) and XOR-reduced.
  const uint32_t D = d.cb_data_bits;
  const uint32_t poly = 0x1864cfbu;
  const uint32_t stride = d.cb_bits_stride;
  {
    // Word w covers TB bits 32w..32w+31; a step of nt words moves (c, j) by (dc, dj).
    const uint32_t nwords = WIDE ? (nbytes + 3u) / 4u : 0u;
    const uint32_t step = 32u * (uint32_t)nt, dc = step / D, dj = step - dc * D;
    uint32_t c = (32u * (uint32_t)t) / D, j = 32u * (uint32_t)t - c * D;
    if (WIDE) {
      uint32_t* out4 = reinterpret_cast<uint32_t*>(out);
      for (uint32_t w = (uint32_t)t; w < nwords; w += (uint32_t)nt) {
        uint32_t cc = c, jj = j, word = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          word |= tb_byte(first, stride, D, cc, jj) << (8 * k);
          jj += 8u;
          if (jj >= D) {
            jj -= D;
            ++cc;
          }
        }
        if (4u * w + 4u <= nbytes) {
          out4[w] = word;
        } else {
          for (uint32_t b = 4u * w; b < nbytes; ++b) {
            out[b] = (uint8_t)(word >> (8 * (b & 3u)));
          }
        }
        c += dc;
        j += dj;
        if (j >= D) {
          j -= D;
          ++c;
        }
      }
    } else {
      const uint32_t bstep = 8u * (uint32_t)nt, bdc = bstep / D, bdj = bstep - bdc * D;
      c = (8u * (uint32_t)t) / D;
      j = 8u * (uint32_t)t - c * D;
      for (uint32_t b = (uint32_t)t; b < nbytes; b += (uint32_t)nt) {
        out[b] = (uint8_t)tb_byte(first, stride, D, c, j);
        c += bdc;
        j += bdj;
        if (j >= D) {
          j -= D;
          ++c;
        }
      }
    }
  }
  const uint32_t chunk = (nbytes + (uint32_t)nt - 1) / (uint32_t)nt;
  const uint32_t a0 = min(nbytes, (uint32_t)t * chunk), a1 = min(nbytes, a0 + chunk);
  uint32_t rem = 0;
  // Codeblock c and bit j of TB bit 8 * b (one division per thread, then stepped); four
  // bytes' loads are issued before their four table steps.
  uint32_t c = (8u * a0) / D, j = 8u * a0 - c * D;
  for (uint32_t b = a0; b < a1; b += 4u) {
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = b + (uint32_t)k < a1 ? tb_byte(first, stride, D, c, j) : 0u;
      j += 8u;
      if (j >= D) {
        j -= D;
        ++c;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (b + (uint32_t)k < a1) {
        rem = ((rem << 8) & 0xffffffu) ^ lut[((rem >> 16) ^ v[k]) & 0xffu];
      }
    }
  }
  uint32_t x = 0, z = 0;
  if (a0 < a1) {
    x = gf_shift_bytes(rem, nbytes - a1, lut + 256, poly);
  }
  block_reduce_xor_or(x, z, red);
  if (t == 0) {
    // The 24-bit TB checksum follows the last codeblock's data bits (:764-769): bits
    // start..start+23 of its MSB-first bytes, read as one 32-bit window (four independent
    // loads; the window's last bits lie inside the codeblock's message, whose CRC follows).
    const uint32_t last_c = d.nof_cbs - 1;
    const uint32_t start = d.tbs_bits - last_c * D;
    const uint8_t* src = first + (size_t)last_c * d.cb_bits_stride + (start >> 3);
    const uint32_t win = ((uint32_t)src[0] << 24) | ((uint32_t)src[1] << 16) | ((uint32_t)src[2] << 8) | src[3];
    const uint32_t ck = (win >> (8u - (start & 7u))) & 0xffffffu;
    status[tb] = (x == ck) ? 1 : -1;
  }
}

// Threads of a TB workgroup: a large TB's concatenation and CRC chunks spread over 16 waves.
constexpr int kTbThreads = 1024;

// TB finalize of transport block tb of n_tbs (one workgroup).  SIGNAL: the outputs are
// host-visible; the codeblock results go to sig.res_out as well, and the last of the
// n_tbs to finish stores sig.seq into *sig.flag (no copy back, no event on the latency
// path).  red (32 words), lut (256 + 96 words) and all_ok are workgroup scratch.
template <bool SIGNAL>
__device__ __forceinline__ void tb_finalize_signal(uint32_t tb, uint32_t n_tbs, const lh_tb_desc_dev* __restrict__ tbs,
                                                   const uint8_t* __restrict__ cb_bits,
                                                   const int32_t* __restrict__ cb_res, uint8_t* __restrict__ tb_out,
                                                   int32_t* __restrict__ status, const uint32_t* __restrict__ lut24a,
                                                   const uint32_t* __restrict__ pow2_24a, tb_signal sig,
                                                   uint32_t* red, uint32_t* lut, int& all_ok)
{
  const lh_tb_desc_dev d = tbs[tb];
  tb_finalize_one<SIGNAL>(d, tb, cb_bits, cb_res, tb_out, status, lut24a, pow2_24a, red, lut, all_ok);
  if constexpr (SIGNAL) {
    for (uint32_t c = threadIdx.x; c < d.nof_cbs; c += blockDim.x) {
      sig.res_out[d.first_cb + c] = cb_res[d.first_cb + c];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t before = atomicAdd(sig.counter, 1u);
      if (before == n_tbs - 1) {
        // Every transport block's writes are fenced before its increment: the launch is whole.
        *sig.counter = 0;
        __threadfence_system();
        __hip_atomic_store(sig.flag, sig.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

// One workgroup per transport block.
template <bool SIGNAL>
__global__ void __launch_bounds__(kTbThreads) k_tb_finalize(const lh_tb_desc_dev* __restrict__ tbs,
                                                     const uint8_t* __restrict__ cb_bits,
                                                     const int32_t* __restrict__ cb_res, uint8_t* __restrict__ tb_out,
                                                     int32_t* __restrict__ status, const uint32_t* __restrict__ lut24a,
                                                     const uint32_t* __restrict__ pow2_24a, tb_signal sig)
{
  __shared__ uint32_t red[32];
  __shared__ uint32_t lut[256 + 16 * kTbShiftDigits];
  __shared__ int all_ok;
  tb_finalize_signal<SIGNAL>(blockIdx.x, gridDim.x, tbs, cb_bits, cb_res, tb_out, status, lut24a, pow2_24a, sig, red,
                             lut, all_ok);
}

// ---- host-side launchers ------------------------------------------------------------
size_t decode_item_lds_bytes(int bg, int Z, int lmax, int G, int stage, bool ext8)
{
  const int K = bg == 1 ? 22 : 10;
  const int ext_edges = bg == 1 ? bgt<1>::rs(lmax) - bgt<1>::rs(4) : bgt<2>::rs(lmax) - bgt<2>::rs(4);
  const int ext = ext8 ? ((ext_edges * Z + 15) & ~15) : 4 * (lmax - 4) * Z;
  return kQueueBytes + (size_t)G * ((((K + lmax) * Z + 15) & ~15) + ext + stage) + kCtlBytes;
}

// The kernel instance of a launch: plain or learned, row pairs or one row, and in the
// one-row mode the int8 extension-layer messages (EXT8) or the compressed words.
const void* decode_kernel(bool learned, bool pairs, bool ext8)
{
  if (pairs) {
    return learned ? (const void*)k_ldpc_decode<true, true, false> : (const void*)k_ldpc_decode<false, true, false>;
  }
  if (ext8) {
    return learned ? (const void*)k_ldpc_decode<true, false, true> : (const void*)k_ldpc_decode<false, false, true>;
  }
  return learned ? (const void*)k_ldpc_decode<true, false, false> : (const void*)k_ldpc_decode<false, false, false>;
}

hipError_t launch_decode(launch_header* d_hdr, const dec_item* d_items, const dec_job* d_jobs, unsigned grid,
                         int nt, size_t lds, int8_t* rm, const int8_t* cw, uint8_t* bits, int32_t* results,
                         const uint32_t* crc_x, const uint32_t* gtab, const uint32_t* scale_tab, int fb_iters,
                         bool pairs, bool ext8, hipStream_t s, const fused_tb_args* fa)
{
  const bool learned = scale_tab != nullptr;
  ext8 = ext8 && !pairs;
  const void* fn = decode_kernel(learned, pairs, ext8);
  // The dynamic-LDS limit of each kernel instance per device only ever rises: a launch
  // within the limit already set makes no runtime call.
  static std::atomic<int> lds_limit[kMaxCachedDevices][6];
  int dev = 0;
  const bool cached = hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxCachedDevices;
  std::atomic<int>* lim = cached ? &lds_limit[dev][(learned ? 3 : 0) + (pairs ? 0 : (ext8 ? 1 : 2))] : nullptr;
  if (lim == nullptr || (int)lds > lim->load(std::memory_order_relaxed)) {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
      return e;
    }
    if (lim != nullptr) {
      int cur = lim->load(std::memory_order_relaxed);
      while ((int)lds > cur && !lim->compare_exchange_weak(cur, (int)lds, std::memory_order_relaxed)) {
      }
    }
  }
  // lds (from decode_item_lds_bytes, a multiple of 4) ends with the claimed-item word.
  const unsigned qslot = (unsigned)lds - kQueueBytes;
#define LH_LAUNCH(L, P, X)                                                                                        \
  hipLaunchKernelGGL((k_ldpc_decode<L, P, X>), dim3(grid), dim3(nt), lds, s, d_hdr, d_items, d_jobs, rm, cw, bits, \
                     results, crc_x, gtab, scale_tab, fb_iters, qslot, fa)
  if (pairs) {
    if (learned) {
      LH_LAUNCH(true, true, false);
    } else {
      LH_LAUNCH(false, true, false);
    }
  } else if (ext8) {
    if (learned) {
      LH_LAUNCH(true, false, true);
    } else {
      LH_LAUNCH(false, false, true);
    }
  } else {
    if (learned) {
      LH_LAUNCH(true, false, false);
    } else {
      LH_LAUNCH(false, false, false);
    }
  }
#undef LH_LAUNCH
  return hipGetLastError();
}

// Zeroes the sub-queue counters of a resident descriptor block (graph replays):
// `words` words from `ctr` on.
__global__ void k_reset_counters(uint32_t* ctr, unsigned words)
{
  for (unsigned i = threadIdx.x; i < words; i += blockDim.x) {
    ctr[i] = 0;
  }
}

hipError_t launch_reset_counters(uint32_t* ctr, unsigned words, hipStream_t s)
{
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(256), 0, s, ctr, words);
  return hipGetLastError();
}

int decode_occupancy(int nt, size_t lds, bool learned, bool pairs, bool ext8)
{
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, decode_kernel(learned, pairs, ext8 && !pairs), nt, lds) !=
      hipSuccess) {
    return 1;
  }
  return n;
}

hipError_t launch_rate_dematch(const rm_params* d_jobs, unsigned n, unsigned maxN, const int8_t* cw, int8_t* rm,
                               hipStream_t s)
{
  const unsigned chunks = (maxN + 1023) / 1024;
  hipLaunchKernelGGL(k_rate_dematch, dim3(chunks, n), dim3(256), 0, s, d_jobs, cw, rm);
  return hipGetLastError();
}

hipError_t launch_tb_finalize(const lh_tb_desc_dev* d_tbs, unsigned n, const uint8_t* cb_bits, const int32_t* cb_res,
                              uint8_t* tb, int32_t* status, const uint32_t* lut, const uint32_t* pow2, hipStream_t s,
                              const tb_signal* sig)
{
  if (sig != nullptr) {
    hipLaunchKernelGGL(k_tb_finalize<true>, dim3(n), dim3(kTbThreads), 0, s, d_tbs, cb_bits, cb_res, tb, status, lut, pow2,
                       *sig);
  } else {
    hipLaunchKernelGGL(k_tb_finalize<false>, dim3(n), dim3(kTbThreads), 0, s, d_tbs, cb_bits, cb_res, tb, status, lut, pow2,
                       tb_signal{});
  }
  return hipGetLastError();
}

hipError_t launch_crc(const uint64_t* d_offs, const uint32_t* d_nbits, unsigned n, const uint8_t* bits,
                      const uint32_t* xtab, int order, uint32_t* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_crc, dim3(n), dim3(256), 0, s, d_offs, d_nbits, bits, xtab, order, out);
  return hipGetLastError();
}

} // namespace lh
