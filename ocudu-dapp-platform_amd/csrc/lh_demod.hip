// Soft demodulation and descrambling of PUSCH codeword blocks: equalized symbols and
// their noise variances -> int8 LLRs with the scrambling sequence removed; then the
// UL-SCH demultiplexing of whole codewords (k_ulsch_demux, below).
//
// Follows the reference's generic (scalar) demappers and the PUSCH descrambling:
//  * demodulation_mapper_impl.cpp:13-47 (BPSK, pi/2-BPSK), demodulation_mapper_qpsk.cpp
//    (demod_QPSK_symbol), demodulation_mapper_qam16.cpp (demod_16QAM_symbol_01/_23, the
//    near-zero rule of demodulate_soft_QAM16), demodulation_mapper_qam64.cpp and
//    demodulation_mapper_qam256.cpp (interval_function over the slope/intercept tables of
//    demodulation_mapper_intervals.h), quantize in log_likelihood_ratio.cpp:72-81;
//  * the Gold sequence of TS38.211 5.2.1 (pseudo_random_generator_impl.cpp, Nc = 1600)
//    initialised with c_init (pusch_demodulator_impl.cpp:262-271) and revert_scrambling
//    (pusch_demodulator_impl.cpp:22-80, 411-414): an LLR is negated where the sequence
//    bit is 1.
//
// The float arithmetic is the reference's, operation by operation, with no contraction
// into fused multiply-adds, so the LLRs are bit-exact with the reference's scalar path.
//
// Work split: one workgroup per tile of 256 symbols of one block, one symbol per thread.
// The host gives each tile the two LFSR states at its first sequence bit (the block's
// start by a jump, then tile to tile by one product with A^(256 Qm) over GF(2), as byte
// tables).  The tile's 256 * Qm sequence bits are made first, one 32-bit word per thread
// for the first 8 Qm threads: word w by a jump of 32 w bits, products with the matrices
// A^(32 * 2^j) for the set bits of w (at most 6); then each thread demaps its symbol and
// applies the sequence bits from LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "lh_internal.h"
#include "lh_demux_rank.h"

namespace lh {

namespace {

constexpr int kTileSymbols = 256;
constexpr int kDemodThreads = 256;
constexpr unsigned kNc = 1600;

// One step of the two LFSRs on a 31-bit state (bit i = x(n + i)).
inline uint32_t x1_feedback(uint32_t s) { return (s ^ (s >> 3)) & 1u; }
inline uint32_t x2_feedback(uint32_t s) { return (s ^ (s >> 1) ^ (s >> 2) ^ (s >> 3)) & 1u; }

// GF(2) matrix as its 31 columns: column i = image of the unit state e_i.
struct gf2m {
  uint32_t col[31];
};

inline uint32_t apply(const gf2m& m, uint32_t s)
{
  uint32_t r = 0;
  for (int i = 0; i < 31; ++i) {
    if ((s >> i) & 1u) {
      r ^= m.col[i];
    }
  }
  return r;
}

inline gf2m square(const gf2m& m)
{
  gf2m r;
  for (int i = 0; i < 31; ++i) {
    r.col[i] = apply(m, m.col[i]);
  }
  return r;
}

// Transition matrices A^(2^k), k = 0 .. kJumpBits - 1, of one LFSR.
constexpr int kJumpBits = 28;
struct jump_set {
  gf2m pow2[kJumpBits];
};

jump_set make_jumps(bool second)
{
  jump_set j;
  for (int i = 0; i < 31; ++i) {
    const uint32_t s = 1u << i;
    const uint32_t fb = second ? x2_feedback(s) : x1_feedback(s);
    j.pow2[0].col[i] = (s >> 1) | (fb << 30);
  }
  for (int k = 1; k < kJumpBits; ++k) {
    j.pow2[k] = square(j.pow2[k - 1]);
  }
  return j;
}

const jump_set& jumps(bool second)
{
  static const jump_set j1 = make_jumps(false);
  static const jump_set j2 = make_jumps(true);
  return second ? j2 : j1;
}

uint32_t jump(bool second, uint32_t s, uint64_t n)
{
  const jump_set& j = jumps(second);
  for (int k = 0; k < kJumpBits && n != 0; ++k, n >>= 1) {
    if (n & 1u) {
      s = apply(j.pow2[k], s);
    }
  }
  return s;
}

// Max-log LLR of Gray-mapped PAM (TS38.211 5.1): for each interval of one bit level, the
// nearest amplitudes a0 (bit 0) and a1 (bit 1) give LLR * noise_var = slope x + intercept
// with slope = 2 (a0 - a1) / sqrt(N) and intercept = (a1^2 - a0^2) / N (amplitudes in
// units of 1/sqrt(N)).  Computed in float as k * (1 / sqrt(N)) and p / q, which are the
// values of the reference's tables.
// Amplitude (units of 1/sqrt(N)) of m PAM bits b[0] (sign) .. b[m-1].
int pam_amplitude(int m, unsigned bits)
{
  int a = 1;
  for (int l = m - 1; l >= 1; --l) {
    const int s = 1 - 2 * (int)((bits >> l) & 1u);
    a = (1 << (m - l)) - s * a;
  }
  return (1 - 2 * (int)(bits & 1u)) * a;
}

void pam_table(int m, int level, int nof_intervals, int width_units, float rsq, int N, float* slope, float* icpt)
{
  for (int idx = 0; idx < nof_intervals; ++idx) {
    const double u = (idx - nof_intervals / 2 + 0.5) * width_units;
    int best[2] = {0, 0};
    double dbest[2] = {1e30, 1e30};
    for (unsigned b = 0; b < (1u << m); ++b) {
      const int a = pam_amplitude(m, b);
      const int bit = (b >> level) & 1u;
      const double d = (u - a) * (u - a);
      if (d < dbest[bit]) {
        dbest[bit] = d;
        best[bit] = a;
      }
    }
    slope[idx] = (float)(2 * (best[0] - best[1])) * rsq;
    icpt[idx] = (float)(best[1] * best[1] - best[0] * best[0]) / (float)N;
  }
}

// A^(256 Qm) of both LFSRs as byte tables (entry [b][v]: image of byte b of the state
// holding v), for the tile-to-tile step.
struct tile_step {
  uint32_t t[2][4][256];
};

tile_step make_tile_step(int qm)
{
  tile_step ts;
  for (int lf = 0; lf < 2; ++lf) {
    gf2m m;
    for (int i = 0; i < 31; ++i) {
      m.col[i] = jump(lf == 1, 1u << i, (uint64_t)kTileSymbols * qm);
    }
    for (int b = 0; b < 4; ++b) {
      for (uint32_t v = 0; v < 256; ++v) {
        ts.t[lf][b][v] = apply(m, (v << (8 * b)) & 0x7fffffffu);
      }
    }
  }
  return ts;
}

const tile_step& tile_steps(int qm)
{
  static const tile_step t1 = make_tile_step(1), t2 = make_tile_step(2), t4 = make_tile_step(4),
                         t6 = make_tile_step(6), t8 = make_tile_step(8);
  return qm == 1 ? t1 : qm == 2 ? t2 : qm == 4 ? t4 : qm == 6 ? t6 : t8;
}

} // namespace

void demod_tile_states(uint32_t c_init, uint64_t pos, int qm, uint32_t nof_tiles, uint32_t* x1, uint32_t* x2)
{
  uint32_t s1 = jump(false, 1u, kNc + pos);
  uint32_t s2 = jump(true, c_init & 0x7fffffffu, kNc + pos);
  const tile_step& ts = tile_steps(qm);
  for (uint32_t k = 0; k < nof_tiles; ++k) {
    x1[k] = s1;
    x2[k] = s2;
    s1 = ts.t[0][0][s1 & 255u] ^ ts.t[0][1][(s1 >> 8) & 255u] ^ ts.t[0][2][(s1 >> 16) & 255u] ^ ts.t[0][3][s1 >> 24];
    s2 = ts.t[1][0][s2 & 255u] ^ ts.t[1][1][(s2 >> 8) & 255u] ^ ts.t[1][2][(s2 >> 16) & 255u] ^ ts.t[1][3][s2 >> 24];
  }
}

void build_demod_tables(uint32_t* out)
{
  // Words 0 .. 2*kDemodJumps*31-1: A^(32 * 2^j) of x1, then of x2.
  for (int lf = 0; lf < 2; ++lf) {
    const jump_set& j = jumps(lf == 1);
    for (int k = 0; k < kDemodJumps; ++k) {
      for (int i = 0; i < 31; ++i) {
        out[(lf * kDemodJumps + k) * 31 + i] = j.pow2[k + 5].col[i];
      }
    }
  }
  // Float tables: 64-QAM levels 0, 1 (8 intervals), 2 (4); 256-QAM levels 0-2 (16), 3 (8).
  float* f = reinterpret_cast<float*>(out + 2 * kDemodJumps * 31);
  const float r42 = 1.0F / std::sqrt(42.0F), r170 = 1.0F / std::sqrt(170.0F);
  pam_table(3, 0, 8, 2, r42, 42, f + 0, f + 8);
  pam_table(3, 1, 8, 2, r42, 42, f + 16, f + 24);
  pam_table(3, 2, 4, 4, r42, 42, f + 32, f + 40);
  pam_table(4, 0, 16, 2, r170, 170, f + 48, f + 64);
  pam_table(4, 1, 16, 2, r170, 170, f + 80, f + 96);
  pam_table(4, 2, 16, 2, r170, 170, f + 112, f + 128);
  pam_table(4, 3, 8, 4, r170, 170, f + 144, f + 152);
  f[160] = r42;
  f[161] = r170;
  f[162] = 1.0F / std::sqrt(10.0F);
}

namespace {

#pragma clang fp contract(off)

// log_likelihood_ratio::quantize: clip to +-range, then round(v / range * LLR_MAX) with the
// maximum finite LLR 120.
__device__ __forceinline__ int quantize_llr(float v, float range)
{
  const float c = fabsf(v) > range ? copysignf(range, v) : v;
  return (int)roundf(__fdiv_rn(c, range) * 120.0F);
}

__device__ __forceinline__ int interval_llr(float x, float rcp_noise, float width, int n, const float* slope,
                                            const float* icpt)
{
  int idx = (int)floorf(__fdiv_rn(x, width)) + n / 2;
  idx = min(max(idx, 0), n - 1);
  float l = slope[idx] * x + icpt[idx];
  l *= rcp_noise;
  return quantize_llr(l, 20.0F);
}

__device__ __forceinline__ int qam16_01(float x, float nv, float r10)
{
  if (!(nv > 0.0F)) {
    return 0;
  }
  float l = (4.0F * r10) * x;
  if (fabsf(x) > 2.0F * r10) {
    l = 2.0F * l - copysignf(0.8F, x);
  }
  l = __fdiv_rn(l, nv);
  return quantize_llr(l, 20.0F);
}

__device__ __forceinline__ int qam16_23(float x, float nv, float r10)
{
  if (!(nv > 0.0F)) {
    return 0;
  }
  float l = 0.8F - (4.0F * r10) * fabsf(x);
  l = __fdiv_rn(l, nv);
  return quantize_llr(l, 20.0F);
}

__device__ __forceinline__ int bpsk_llr(float re, float im, float nv)
{
  if (!(nv > 0.0F)) {
    return 0;
  }
  const float k = 2.0F * 1.41421356237309504880F;
  return quantize_llr(__fdiv_rn(k * (re + im), nv), 24.0F);
}

// The Qm LLRs of one symbol into llr[0 .. Qm-1].  `odd`: odd position in its block (pi/2-BPSK).
__device__ __forceinline__ void demap_symbol(int mod, float re, float im, float nv, bool odd, const float* ft,
                                             int* llr)
{
  if (mod == 0) { // pi/2-BPSK: odd symbols rotated by -pi/2
    llr[0] = odd ? bpsk_llr(im, -re, nv) : bpsk_llr(re, im, nv);
    return;
  }
  if (mod == 1) {
    llr[0] = bpsk_llr(re, im, nv);
    return;
  }
  if (mod == 2) {
    const float k = 2.0F * 1.41421356237309504880F;
    const bool ok = nv > 0.0F;
    llr[0] = ok ? quantize_llr(__fdiv_rn(k * re, nv), 24.0F) : 0;
    llr[1] = ok ? quantize_llr(__fdiv_rn(k * im, nv), 24.0F) : 0;
    return;
  }
  const bool zero = re * re + im * im < 1e-9F;
  if (mod == 4) {
    const float r10 = ft[162];
    llr[0] = zero ? 0 : qam16_01(re, nv, r10);
    llr[1] = zero ? 0 : qam16_01(im, nv, r10);
    llr[2] = zero ? 0 : qam16_23(re, nv, r10);
    llr[3] = zero ? 0 : qam16_23(im, nv, r10);
    return;
  }
  const float rcp = nv > 0.0F ? __fdiv_rn(1.0F, nv) : 0.0F;
  if (mod == 6) {
    const float w2 = 2.0F * ft[160], w4 = 4.0F * ft[160];
    llr[0] = interval_llr(re, rcp, w2, 8, ft + 0, ft + 8);
    llr[1] = interval_llr(im, rcp, w2, 8, ft + 0, ft + 8);
    llr[2] = interval_llr(re, rcp, w2, 8, ft + 16, ft + 24);
    llr[3] = interval_llr(im, rcp, w2, 8, ft + 16, ft + 24);
    llr[4] = interval_llr(re, rcp, w4, 4, ft + 32, ft + 40);
    llr[5] = interval_llr(im, rcp, w4, 4, ft + 32, ft + 40);
  } else {
    const float w2 = 2.0F * ft[161], w4 = 4.0F * ft[161];
    llr[0] = interval_llr(re, rcp, w2, 16, ft + 48, ft + 64);
    llr[1] = interval_llr(im, rcp, w2, 16, ft + 48, ft + 64);
    llr[2] = interval_llr(re, rcp, w2, 16, ft + 80, ft + 96);
    llr[3] = interval_llr(im, rcp, w2, 16, ft + 80, ft + 96);
    llr[4] = interval_llr(re, rcp, w2, 16, ft + 112, ft + 128);
    llr[5] = interval_llr(im, rcp, w2, 16, ft + 112, ft + 128);
    llr[6] = interval_llr(re, rcp, w4, 8, ft + 144, ft + 152);
    llr[7] = interval_llr(im, rcp, w4, 8, ft + 144, ft + 152);
  }
  if (zero) {
    for (int b = 0; b < 8; ++b) {
      llr[b] = 0;
    }
  }
}

#pragma clang fp contract(on)

__device__ __forceinline__ uint32_t gf2_apply(const uint32_t* __restrict__ col, uint32_t s)
{
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 31; ++i) {
    r ^= ((s >> i) & 1u) ? col[i] : 0u;
  }
  return r;
}

__global__ void __launch_bounds__(kDemodThreads)
    k_demod_descramble(const demod_tile* __restrict__ tiles, const demod_job* __restrict__ jobs,
                       const uint32_t* __restrict__ tab, const float2* __restrict__ symbols,
                       const float* __restrict__ noise, int8_t* __restrict__ llrs)
{
  __shared__ uint32_t seq[8 * kTileSymbols / 32 + 1];
  __shared__ float ft[kDemodFloats];
  const demod_tile t = tiles[blockIdx.x];
  const demod_job jb = jobs[t.job];
  const int mod = (int)jb.modulation;
  const int qm = mod == 0 ? 1 : mod;
  const uint32_t tid = threadIdx.x;
  const float* gft = reinterpret_cast<const float*>(tab + 2 * kDemodJumps * 31);
  for (uint32_t i = tid; i < kDemodFloats; i += kDemodThreads) {
    ft[i] = gft[i];
  }
  // Sequence words of the tile: word w starts at bit 32 w of the tile.
  const uint32_t nwords = (uint32_t)(kTileSymbols / 32 * qm);
  if (tid < nwords) {
    uint32_t s1 = t.x1, s2 = t.x2;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if ((tid >> k) & 1u) {
        s1 = gf2_apply(tab + k * 31, s1);
        s2 = gf2_apply(tab + (kDemodJumps + k) * 31, s2);
      }
    }
    const uint32_t w1 = s1 | (((s1 ^ (s1 >> 3)) & 1u) << 31);
    const uint32_t w2 = s2 | (((s2 ^ (s2 >> 1) ^ (s2 >> 2) ^ (s2 >> 3)) & 1u) << 31);
    seq[tid] = w1 ^ w2;
  }
  if (tid == 0) {
    seq[nwords] = 0;
  }
  __syncthreads();
  const uint32_t sym = t.tile * kTileSymbols + tid;
  if (sym < jb.nof_symbols) {
    const float2 z = symbols[jb.sym_off + sym];
    const float nv = noise[jb.sym_off + sym];
    int l[8];
    demap_symbol(mod, z.x, z.y, nv, (sym & 1u) != 0, ft, l);
    const uint32_t pos = tid * (uint32_t)qm;
    const uint64_t win = (uint64_t)seq[pos >> 5] | ((uint64_t)seq[(pos >> 5) + 1] << 32);
    const uint32_t c = (uint32_t)(win >> (pos & 31u));
    int8_t* out = llrs + jb.llr_off + (uint64_t)sym * qm;
    for (int b = 0; b < qm; ++b) {
      const int v = ((c >> b) & 1u) ? -l[b] : l[b];
      out[b] = (int8_t)v;
    }
  }
}

// ---- UL-SCH demultiplexing ---------------------------------------------------------
//
// One thread per unit (1, 2 or 4 soft bits of one RE) of one OFDM symbol; the RE's fields
// and ranks come from demux_classify (lh_demux_rank.h).

__device__ __forceinline__ uint32_t seq_word(uint32_t s1, uint32_t s2)
{
  const uint32_t w1 = s1 | (((s1 ^ (s1 >> 3)) & 1u) << 31);
  const uint32_t w2 = s2 | (((s2 ^ (s2 >> 1) ^ (s2 >> 2) ^ (s2 >> 3)) & 1u) << 31);
  return w1 ^ w2;
}

// Sequence bits 32 w .. 32 w + 63 after the symbol's first soft bit.
__device__ uint64_t demux_seq_window(uint32_t s1, uint32_t s2, const uint32_t* __restrict__ tab, uint32_t w)
{
  for (int k = 0; k < kDemodJumps && (w >> k) != 0; ++k) {
    if ((w >> k) & 1u) {
      s1 = gf2_apply(tab + k * 31, s1);
      s2 = gf2_apply(tab + (kDemodJumps + k) * 31, s2);
    }
  }
  const uint32_t lo = seq_word(s1, s2);
  const uint32_t hi = seq_word(gf2_apply(tab, s1), gf2_apply(tab + kDemodJumps * 31, s2));
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

// on_uci_placeholder_1bit / _2bit (ulsch_demultiplex_impl.cpp:86-176) on the unit of u soft
// bits starting at symbol position r: bit 0 of each modulation symbol kept; bit 1 kept
// (2-bit payload) or negated where c(n - 1) ^ c(n) = 1 (1-bit payload); bits 2.. negated
// where c(n) = 1.
__device__ uint32_t demux_placeholder(uint32_t v, uint32_t r, uint32_t u, uint32_t qm, uint32_t payload,
                                      uint32_t x1, uint32_t x2, const uint32_t* __restrict__ tab)
{
  if (payload == 0 || v == 0) {
    return v;
  }
  const uint32_t base = r == 0 ? 0u : ((r - 1u) & ~31u);
  const uint64_t win = demux_seq_window(x1, x2, tab, base >> 5);
  uint32_t out = 0;
  for (uint32_t i = 0; i < u; ++i) {
    const uint32_t pos = r + i;
    const uint32_t k = pos % qm;
    const uint32_t c = (uint32_t)(win >> (pos - base)) & 1u;
    uint32_t neg = 0;
    if (k == 1 && payload == 1) {
      neg = c ^ ((uint32_t)(win >> (pos - 1u - base)) & 1u);
    } else if (k >= 2) {
      neg = c;
    }
    const int8_t b = (int8_t)(v >> (8 * i));
    out |= (uint32_t)(uint8_t)(neg ? (int8_t)-b : b) << (8 * i);
  }
  return out;
}

__device__ __forceinline__ uint32_t load_unit(const int8_t* p, uint32_t u)
{
  return u == 4 ? *reinterpret_cast<const uint32_t*>(p)
                : (u == 2 ? (uint32_t)*reinterpret_cast<const uint16_t*>(p) : (uint32_t)(uint8_t)*p);
}

__device__ __forceinline__ void store_unit(int8_t* p, uint32_t v, uint32_t u)
{
  if (u == 4) {
    *reinterpret_cast<uint32_t*>(p) = v;
  } else if (u == 2) {
    *reinterpret_cast<uint16_t*>(p) = (uint16_t)v;
  } else {
    *p = (int8_t)v;
  }
}

__global__ void __launch_bounds__(kDemuxThreads)
    k_ulsch_demux(const demux_tile* __restrict__ tiles, const demux_sym* __restrict__ syms,
                  const uint32_t* __restrict__ tab, const int8_t* __restrict__ llrs, int8_t* __restrict__ sch,
                  int8_t* __restrict__ ack, int8_t* __restrict__ c1, int8_t* __restrict__ c2)
{
  const demux_tile t = tiles[blockIdx.x];
  const demux_sym& s = syms[t.sym];
  const uint32_t u = s.unit, nbpr = s.nbpr;
  const uint32_t r = (t.tile * kDemuxTileUnits + threadIdx.x) * u;
  if (r >= s.M * nbpr) {
    return;
  }
  const uint32_t v = load_unit(llrs + s.llr_in + r, u);
  if (s.plain) {
    store_unit(sch + s.sch_out + r, v, u);
    return;
  }
  const uint32_t x = r / nbpr, b = r - x * nbpr;
  const demux_re f = demux_classify(s, x);
  const uint32_t vz = f.punctured ? 0u : v; // a 1- or 2-bit HARQ-ACK RE reads zero after extraction
  if (f.ack != kDemuxNone) {
    store_unit(ack + s.ack_out + (uint64_t)f.ack * nbpr + b,
               demux_placeholder(v, r, u, s.qm, s.ph_ack, s.x1, s.x2, tab), u);
  }
  if (f.c1 != kDemuxNone) {
    store_unit(c1 + s.c1_out + (uint64_t)f.c1 * nbpr + b, demux_placeholder(v, r, u, s.qm, s.ph_c1, s.x1, s.x2, tab),
               u);
  }
  if (f.c2 != kDemuxNone) {
    store_unit(c2 + s.c2_out + (uint64_t)f.c2 * nbpr + b,
               demux_placeholder(vz, r, u, s.qm, s.ph_c2, s.x1, s.x2, tab), u);
  }
  if (f.sch != kDemuxNone) {
    store_unit(sch + s.sch_out + (uint64_t)f.sch * nbpr + b, vz, u);
  }
}

} // namespace

hipError_t launch_demux(const demux_tile* tiles, unsigned nof_tiles, const demux_sym* syms, const uint32_t* tab,
                        const int8_t* llrs, int8_t* sch, int8_t* ack, int8_t* c1, int8_t* c2, hipStream_t stream)
{
  hipLaunchKernelGGL(k_ulsch_demux, dim3(nof_tiles), dim3(kDemuxThreads), 0, stream, tiles, syms, tab, llrs, sch, ack,
                     c1, c2);
  return hipGetLastError();
}

hipError_t launch_demod(const demod_tile* tiles, unsigned nof_tiles, const demod_job* jobs, const uint32_t* tab,
                        const float* symbols, const float* noise, int8_t* llrs, hipStream_t stream)
{
  hipLaunchKernelGGL(k_demod_descramble, dim3(nof_tiles), dim3(kDemodThreads), 0, stream, tiles, jobs, tab,
                     reinterpret_cast<const float2*>(symbols), noise, llrs);
  return hipGetLastError();
}

} // namespace lh
