// Synthetic codeblock generator for gfx950: message bits, CRC, LDPC encoding, rate
// matching, Gray QAM mapping, AWGN and max-log soft demapping to int8 LLRs, one
// workgroup per codeblock with the whole codeword in LDS.
//
// It produces the inputs of the decoder benchmarks and of the learned-decoder training
// (TS38.212 5.3.2 encoding as in lib/phy/upper/channel_coding/ldpc/ldpc_encoder_generic.cpp,
// 5.4.2 rate matching as in ldpc_rate_matcher_impl.cpp, TS38.211 5.1 mapping as in
// modulation_mapper_lut_impl.cpp, quantisation as log_likelihood_ratio::quantize).
//
// Random streams are splitmix64 in counter form: output k of the stream seeded with s
// is mix(s + (k + 1) * 0x9e3779b97f4a7c15), so every thread draws its own outputs
// without a sequential walk.  Message bit j is bit 63 - (j mod 64) of output j / 64;
// Gaussian component i (I and Q interleaved) uses outputs 2i and 2i + 1 (Box-Muller).
#include <hip/hip_runtime.h>

#include "bg_tables.inc"
#include "lh_gf.h"
#include "lh_internal.h"

#pragma clang fp contract(off)

namespace lh {

namespace {

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t stream_at(uint64_t seed, uint64_t k) { return mix64(seed + (k + 1) * kGamma); }

__device__ __forceinline__ double unit_open(uint64_t x) { return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t i)
{
  const double u1 = unit_open(stream_at(seed, 2 * i));
  const double u2 = unit_open(stream_at(seed, 2 * i + 1));
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Gray PAM amplitude of one component from its bits b_0 (sign), b_1, ... (nb bits, b_k
// = bit k of the mask).
__device__ __forceinline__ int pam_level(uint32_t b, int nb)
{
  int v = 1;
  for (int k = nb - 1; k >= 1; --k) {
    v = (1 << (nb - k)) - (1 - 2 * (int)((b >> k) & 1u)) * v;
  }
  return (1 - 2 * (int)(b & 1u)) * v;
}

__device__ __forceinline__ double qam_norm(int qm)
{
  return qm == 2 ? 1.0 / sqrt(2.0) : qm == 4 ? 1.0 / sqrt(10.0) : qm == 6 ? 1.0 / sqrt(42.0) : 1.0 / sqrt(170.0);
}

__device__ __forceinline__ int8_t quantize(float v, float range)
{
  const float c = fabsf(v) > range ? copysignf(range, v) : v;
  return (int8_t)roundf(c / range * 120.0f);
}

__device__ __forceinline__ uint32_t block_xor(uint32_t x, uint32_t* red)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    x ^= (uint32_t)__shfl_xor((int)x, o);
  }
  const int w = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = x;
  }
  __syncthreads();
  x = 0;
  for (int i = 0; i < nw; ++i) {
    x ^= red[i];
  }
  return x;
}

// Shift of the edge of row `row` on column `col`, or -1 (uniform scan of the row).
__device__ __forceinline__ int edge_shift(const uint32_t* gt, int rs0, int rs1, uint32_t colZ)
{
  int s = -1;
  for (int e = rs0; e < rs1; ++e) {
    const uint32_t g = gt[e];
    if ((g & 0xffffu) == colZ) {
      s = (int)(g >> 16);
    }
  }
  return s;
}

constexpr int kGenMaxBytes = 68 * kMaxZ; // (K_b + M) * Z of BG1 at Z = 384

// --- Packed-word encoder (lifting sizes Z >= 32) -------------------------------------
// A lifted column is a Z-bit vector held in kColWords words (bit j of the column = bit
// j mod 32 of word j / 32, bits past Z zero), one zero word behind it.  The cyclic shift
// of an edge is then a 32-bit window of the vector: one funnel shift, plus the wrapped
// start of the vector OR-ed in near its end.
constexpr int kColWords = kMaxZ / 32 + 1;              // words per column, with the zero word
constexpr int kMsgWords = 22 * kMaxZ / 32 + 1;         // packed message of BG1 at Z = 384, + 1
constexpr int kPackedWords = kMsgWords + (22 + 4 + 5) * kColWords + 316; // message, columns, lambda, edges

// Bits a .. a + 31 (mod Z) of the column vector v (0 <= a < Z, Z >= 32).
__device__ __forceinline__ uint32_t cyc32(const uint32_t* v, int Z, int a)
{
  const int q = a >> 5;
  uint32_t x = __builtin_amdgcn_alignbit(v[q + 1], v[q], (uint32_t)(a & 31));
  if (a + 32 > Z) {
    x |= v[0] << (Z - a);
  }
  return x;
}

// Valid bits of word w of a Z-bit vector.
__device__ __forceinline__ uint32_t word_mask(int w, int Z)
{
  const int n = Z - 32 * w;
  return n >= 32 ? 0xffffffffu : ((1u << n) - 1u);
}

// (x mod Z) for -Z <= x < 2Z.
__device__ __forceinline__ int wrap_z(int x, int Z)
{
  x = x < 0 ? x + Z : x;
  return x >= Z ? x - Z : x;
}

// Bytes 0/1 of 32 bits, four per word store: bit k of x to byte k of the four at `out`.
__device__ __forceinline__ void spread_bits(uint32_t x, uint8_t* out, int n)
{
  if ((reinterpret_cast<uintptr_t>(out) & 3) == 0 && n == 32) {
    uint32_t* o4 = reinterpret_cast<uint32_t*>(out);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      // Bit k of the nibble to bit 8k: the products of distinct bits never meet.
      o4[q] = (((x >> (4 * q)) & 15u) * 0x00204081u) & 0x01010101u;
    }
  } else {
    for (int k = 0; k < n; ++k) {
      out[k] = (uint8_t)((x >> k) & 1u);
    }
  }
}

// Column of a lifted edge word (col * Z in the low half): col * Z / Z.
__device__ __forceinline__ int edge_col(uint32_t g, float rz)
{
  return (int)((float)(g & 0xffffu) * rz + 0.5f);
}

// LDPC encoding of the message of jb (msg_input) into the codeword bytes cw (one bit per
// byte, (K + M) Z), TS38.212 5.3.2 with the double-diagonal core of
// ldpc_encoder_generic.cpp: lambda_i = systematic sums of the four core rows, column K
// from their sum, the other core columns one row at a time, then every extension row
// independently.  Each step is word-parallel over the 32-row words of a column.  pk:
// kPackedWords words of LDS.  Z >= 32.
template <int BG>
__device__ void encode_packed(const gen_job& jb, const uint8_t* __restrict__ msgs, const uint32_t* __restrict__ gtab,
                              uint8_t* cw, uint32_t* pk)
{
  constexpr int K = BG == 1 ? 22 : 10, M = BG == 1 ? 46 : 42;
  const uint16_t* rs = BG == 1 ? lh_bg1_row_start : lh_bg2_row_start;
  const int t = (int)threadIdx.x, NT = (int)blockDim.x;
  const int Z = jb.Z;
  const int W = (Z + 31) >> 5;
  const int kz = K * Z;
  const int nmw = (kz + 31) >> 5;
  const float rz = 1.0f / (float)Z;
  uint32_t* pm = pk;                         // packed message, LSB first, + one zero word
  uint32_t* col = pk + kMsgWords;            // columns 0 .. K + 3
  uint32_t* lam = col + (K + 4) * kColWords; // lambda_0..3, then their sum
  uint32_t* gt = lam + 5 * kColWords;        // the graph's lifted edge table, staged in LDS
  {
    const uint32_t* g = gtab + jb.gt_off;
    for (int e = t; e < rs[M]; e += NT) {
      gt[e] = g[e];
    }
  }

  // 1. Message bytes (MSB first) to LSB-first words, and to the codeword bytes.
  {
    const uint8_t* src = msgs + jb.msg_off;
    const int nbytes = (kz + 7) >> 3;
    const bool al4 = (jb.msg_off & 3u) == 0;
    for (int w = t; w <= nmw; w += NT) {
      uint32_t x = 0;
      if (w < nmw) {
        if (al4 && 4 * w + 4 <= nbytes) {
          x = reinterpret_cast<const uint32_t*>(src)[w];
        } else {
          for (int k = 0; k < 4 && 4 * w + k < nbytes; ++k) {
            x |= (uint32_t)src[4 * w + k] << (8 * k);
          }
        }
        // Byte k of x holds bits 8k .. 8k + 7 MSB first: reverse the bits of each byte.
        x = __builtin_bitreverse32(__builtin_bswap32(x));
        if (32 * w + 32 > kz) {
          x &= (1u << (kz - 32 * w)) - 1u;
        }
        spread_bits(x, cw + 32 * w, min(32, kz - 32 * w));
      }
      pm[w] = x;
    }
  }
  __syncthreads();
  // 2. Info columns from the message string; zero words behind every column.
  for (int task = t; task < (K + 4) * kColWords; task += NT) {
    const int c = task / kColWords, w = task - c * kColWords;
    uint32_t x = 0;
    if (c < K && w < W) {
      const int a = c * Z + 32 * w;
      x = __builtin_amdgcn_alignbit(pm[(a >> 5) + 1], pm[a >> 5], (uint32_t)(a & 31)) & word_mask(w, Z);
    }
    col[task] = x;
  }
  for (int task = t; task < 5 * kColWords; task += NT) {
    lam[task] = 0;
  }
  __syncthreads();
  // 3. lambda_i[r] = XOR over the systematic edges of core row i of column[(r + s) mod Z].
  for (int task = t; task < 4 * W; task += NT) {
    const int i = task / W, w = task - i * W;
    uint32_t x = 0;
    for (int e = rs[i]; e < rs[i + 1]; ++e) {
      const uint32_t g = gt[e];
      const int c = edge_col(g, rz);
      if (c < K) {
        x ^= cyc32(col + c * kColWords, Z, wrap_z(32 * w + (int)(g >> 16), Z));
      }
    }
    lam[i * kColWords + w] = x & word_mask(w, Z);
  }
  __syncthreads();
  // Shifts of the core columns in the core rows (-1: no edge), uniform.
  int sh[4][4];
  for (int i = 0; i < 4; ++i) {
    for (int c = 0; c < 4; ++c) {
      sh[i][c] = edge_shift(gt, rs[i], rs[i + 1], (uint32_t)((K + c) * Z));
    }
  }
  // The sum of the four core rows leaves one circulant on column K: the shift that occurs
  // an odd number of times among the column-K entries.
  int tshift = -1;
  for (int i = 0; i < 4; ++i) {
    if (sh[i][0] < 0) {
      continue;
    }
    int dup = 0;
    for (int i2 = 0; i2 < 4; ++i2) {
      dup += (i2 != i && sh[i2][0] == sh[i][0]) ? 1 : 0;
    }
    if ((dup & 1) == 0) {
      tshift = sh[i][0];
    }
  }
  uint32_t* lsum = lam + 4 * kColWords;
  for (int w = t; w < W; w += NT) {
    lsum[w] = lam[w] ^ lam[kColWords + w] ^ lam[2 * kColWords + w] ^ lam[3 * kColWords + w];
  }
  __syncthreads();
  // 4. Column K: p[q] = sum[(q - tshift) mod Z].
  for (int w = t; w < W; w += NT) {
    col[K * kColWords + w] = cyc32(lsum, Z, wrap_z(32 * w - tshift, Z)) & word_mask(w, Z);
  }
  __syncthreads();
  // The other core columns: a core row with exactly one unknown column gives it,
  // column[j] = lambda_i[(j - s_u) mod Z] ^ XOR over the known core columns c of
  // column_c[(j - s_u + s_c) mod Z].
  bool known[4] = {true, false, false, false};
  for (int pass = 0; pass < 4; ++pass) {
    for (int i = 0; i < 4; ++i) {
      int unk = -1, nunk = 0;
      for (int c = 0; c < 4; ++c) {
        if (sh[i][c] >= 0 && !known[c]) {
          unk = c;
          ++nunk;
        }
      }
      if (nunk != 1) {
        continue;
      }
      const int su = sh[i][unk];
      for (int w = t; w < W; w += NT) {
        uint32_t x = cyc32(lam + i * kColWords, Z, wrap_z(32 * w - su, Z));
        for (int c = 0; c < 4; ++c) {
          if (sh[i][c] >= 0 && known[c]) {
            x ^= cyc32(col + (K + c) * kColWords, Z, wrap_z(wrap_z(32 * w - su, Z) + sh[i][c], Z));
          }
        }
        col[(K + unk) * kColWords + w] = x & word_mask(w, Z);
      }
      known[unk] = true;
      __syncthreads();
    }
  }
  // 5. Core columns to the codeword bytes; extension rows i >= 4 (identity part on column
  //    K + i with shift s_own): column_{K+i}[j] = XOR over the other edges of
  //    column_c[(j - s_own + s) mod Z].  Without the codeword output and when the rate
  //    matcher's ranks [start, start + E) do not wrap, only the rows up to the column of
  //    the last position it reads (past the fillers) are needed: the other columns are
  //    neither computed nor written.
  const int rows = encode_parity_rows(jb.rm, K, M, Z, jb.want_cw != 0);
  for (int task = t; task < rows * W; task += NT) {
    const int row = task / W, w = task - row * W;
    const int n = min(32, Z - 32 * w);
    if (row < 4) {
      spread_bits(col[(K + row) * kColWords + w], cw + (K + row) * Z + 32 * w, n);
      continue;
    }
    const int i = row;
    int own = 0;
    for (int e = rs[i]; e < rs[i + 1]; ++e) {
      const uint32_t g = gt[e];
      if (edge_col(g, rz) == K + i) {
        own = (int)(g >> 16);
      }
    }
    const int base = wrap_z(32 * w - own, Z);
    uint32_t x = 0;
    for (int e = rs[i]; e < rs[i + 1]; ++e) {
      const uint32_t g = gt[e];
      const int c = edge_col(g, rz);
      if (c < K + 4) {
        x ^= cyc32(col + c * kColWords, Z, wrap_z(base + (int)(g >> 16), Z));
      }
    }
    spread_bits(x & word_mask(w, Z), cw + (K + i) * Z + 32 * w, n);
  }
  __syncthreads();
}

} // namespace

// ENCODE: the transmit side only (ldpc_encoder + ldpc_rate_matcher): the message comes
// from msgs, and the rate-matched bits are written packed (MSB first) to out_bits +
// rm.cw_off, with the codeword (N_short * Z bits, one per byte, as
// ldpc_encoder_buffer::write_codeblock) at cw_bits + cwb_off when asked.
template <int BG, bool ENCODE, bool PACKED = false>
__device__ void gen_codeblock(const gen_job& jb, int8_t* __restrict__ llrs, uint8_t* __restrict__ msgs,
                              const uint32_t* __restrict__ crc_x, const uint32_t* __restrict__ gtab, uint8_t* cw,
                              uint8_t* lam, uint32_t* red, uint8_t* __restrict__ out_bits = nullptr,
                              uint8_t* __restrict__ cw_bits = nullptr, uint32_t* pk = nullptr)
{
  constexpr int K = BG == 1 ? 22 : 10, M = BG == 1 ? 46 : 42;
  const uint16_t* rs = BG == 1 ? lh_bg1_row_start : lh_bg2_row_start;
  const int t = (int)threadIdx.x, NT = (int)blockDim.x;
  const int Z = jb.Z;
  const int order = jb.crc_idx == 0 ? 0 : (jb.crc_idx == 3 ? 16 : 24);
  const int kd = K * Z - jb.F - order;
  const uint32_t* gt = gtab + jb.gt_off;

  // The transmit side encodes word-parallel for lifting sizes from 32 on (PACKED: the
  // message is an input and Z >= 32).
  constexpr bool packed = ENCODE && PACKED;
  if constexpr (packed) {
    encode_packed<BG>(jb, msgs, gtab, cw, pk);
  }

  // 1. Message: data bits, CRC, zero fillers; the parity part starts zeroed.  With
  //    msg_input the whole message (data, CRC, fillers) comes from msgs + msg_off.
  if (jb.msg_input && !packed) {
    // One 4-byte load per thread step (the message is read once, not once per bit) where
    // the message is 4-byte aligned and the word lies inside it; bytes otherwise.
    const uint8_t* src = msgs + jb.msg_off;
    const int kz = K * Z, nbytes = (kz + 7) / 8;
    const bool al4 = (jb.msg_off & 3u) == 0;
    for (int w = t; 4 * w < nbytes; w += NT) {
      uint32_t x = 0;
      if (al4 && 4 * w + 4 <= nbytes) {
        x = reinterpret_cast<const uint32_t*>(src)[w];
      } else {
        for (int k = 0; k < 4 && 4 * w + k < nbytes; ++k) {
          x |= (uint32_t)src[4 * w + k] << (8 * k);
        }
      }
      for (int k = 0; k < 32 && 32 * w + k < kz; ++k) {
        cw[32 * w + k] = (uint8_t)((x >> (8 * (k >> 3) + 7 - (k & 7))) & 1u);
      }
    }
    for (int j = kz + t; j < (K + M) * Z; j += NT) {
      cw[j] = 0;
    }
    __syncthreads();
  }
  for (int w = t; !jb.msg_input && 64 * w < kd; w += NT) {
    const uint64_t x = stream_at(jb.seed_msg, (uint64_t)w);
    const int n = min(64, kd - 64 * w);
    for (int b = 0; b < n; ++b) {
      cw[64 * w + b] = (uint8_t)((x >> (63 - b)) & 1u);
    }
  }
  for (int j = kd + t; !jb.msg_input && j < (K + M) * Z; j += NT) {
    cw[j] = 0;
  }
  __syncthreads();
  if (order != 0 && !jb.msg_input) {
    const uint32_t* xt = crc_x + (jb.crc_idx - 1) * kCrcXLen + (kd - 1 + order);
    uint32_t x = 0;
    for (int j = t; j < kd; j += NT) {
      x ^= cw[j] ? xt[-j] : 0u;
    }
    const uint32_t crc = block_xor(x, red);
    for (int k = t; k < order; k += NT) {
      cw[kd + k] = (uint8_t)((crc >> (order - 1 - k)) & 1u);
    }
    __syncthreads();
  }

  // 2. Core parity (rows 0..3, double diagonal): lambda_i = systematic sum of row i.
  for (int r = t; !packed && r < Z; r += NT) {
    for (int i = 0; i < 4; ++i) {
      uint8_t acc = 0;
      for (int e = rs[i]; e < rs[i + 1]; ++e) {
        const uint32_t g = gt[e];
        const uint32_t colZ = g & 0xffffu;
        if (colZ < (uint32_t)(K * Z)) {
          int j = r + (int)(g >> 16);
          j = j >= Z ? j - Z : j;
          acc ^= cw[colZ + j];
        }
      }
      lam[i * kMaxZ + r] = acc;
    }
  }
  // The sum of the four core rows leaves one circulant on column K: its shift is the one
  // that appears an odd number of times among the column-K entries.
  int sh[4][4]; // sh[row][c] = shift of core column K + c in core row, or -1
  for (int i = 0; i < 4; ++i) {
    for (int c = 0; c < 4; ++c) {
      sh[i][c] = edge_shift(gt, rs[i], rs[i + 1], (uint32_t)((K + c) * Z));
    }
  }
  int tshift = -1;
  for (int i = 0; i < 4; ++i) {
    if (sh[i][0] < 0) {
      continue;
    }
    int dup = 0;
    for (int i2 = 0; i2 < 4; ++i2) {
      dup += (i2 != i && sh[i2][0] == sh[i][0]) ? 1 : 0;
    }
    if ((dup & 1) == 0) {
      tshift = sh[i][0];
    }
  }
  __syncthreads();
  for (int q = t; !packed && q < Z; q += NT) {
    int r = q - tshift;
    r = r < 0 ? r + Z : r;
    cw[K * Z + q] = lam[r] ^ lam[kMaxZ + r] ^ lam[2 * kMaxZ + r] ^ lam[3 * kMaxZ + r];
  }
  __syncthreads();
  // Remaining core columns: repeatedly a core row with exactly one unknown column.
  bool known[4] = {packed, packed, packed, packed};
  known[0] = true;
  for (int pass = 0; pass < 4; ++pass) {
    for (int i = 0; i < 4; ++i) {
      int unk = -1, nunk = 0;
      for (int c = 0; c < 4; ++c) {
        if (sh[i][c] >= 0 && !known[c]) {
          unk = c;
          ++nunk;
        }
      }
      if (nunk != 1) {
        continue;
      }
      for (int r = t; r < Z; r += NT) {
        uint8_t acc = lam[i * kMaxZ + r];
        for (int c = 0; c < 4; ++c) {
          if (sh[i][c] >= 0 && known[c]) {
            int j = r + sh[i][c];
            j = j >= Z ? j - Z : j;
            acc ^= cw[(K + c) * Z + j];
          }
        }
        int j = r + sh[i][unk];
        j = j >= Z ? j - Z : j;
        cw[(K + unk) * Z + j] = acc;
      }
      known[unk] = true;
      __syncthreads();
    }
  }

  // 3. Extension parity (rows 4..M-1, identity diagonal): independent rows.
  for (int idx = t; !packed && idx < (M - 4) * Z; idx += NT) {
    const int i = 4 + idx / Z, r = idx - (i - 4) * Z;
    uint8_t acc = 0;
    int own = 0;
    for (int e = rs[i]; e < rs[i + 1]; ++e) {
      const uint32_t g = gt[e];
      const uint32_t colZ = g & 0xffffu;
      if (colZ == (uint32_t)((K + i) * Z)) {
        own = (int)(g >> 16);
        continue;
      }
      int j = r + (int)(g >> 16);
      j = j >= Z ? j - Z : j;
      acc ^= cw[colZ + j];
    }
    int j = r + own;
    j = j >= Z ? j - Z : j;
    cw[(K + i) * Z + j] = acc;
  }
  __syncthreads();

  if constexpr (ENCODE) {
    // Codeword and rate matching (TS38.212 5.4.2, as the generator's step 5 below): output
    // bit j = s Qm + q is rate-matched bit e_i, i = q Kq + s, read from the circular
    // buffer at rank start + i (mod P), fillers skipped.
    const rm_params& rp = jb.rm;
    const uint8_t* d = cw + 2 * Z;
    if (jb.want_cw) {
      uint8_t* co = cw_bits + jb.cwb_off;
      for (int p = t; p < (K + M - 2) * Z; p += NT) {
        co[p] = d[p];
      }
    }
    uint8_t* ro = out_bits + rp.cw_off;
    const uint32_t qm = rp.Qm;
    const uint32_t F = rp.nsys - rp.info;
    // Bit e_i, i = q Kq + s: circular-buffer rank start + i (mod P), fillers skipped.
    const auto rm_bit = [&](uint32_t q, uint32_t s) -> uint32_t {
      uint32_t rank = rp.start + q * rp.Kq + s;
      if (rank >= rp.P) {
        rank = rank - rp.P < rp.P ? rank - rp.P : rank % rp.P;
      }
      return d[rank < rp.info ? rank : rank + F];
    };
    // Qm a power of two: output byte b holds symbols 8b / Qm .. of whole bits (s, q) known
    // per bit position.
    const bool p2 = (qm & (qm - 1u)) == 0;
    const uint32_t lq = (uint32_t)__builtin_ctz(qm);
    for (uint32_t b = (uint32_t)t; 8 * b < rp.E; b += (uint32_t)NT) {
      uint32_t v = 0;
      if (p2 && 8 * b + 8 <= rp.E) {
        const uint32_t s0 = (8 * b) >> lq;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
          v |= rm_bit(k & (qm - 1u), s0 + (k >> lq)) << (7 - k);
        }
      } else {
        uint32_t s = (8 * b) / qm, q = 8 * b - s * qm;
        for (uint32_t k = 0; k < 8 && 8 * b + k < rp.E; ++k) {
          v |= rm_bit(q, s) << (7 - k);
          if (++q == qm) {
            q = 0;
            ++s;
          }
        }
      }
      ro[b] = (uint8_t)v;
    }
    return;
  }

  // 4. Packed message bits (MSB first).
  if (jb.write_msg && !jb.msg_input) {
    uint8_t* out = msgs + jb.msg_off;
    const int nbytes = (K * Z + 7) / 8;
    for (int b = t; b < nbytes; b += NT) {
      uint32_t v = 0;
      for (int k = 0; k < 8; ++k) {
        const int j = 8 * b + k;
        v |= (j < K * Z ? (uint32_t)cw[j] : 0u) << (7 - k);
      }
      out[b] = (uint8_t)v;
    }
  }

  // 5. Rate matching (circular buffer from k0 without fillers, bit interleaving by Qm),
  //    mapping, AWGN, max-log demapping and quantisation, one symbol per thread step.
  const rm_params& rp = jb.rm;
  const int qm = rp.Qm, half = qm / 2;
  const double nrm = qam_norm(qm);
  const double sd = sqrt(jb.n0 / 2.0);
  const float n0f = (float)jb.n0;
  const float range = qm == 2 ? 24.0f : 20.0f;
  const uint8_t* d = cw + 2 * Z;
  int8_t* out = llrs + rp.cw_off;
  for (uint32_t s = (uint32_t)t; s < rp.Kq; s += (uint32_t)NT) {
    uint32_t bits = 0; // bit q = rate-matched bit q of the symbol
    for (int q = 0; q < qm; ++q) {
      const uint32_t i = (uint32_t)q * rp.Kq + s;
      const uint32_t rank = (uint32_t)(((uint64_t)rp.start + i) % rp.P);
      const uint32_t p = rank < rp.info ? rank : rank + (rp.nsys - rp.info);
      bits |= (uint32_t)d[p] << q;
    }
    for (int c = 0; c < 2; ++c) {
      uint32_t cb = 0; // component bits: I = even symbol bits, Q = odd
      for (int k = 0; k < half; ++k) {
        cb |= ((bits >> (2 * k + c)) & 1u) << k;
      }
      const float tx = (float)((double)pam_level(cb, half) * nrm);
      const float x = (float)((double)tx + sd * normal_at(jb.seed_noise, 2 * (uint64_t)s + (uint64_t)c));
      const double xd = (double)x;
      for (int k = 0; k < half; ++k) {
        double d0 = 1e30, d1 = 1e30;
        for (int pt = 0; pt < (1 << half); ++pt) {
          const double lev = (double)pam_level((uint32_t)pt, half) * nrm;
          const double dd = (xd - lev) * (xd - lev);
          if ((pt >> k) & 1) {
            d1 = dd < d1 ? dd : d1;
          } else {
            d0 = dd < d0 ? dd : d0;
          }
        }
        const float l = (float)((d1 - d0) / (double)n0f);
        out[s * (uint32_t)qm + 2 * k + c] = quantize(l, range);
      }
    }
  }
}

__global__ void __launch_bounds__(384) k_gen_codeblocks(const gen_job* __restrict__ jobs, int8_t* __restrict__ llrs,
                                                        uint8_t* __restrict__ msgs, const uint32_t* __restrict__ crc_x,
                                                        const uint32_t* __restrict__ gtab)
{
  __shared__ uint8_t cw[kGenMaxBytes];
  __shared__ uint8_t lam[4 * kMaxZ];
  __shared__ uint32_t red[16];
  const gen_job jb = jobs[blockIdx.x];
  if (jb.bg == 1) {
    gen_codeblock<1, false>(jb, llrs, msgs, crc_x, gtab, cw, lam, red);
  } else {
    gen_codeblock<2, false>(jb, llrs, msgs, crc_x, gtab, cw, lam, red);
  }
}

// LDPC encoder + rate matcher, one workgroup per codeblock (message from msgs).  PACKED:
// the codeblocks with Z >= 32 (word-parallel encoder), otherwise those with Z < 32; each
// instance leaves the others' codeblocks alone.
// PACKED (lifting sizes from 32 on): the packed words and then the codeword bytes in
// dynamic LDS, sized by the host for the columns the launch's codeblocks need.
template <bool PACKED>
__global__ void __launch_bounds__(384) k_encode_rate_match(const gen_job* __restrict__ jobs, uint8_t* __restrict__ msgs,
                                                           uint8_t* __restrict__ out_bits, uint8_t* __restrict__ cw_bits,
                                                           const uint32_t* __restrict__ crc_x,
                                                           const uint32_t* __restrict__ gtab)
{
  extern __shared__ __align__(16) uint8_t enc_lds[];
  __shared__ __align__(16) uint8_t cw_static[PACKED ? 16 : kGenMaxBytes];
  __shared__ uint8_t lam[PACKED ? 1 : 4 * kMaxZ];
  __shared__ uint32_t red[16];
  uint32_t* pk = PACKED ? reinterpret_cast<uint32_t*>(enc_lds) : nullptr;
  uint8_t* cw = PACKED ? enc_lds + 4 * kPackedWords : cw_static;
  const gen_job jb = jobs[blockIdx.x];
  if ((jb.Z >= 32) != PACKED) {
    return;
  }
  if (jb.bg == 1) {
    gen_codeblock<1, true, PACKED>(jb, nullptr, msgs, crc_x, gtab, cw, lam, red, out_bits, cw_bits, pk);
  } else {
    gen_codeblock<2, true, PACKED>(jb, nullptr, msgs, crc_x, gtab, cw, lam, red, out_bits, cw_bits, pk);
  }
}

// Transport-block generator: TB bytes from a splitmix64 stream (byte b = byte b mod 8,
// most significant first, of output b / 8), the TB CRC (CRC16 up to 3824 bits, else
// CRC24A), and the codeblock messages of the TX segmentation (TS38.212 5.2.2 as in
// lib/phy/upper/channel_coding/ldpc/ldpc_segmenter_tx_impl.cpp:124-198): codeblock c
// carries bits [c D, c D + D) of TB || TB CRC || zero padding, then its CRC24B when there
// are several codeblocks, then F zero fillers.  One workgroup (256 threads) per TB; the TB
// bytes go to tb (required: the codeblock bits are read back from there).
__device__ __forceinline__ uint32_t tb_bit(const uint8_t* tb, uint32_t tbs, uint32_t crc, uint32_t L, uint32_t p)
{
  if (p < tbs) {
    return (tb[p >> 3] >> (7 - (p & 7))) & 1u;
  }
  p -= tbs;
  return p < L ? (crc >> (L - 1 - p)) & 1u : 0u;
}

__global__ void __launch_bounds__(256) k_gen_tb(const gen_tb_job* __restrict__ jobs, uint8_t* __restrict__ tb_base,
                                                uint8_t* __restrict__ msgs, const uint32_t* __restrict__ crc_x,
                                                const uint32_t* __restrict__ lut24a, const uint32_t* __restrict__ pow2)
{
  __shared__ uint32_t red[16];
  __shared__ uint32_t lut[256];
  const gen_tb_job jb = jobs[blockIdx.x];
  const int t = (int)threadIdx.x, NT = (int)blockDim.x;
  uint8_t* tb = tb_base + jb.tb_off;
  const uint32_t nbytes = jb.tbs_bits / 8;
  lut[t] = lut24a[t];
  for (uint32_t b = (uint32_t)t; b < nbytes; b += (uint32_t)NT) {
    const uint64_t w = stream_at(jb.seed, b >> 3);
    tb[b] = (uint8_t)(w >> (56 - 8 * (b & 7)));
  }
  __threadfence_block();
  __syncthreads();
  // TB CRC.
  const bool crc16 = jb.tbs_bits <= 3824;
  const uint32_t L = crc16 ? 16u : 24u;
  uint32_t x = 0;
  if (crc16) {
    const uint32_t* xt = crc_x + 2 * kCrcXLen + (jb.tbs_bits - 1 + 16);
    for (uint32_t i = (uint32_t)t; i < jb.tbs_bits; i += (uint32_t)NT) {
      x ^= ((tb[i >> 3] >> (7 - (i & 7))) & 1u) ? xt[-(int)i] : 0u;
    }
  } else {
    const uint32_t chunk = (nbytes + (uint32_t)NT - 1) / (uint32_t)NT;
    const uint32_t a0 = (uint32_t)t * chunk, a1 = min(nbytes, a0 + chunk);
    uint32_t rem = 0;
    for (uint32_t b = a0; b < a1; ++b) {
      rem = ((rem << 8) & 0xffffffu) ^ lut[((rem >> 16) ^ tb[b]) & 0xffu];
    }
    if (a0 < a1) {
      x = gf_mulmod(rem, gf_xpow(8u * (nbytes - a1), pow2, 0x1864cfbu, 24), 0x1864cfbu, 24);
    }
  }
  const uint32_t tbcrc = block_xor(x, red);
  // Codeblock messages.
  const uint32_t C = jb.nof_cbs, K = jb.segment_length, F = jb.nof_filler_bits;
  const uint32_t D = C > 1 ? K - F - 24 : K - F;
  const uint32_t kbytes = (K + 7) / 8;
  const uint32_t* xb = crc_x + kCrcXLen + (D - 1 + 24); // CRC24B weights
  for (uint32_t c = 0; c < C; ++c) {
    uint8_t* m = msgs + jb.msg_off + (size_t)c * jb.msg_stride;
    uint32_t cbcrc = 0;
    if (C > 1) {
      uint32_t y = 0;
      for (uint32_t i = (uint32_t)t; i < D; i += (uint32_t)NT) {
        y ^= tb_bit(tb, jb.tbs_bits, tbcrc, L, c * D + i) ? xb[-(int)i] : 0u;
      }
      __syncthreads(); // red is reused
      cbcrc = block_xor(y, red);
    }
    for (uint32_t b = (uint32_t)t; b < kbytes; b += (uint32_t)NT) {
      uint32_t v = 0;
      for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t i = 8 * b + k;
        uint32_t bit = 0;
        if (i < D) {
          bit = tb_bit(tb, jb.tbs_bits, tbcrc, L, c * D + i);
        } else if (C > 1 && i < D + 24) {
          bit = (cbcrc >> (23 - (i - D))) & 1u;
        }
        v |= bit << (7 - k);
      }
      m[b] = (uint8_t)v;
    }
    __syncthreads();
  }
}

hipError_t launch_gen_tb(const gen_tb_job* d_jobs, unsigned n, uint8_t* tb, uint8_t* msgs, const uint32_t* crc_x,
                         const uint32_t* lut24a, const uint32_t* pow2, hipStream_t s)
{
  hipLaunchKernelGGL(k_gen_tb, dim3(n), dim3(256), 0, s, d_jobs, tb, msgs, crc_x, lut24a, pow2);
  return hipGetLastError();
}

hipError_t launch_gen_codeblocks(const gen_job* d_jobs, unsigned n, int8_t* llrs, uint8_t* msgs, const uint32_t* crc_x,
                                 const uint32_t* gtab, hipStream_t s)
{
  hipLaunchKernelGGL(k_gen_codeblocks, dim3(n), dim3(kMaxThreads), 0, s, d_jobs, llrs, msgs, crc_x, gtab);
  return hipGetLastError();
}

// Threads of a word-parallel encoder workgroup: its phases have at most a few hundred
// tasks, and small workgroups with small LDS keep more codeblocks resident per CU.
constexpr int kEncodeThreads = 192;

hipError_t launch_encode_rate_match(const gen_job* d_jobs, unsigned n, uint8_t* msgs, uint8_t* out_bits,
                                    uint8_t* cw_bits, const uint32_t* crc_x, const uint32_t* gtab, bool any_small,
                                    size_t packed_cw_bytes, hipStream_t s)
{
  const size_t lds = 4 * (size_t)kPackedWords + ((packed_cw_bytes + 15) & ~(size_t)15);
  hipLaunchKernelGGL(k_encode_rate_match<true>, dim3(n), dim3(kEncodeThreads), lds, s, d_jobs, msgs, out_bits,
                     cw_bits, crc_x, gtab);
  if (any_small) {
    hipLaunchKernelGGL(k_encode_rate_match<false>, dim3(n), dim3(kMaxThreads), 0, s, d_jobs, msgs, out_bits, cw_bits,
                       crc_x, gtab);
  }
  return hipGetLastError();
}

} // namespace lh
