// HIP kernels of the PUSCH LDPC receive chain for gfx950 (MI355X).
//
// Decoder: one workgroup per codeblock, one thread per lifted check row r (r < Z).
// Thread r owns, for every layer, the check-to-variable messages of check row r, so
// the layered min-sum update needs no cross-lane exchange: the cyclic shift of each
// edge becomes the LDS address (col*Z + (r + shift) mod Z) of the variable-node soft
// bit that row r reads and writes.  Soft bits live in LDS as int8; the four core
// layers keep their compressed check-to-variable state in VGPRs, the extension layers
// in thread-private LDS words.  Barriers separate layers (strict layer order).
#include <hip/hip_runtime.h>

#include "bg_tables.inc"
#include "lh_internal.h"

namespace lh {

template <int BG>
struct bgt;

template <>
struct bgt<1> {
  static constexpr int K = 22, M = 46, NNZ = 316, DEG_EXT = 10, NSHORT = 66;
  __host__ __device__ static constexpr int rs(int i) { return lh_bg1_row_start[i]; }
  __host__ __device__ static constexpr int col(int i) { return lh_bg1_col[i]; }
  __host__ __device__ static constexpr int v(int ils, int i) { return lh_bg1_v[ils][i]; }
};

template <>
struct bgt<2> {
  static constexpr int K = 10, M = 42, NNZ = 197, DEG_EXT = 6, NSHORT = 50;
  __host__ __device__ static constexpr int rs(int i) { return lh_bg2_row_start[i]; }
  __host__ __device__ static constexpr int col(int i) { return lh_bg2_col[i]; }
  __host__ __device__ static constexpr int v(int ils, int i) { return lh_bg2_v[ils][i]; }
};

// --- int8 log_likelihood_ratio arithmetic (values held in int) --------------------
// Infinite LLR: |x| > 120 (only +-127 occurs).
__device__ __forceinline__ bool llr_isinf(int x) { return (unsigned)(x + 120) > 240u; }
// Saturated sum with the infinity rules; callers guarantee b is finite.
__device__ __forceinline__ int llr_sat(int x) { return min(max(x, -120), 120); }
// Promotion: |x| > 120 becomes +-127.
__device__ __forceinline__ int llr_promote(int x) { return x > 120 ? 127 : (x < -120 ? -127 : x); }
// round(0.8 * m) for 0 <= m <= 127 (no ties occur: 0.8m has fraction 0, .2, .4, .6, .8).
__device__ __forceinline__ int scale08(int m) { return (4 * m + 2) / 5; }

// One layer update for check row r.  DEG compile-time (core rows) or a runtime deg
// <= DEG for extension rows.  State word layout: bits 0-6 scaled min1, 7-13 scaled
// min2, 14-18 argmin edge, signs of the check-to-variable messages in `sgn`.
// A zero state reproduces the reference's first visit (v2c = soft).
template <int DEG, bool RUNTIME_DEG>
__device__ __forceinline__ void layer_update(int8_t* __restrict__ soft, const uint32_t* __restrict__ et_row, int deg,
                                             int r, int Z, uint32_t& st, uint32_t& sgn)
{
  int v[DEG];
  int addr[DEG];
  const int s1o = st & 127, s2o = (st >> 7) & 127, amo = (st >> 14) & 31;
  // Composite keys |v2c| << 5 | edge: min over keys gives min1 and the first index
  // reaching it (the reference's strict '<'); inits are the reference's LLR_MAX.
  uint32_t m1 = 120u << 5, m2 = (120u << 5) | 31u;
  int sp = 0;
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (RUNTIME_DEG && e >= deg) {
      break;
    }
    const uint32_t et = et_row[e];
    int j = r + (int)(et >> 16);
    j = j >= Z ? j - Z : j;
    addr[e] = (int)(et & 0xffffu) + j;
    const int sv = soft[addr[e]];
    const int mag = (e == amo) ? s2o : s1o;
    const int c = ((sgn >> e) & 1u) ? -mag : mag;
    const int x = llr_isinf(sv) ? sv : llr_sat(sv - c);
    v[e] = x;
    const uint32_t key = ((uint32_t)abs(x) << 5) | (uint32_t)e;
    m2 = max(m1, min(m2, key));
    m1 = min(m1, key);
    sp ^= x;
  }
  const int amin = (int)(m1 & 31u);
  const int s1 = scale08((int)(m1 >> 5)), s2 = scale08((int)(m2 >> 5));
  uint32_t ns = 0;
#pragma unroll
  for (int e = 0; e < DEG; ++e) {
    if (RUNTIME_DEG && e >= deg) {
      break;
    }
    const int mag = (e == amin) ? s2 : s1;
    const int neg = (sp ^ v[e]) >> 31; // 0 or -1: sign product excluding edge e
    const int c = (mag ^ neg) - neg;
    ns |= (uint32_t)(neg & 1) << e;
    const int x = v[e];
    const int sb = llr_isinf(x) ? x : llr_promote(c + x);
    soft[addr[e]] = (int8_t)sb;
  }
  st = (uint32_t)s1 | ((uint32_t)s2 << 7) | ((uint32_t)amin << 14);
  sgn = ns;
}

template <int BG, int ROW>
__device__ __forceinline__ void core_layer(int8_t* soft, const uint32_t* etab, int r, int Z, uint32_t& st,
                                           uint32_t& sgn)
{
  constexpr int E0 = bgt<BG>::rs(ROW);
  constexpr int DEG = bgt<BG>::rs(ROW + 1) - E0;
  layer_update<DEG, false>(soft, etab + E0, DEG, r, Z, st, sgn);
}

// Wave64 + workgroup reductions.  `red` must hold 2 * (NT / 64) words.
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
    red[8 + w] = z;
  }
  __syncthreads();
  x = 0;
  z = 0;
  for (int i = 0; i < nw; ++i) {
    x ^= red[i];
    z |= red[8 + i];
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
// (check_syndrome, ldpc_decoder_impl.cpp:329-387).  Returns nonzero when it fails.
template <int BG>
__device__ uint32_t syndrome_fail(const int8_t* soft, const uint32_t* etab, int r, int Z, int L)
{
  uint32_t bad = 0;
  if (r < Z) {
    for (int l = 0; l < L; ++l) {
      const int e0 = bgt<BG>::rs(l), e1 = bgt<BG>::rs(l + 1);
      uint32_t p = 0;
      for (int e = e0; e < e1; ++e) {
        const uint32_t et = etab[e];
        int j = r + (int)(et >> 16);
        j = j >= Z ? j - Z : j;
        const int s = soft[(int)(et & 0xffffu) + j];
        bad |= (s == 0);
        p ^= (s <= 0);
      }
      bad |= p;
    }
  }
  return bad;
}

// Writes the hard decisions of soft[0, nbits) MSB-first; bits of the last byte beyond
// nbits keep their previous value.  all_ones: write ones instead (no-CRC failure path).
__device__ void write_hard_bits(const int8_t* soft, uint8_t* out, int nbits, bool all_ones)
{
  const int nbytes = (nbits + 7) >> 3;
  for (int b = threadIdx.x; b < nbytes; b += blockDim.x) {
    uint32_t byte = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = 8 * b + k;
      if (i < nbits) {
        byte |= (uint32_t)(all_ones || soft[i] <= 0) << (7 - k);
      }
    }
    const int valid = nbits - 8 * b;
    if (valid < 8) {
      const uint32_t keep = 0xffu >> valid;
      byte |= out[b] & keep;
    }
    out[b] = (uint8_t)byte;
  }
}

// Layered normalized min-sum decoder, behaviour of ldpc_decoder_impl::decode
// (lib/phy/upper/channel_coding/ldpc/ldpc_decoder_impl.cpp:41-145) with the arithmetic
// of ldpc_decoder_generic.cpp:12-104.
template <int BG>
__global__ void __launch_bounds__(kMaxThreads) k_ldpc_decode(const dec_job* __restrict__ jobs,
                                                             const int8_t* __restrict__ in_base,
                                                             uint8_t* __restrict__ bits_base,
                                                             int32_t* __restrict__ results,
                                                             const uint32_t* __restrict__ crc_x, int lmax)
{
  using T = bgt<BG>;
  constexpr int K = T::K;
  constexpr int NNZ_PAD = (T::NNZ + 3) & ~3;
  const dec_job jb = jobs[blockIdx.x];
  const int Z = jb.Z;
  const int r = threadIdx.x, NT = blockDim.x;

  extern __shared__ __align__(16) uint8_t lds[];
  const int soft_bytes = ((K + lmax) * Z + 15) & ~15;
  int8_t* soft = reinterpret_cast<int8_t*>(lds);
  uint32_t* etab = reinterpret_cast<uint32_t*>(lds + soft_bytes);
  uint32_t* ext = etab + NNZ_PAD;
  uint32_t* red = ext + (lmax - 4) * NT;

  // Lifted edge table: (col*Z) | (shift << 16), shift = V mod Z (ldpc_luts_impl.cpp:4505-4523).
  for (int e = r; e < T::NNZ; e += NT) {
    etab[e] = (uint32_t)(T::col(e) * Z) | ((uint32_t)(T::v(jb.ils, e) % Z) << 16);
  }

  // load_soft_bits (ldpc_decoder_impl.cpp:147-188) and the last nonzero input (:67-71).
  const int8_t* in = in_base + jb.in_off;
  const int n_in = (int)jb.n_in;
  const int full = (n_in / Z) * Z;
  const int cap = (K + lmax - 2) * Z;
  int last = -1;
  for (int i = r; i < n_in; i += NT) {
    const int x = in[i];
    if (x != 0) {
      last = i;
    }
    if (i < cap) {
      soft[2 * Z + i] = (int8_t)(i < full ? min(max(x, -64), 64) : x);
    }
  }
  for (int i = r; i < 2 * Z; i += NT) {
    soft[i] = 0;
  }
  for (int i = n_in + r; i < cap; i += NT) {
    soft[2 * Z + i] = 0;
  }
  const int input_size = block_reduce_max(last, reinterpret_cast<int*>(red)) + 1;

  const int msg = K * Z;
  uint8_t* out = bits_base + jb.bits_off;
  if (input_size < msg && !(jb.flags & kFlagForce)) {
    // Without a CRC calculator the output is set to ones (ldpc_decoder_impl.cpp:75-81).
    if (jb.crc_idx == 0 || (jb.flags & kFlagCrcAfter)) {
      write_hard_bits(soft, out, msg, true);
    }
    if (r == 0) {
      results[jb.out_idx] = 0;
    }
    return;
  }
  int cbl = max(input_size + 2 * Z, msg + 4 * Z);
  cbl = (cbl + Z - 1) / Z * Z;
  const int L = cbl / Z - K;
  if (L > lmax) { // host bound violated: refuse rather than overrun LDS
    if (r == 0) {
      results[jb.out_idx] = -1;
    }
    return;
  }

  for (int l = 4; l < L; ++l) {
    ext[(l - 4) * NT + r] = 0;
  }
  uint32_t st0 = 0, sg0 = 0, st1 = 0, sg1 = 0, st2 = 0, sg2 = 0, st3 = 0, sg3 = 0;

  // CRC weights: hard bit i of the message contributes x^(len-1-i+order) mod G.
  const int len = msg - (int)jb.F;
  const bool use_crc = jb.crc_idx != 0;
  const int order = jb.crc_idx == 3 ? 16 : 24;
  uint32_t xw[K];
#pragma unroll
  for (int n = 0; n < K; ++n) {
    const int i = n * Z + r;
    xw[n] = (use_crc && r < Z && i < len) ? crc_x[(jb.crc_idx - 1) * kCrcXLen + (len - 1 - i + order)] : 0u;
  }
  __syncthreads();

  int result = 0;
  const int max_iter = jb.max_iter;
  const bool es_crc = use_crc && !(jb.flags & kFlagCrcAfter);
  const bool es_syn = !use_crc && (jb.flags & kFlagEsSyndrome);
  for (int it = 0; it < max_iter; ++it) {
    if (r < Z) {
      core_layer<BG, 0>(soft, etab, r, Z, st0, sg0);
    }
    __syncthreads();
    if (r < Z) {
      core_layer<BG, 1>(soft, etab, r, Z, st1, sg1);
    }
    __syncthreads();
    if (r < Z) {
      core_layer<BG, 2>(soft, etab, r, Z, st2, sg2);
    }
    __syncthreads();
    if (r < Z) {
      core_layer<BG, 3>(soft, etab, r, Z, st3, sg3);
    }
    __syncthreads();
    for (int l = 4; l < L; ++l) {
      if (r < Z) {
        const int e0 = T::rs(l);
        uint32_t w = ext[(l - 4) * NT + r];
        uint32_t st = w & 0x7ffffu, sg = w >> 19;
        layer_update<T::DEG_EXT, true>(soft, etab + e0, T::rs(l + 1) - e0, r, Z, st, sg);
        ext[(l - 4) * NT + r] = st | (sg << 19);
      }
      __syncthreads();
    }

    if (es_crc) {
      // get_hard_bits + CRC (ldpc_decoder_impl.cpp:112-119).
      uint32_t x = 0, z = 0;
      if (r < Z) {
#pragma unroll
        for (int n = 0; n < K; ++n) {
          const int s = soft[n * Z + r];
          z |= (s == 0);
          x ^= (s <= 0) ? xw[n] : 0u;
        }
      }
      block_reduce_xor_or(x, z, red);
      if (z == 0 && x == 0) {
        result = it + 1;
        break;
      }
    } else if (es_syn) {
      uint32_t x = 0, z = 0;
      if (r < Z) {
        for (int n = 0; n < K; ++n) {
          z |= (soft[n * Z + r] == 0);
        }
      }
      z |= syndrome_fail<BG>(soft, etab, r, Z, L);
      block_reduce_xor_or(x, z, red);
      if (z == 0) {
        result = it + 1;
        break;
      }
    }
    __syncthreads();
  }

  if (!es_crc && !es_syn) {
    // No early stop: hard bits + syndrome after the last iteration (:130-144), then the
    // codeblock CRC for the PUSCH path without early stop (pusch_codeblock_decoder.cpp:47-57).
    uint32_t x = 0, z = 0;
    if (r < Z) {
      for (int n = 0; n < K; ++n) {
        const int s = soft[n * Z + r];
        z |= (s == 0);
        x ^= (s <= 0) ? xw[n] : 0u;
      }
    }
    z |= syndrome_fail<BG>(soft, etab, r, Z, L);
    block_reduce_xor_or(x, z, red);
    result = (z == 0 && (!use_crc || x == 0)) ? max_iter : 0;
  }
  __syncthreads();
  write_hard_bits(soft, out, msg, false);
  if (r == 0) {
    results[jb.out_idx] = result;
  }
}

// Rate dematcher, behaviour of ldpc_rate_dematcher_impl::rate_dematch
// (ldpc_rate_dematcher_impl.cpp:29-250) in gather form: output position p (not a
// filler, below Ncb) has rank q among the non-filler positions; input e-index i lands
// on rank (start + i) mod P, P = Ncb - F.  New data starts from 0, otherwise from the
// stored soft bit; every input reaching p is added with saturation, in input order.
// Grid: x = position chunks, y = codeblock.
__global__ void __launch_bounds__(256) k_rate_dematch(const rm_job* __restrict__ jobs, const int8_t* __restrict__ cw,
                                                      int8_t* __restrict__ rm)
{
  const rm_job jb = jobs[blockIdx.y];
  const int8_t* in = cw + jb.cw_off;
  int8_t* out = rm + jb.rm_off;
  const uint32_t F = jb.nsys - jb.info;
  const uint32_t P = jb.Ncb - F;
  const uint32_t start = jb.k0 < jb.info ? jb.k0 : (jb.k0 < jb.nsys ? jb.info : jb.k0 - F);
  const uint32_t Kq = jb.E / jb.Qm;
  const bool new_data = jb.flags & kFlagNewData;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < jb.N; p += gridDim.x * blockDim.x) {
    int val;
    if (p >= jb.info && p < jb.nsys) {
      val = 127;
    } else if (p >= jb.Ncb) {
      val = new_data ? 0 : out[p];
    } else {
      const uint32_t q = p < jb.info ? p : p - F;
      val = new_data ? 0 : out[p];
      for (uint32_t i = (q + P - start) % P; i < jb.E; i += P) {
        const int x = in[(i % Kq) * jb.Qm + i / Kq];
        // saturated sum with the infinity rules (log_likelihood_ratio.cpp:22-55)
        if (val == -x) {
          val = 0;
        } else if (llr_isinf(val)) {
        } else if (llr_isinf(x)) {
          val = x;
        } else {
          val = llr_sat(val + x);
        }
      }
    }
    out[p] = (int8_t)val;
  }
}

// CRC remainders of packed bit buffers, one workgroup per buffer.
__global__ void __launch_bounds__(256) k_crc(const uint64_t* __restrict__ offs, const uint32_t* __restrict__ nbits,
                                             const uint8_t* __restrict__ bits, const uint32_t* __restrict__ xtab,
                                             int order, uint32_t* __restrict__ out)
{
  __shared__ uint32_t red[16];
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

// ---- host-side launchers ------------------------------------------------------------
size_t decode_lds_bytes(int bg, int maxZ, int lmax, int nt)
{
  const int K = bg == 1 ? 22 : 10;
  const int nnz_pad = bg == 1 ? 316 : 200;
  return (size_t)(((K + lmax) * maxZ + 15) & ~15) + (size_t)nnz_pad * 4 + (size_t)(lmax - 4) * nt * 4 + 16 * 4;
}

hipError_t launch_decode(int bg, const dec_job* d_jobs, unsigned n, int maxZ, int lmax, const int8_t* in,
                         uint8_t* bits, int32_t* results, const uint32_t* crc_x, hipStream_t s)
{
  const int nt = ((maxZ + 63) / 64) * 64;
  const size_t lds = decode_lds_bytes(bg, maxZ, lmax, nt);
  if (bg == 1) {
    hipError_t e = hipFuncSetAttribute((const void*)k_ldpc_decode<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) {
      return e;
    }
    hipLaunchKernelGGL(k_ldpc_decode<1>, dim3(n), dim3(nt), lds, s, d_jobs, in, bits, results, crc_x, lmax);
  } else {
    hipError_t e = hipFuncSetAttribute((const void*)k_ldpc_decode<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) {
      return e;
    }
    hipLaunchKernelGGL(k_ldpc_decode<2>, dim3(n), dim3(nt), lds, s, d_jobs, in, bits, results, crc_x, lmax);
  }
  return hipGetLastError();
}

hipError_t launch_rate_dematch(const rm_job* d_jobs, unsigned n, unsigned maxN, const int8_t* cw, int8_t* rm,
                               hipStream_t s)
{
  const unsigned chunks = (maxN + 1023) / 1024;
  hipLaunchKernelGGL(k_rate_dematch, dim3(chunks, n), dim3(256), 0, s, d_jobs, cw, rm);
  return hipGetLastError();
}

hipError_t launch_crc(const uint64_t* d_offs, const uint32_t* d_nbits, unsigned n, const uint8_t* bits,
                      const uint32_t* xtab, int order, uint32_t* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_crc, dim3(n), dim3(256), 0, s, d_offs, d_nbits, bits, xtab, order, out);
  return hipGetLastError();
}

} // namespace lh
