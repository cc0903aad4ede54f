// Shared host/device definitions of the HIP LDPC receive chain (not part of the C ABI).
#pragma once
#include <stdint.h>

namespace lh {

constexpr int kMaxZ = 384;
constexpr int kMaxThreads = 384;
// x^e mod G tables: e in [0, 22*384 + 24).
constexpr int kCrcXLen = 22 * 384 + 24;
constexpr uint32_t kMaxRmLength = 22 * 384 * 35; // MAX_CODEBLOCK_RM_SIZE

// job flags
constexpr uint8_t kFlagForce = 1;      // force_decoding
constexpr uint8_t kFlagEsSyndrome = 2; // early_stop_syndrome (no CRC)
constexpr uint8_t kFlagCrcAfter = 4;   // PUSCH without early stop: syndrome, then CRC
constexpr uint8_t kFlagNewData = 8;    // rate dematch: copy instead of combine

// One codeblock decode, as seen by the decoder kernel (48 bytes).
struct alignas(16) dec_job {
  uint64_t in_off;   // decoder input (rate-dematched soft bits) byte offset
  uint64_t bits_off; // packed output byte offset
  uint32_t n_in;     // number of input soft bits
  uint16_t Z;
  uint8_t ils;       // lifting set index 0..7
  uint8_t max_iter;
  uint16_t F;        // filler bits
  uint8_t crc_idx;   // 0 none, 1 CRC24A, 2 CRC24B, 3 CRC16
  uint8_t flags;
  uint32_t lbound;   // upper bound of the layer count for this job (4..M)
  uint32_t out_idx;  // index of this codeblock's result
  uint64_t pad1;
};
static_assert(sizeof(dec_job) == 48, "dec_job layout");

// One codeblock rate dematch (48 bytes).
struct alignas(16) rm_job {
  uint64_t cw_off;   // E LLRs, interleaved by Qm
  uint64_t rm_off;   // N soft bits
  uint32_t E;
  uint32_t N;        // N_short * Z
  uint32_t Ncb;      // circular buffer length
  uint32_t k0;
  uint32_t info;     // (K_b-2)Z - F: first filler position
  uint32_t nsys;     // (K_b-2)Z
  uint16_t Qm;
  uint8_t flags;
  uint8_t pad[5];
};
static_assert(sizeof(rm_job) == 48, "rm_job layout");

inline int crc_index(uint32_t poly)
{
  return poly == 0x1864cfbu ? 1 : poly == 0x1800063u ? 2 : poly == 0x11021u ? 3 : (poly == 0 ? 0 : -1);
}

} // namespace lh
