// Shared host/device definitions of the HIP LDPC receive chain (not part of the C ABI).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "ocudu_hip_ldpc.h"

namespace lh {

constexpr int kMaxZ = 384;
constexpr int kMaxThreads = 384;
// Devices whose per-kernel launch attributes the library caches (others set them per launch).
constexpr int kMaxCachedDevices = 16;
// x^e mod G tables: e in [0, 22*384 + 24).
constexpr int kCrcXLen = 22 * 384 + 24;
// The decoder's CRC weights: per polynomial kCrcPad zeros, then x^e mod G for e in
// [0, kCrcXLen) with the entries e < order zeroed, so that the weight index of a filler
// bit (at least order - F > -kCrcPad) reads zero.
constexpr int kCrcPad = 20 * 384;
constexpr int kCrcXZLen = kCrcPad + kCrcXLen;
// The TB CRC moves a chunk's remainder past the m bytes that follow it with one table
// factor per base-16 digit of m: m < 16^kTbShiftDigits (16 MiB, above any TB).
constexpr int kTbShiftDigits = 6;
constexpr uint32_t kMaxRmLength = 22 * 384 * 35; // MAX_CODEBLOCK_RM_SIZE
constexpr int kNofLiftingSizes = 51;
// Lifted edge tables of all graphs: BG1 graphs first (316 edges each), then BG2 (197).
constexpr int kGraphTableLen = kNofLiftingSizes * (316 + 197);
// Words after the tables (plain, row pairs, one row): an extension layer loads the edge words
// of its degree class's largest row at once, up to 2 x 9 words past a row's last edge.
constexpr int kGraphTablePad = 32;

// job flags
constexpr uint8_t kFlagForce = 1;      // force_decoding
constexpr uint8_t kFlagEsSyndrome = 2; // early_stop_syndrome (no CRC)
constexpr uint8_t kFlagCrcAfter = 4;   // PUSCH without early stop: syndrome, then CRC
constexpr uint8_t kFlagNewData = 8;    // rate dematch: copy instead of combine
constexpr uint8_t kFlagFused = 16;     // decoder dematches its own input from the codeword LLRs
constexpr uint8_t kFlagHarqWrite = 32; // fused: write the combined soft bits back (HARQ buffer)
constexpr uint8_t kFlagSkipPassed = 64; // fallback pass: leave codeblocks whose result is > 0
constexpr uint8_t kFlagPassed = 128;    // CRC passed in an earlier transmission: combine the new LLRs into
                                        // the HARQ buffer only; the result is `preset`, the bits stay

// Learned normalisation tables: kLearnedIters rows of M words (A | B << 16) per base
// graph, BG1 (46 layers) first, then BG2 (42).
constexpr int kLearnedIters = 8;
constexpr int kScaleTableWords = kLearnedIters * (46 + 42);

// Rate-dematching parameters of one codeblock (64 bytes).
struct alignas(16) rm_params {
  uint64_t cw_off; // E LLRs, interleaved by Qm
  uint64_t rm_off; // N soft bits (HARQ buffer)
  uint32_t E;
  uint32_t N;      // N_short * Z
  uint32_t Ncb;    // circular buffer length
  uint32_t start;  // rank of k0 among the non-filler positions
  uint32_t info;   // (K_b-2)Z - F: first filler position
  uint32_t nsys;   // (K_b-2)Z
  uint32_t P;      // Ncb - F: non-filler positions of the circular buffer
  uint32_t Kq;     // E / Qm
  float rcp_kq;    // 1 / Kq
  uint16_t Qm;
  uint8_t flags;
  uint8_t pad0;
  uint64_t pad1;
};
static_assert(sizeof(rm_params) == 64, "rm_params layout");

#if defined(__HIP__)
#define LH_HOST_DEVICE __host__ __device__
#else
#define LH_HOST_DEVICE
#endif

// Parity rows (of M) the word-parallel encoder computes for one codeblock (K_b systematic
// columns, lifting size Z): all of them with the codeword output or when the rate
// matcher's ranks [start, start + E) wrap; otherwise the rows up to the column of the
// last position it reads (past the fillers), at least the four core rows.  The host
// sizes the launch's LDS from it.
LH_HOST_DEVICE inline int encode_parity_rows(const rm_params& rm, int K, int M, int Z, bool want_cw)
{
  if (want_cw || rm.start + rm.E > rm.P) {
    return M;
  }
  const uint32_t last = rm.start + rm.E - 1u;
  const uint32_t pos = last < rm.info ? last : last + (rm.nsys - rm.info);
  const int rows = (int)(pos / (uint32_t)Z) + 2 - K + 1;
  return rows < 4 ? 4 : (rows > M ? M : rows);
}

// One codeblock decode, as seen by the decoder kernel (112 bytes).
struct alignas(16) dec_job {
  uint64_t in_off;   // non-fused: decoder input (rate-dematched soft bits) byte offset
  uint64_t bits_off; // packed output byte offset
  uint32_t n_in;     // number of input soft bits (N when fused)
  uint32_t gt_off;   // offset of this graph's lifted edge table
  uint16_t Z;
  uint8_t ils;       // lifting set index 0..7
  uint8_t max_iter;
  uint16_t F;        // filler bits
  uint8_t crc_idx;   // 0 none, 1 CRC24A, 2 CRC24B, 3 CRC16
  uint8_t flags;
  uint16_t lbound;   // upper bound of the layer count for this job (4..M)
  uint16_t preset;   // kFlagPassed: the result reported (the earlier iteration count)
  uint32_t out_idx;  // index of this codeblock's result
  rm_params rm;      // fused dematching parameters
};
static_assert(sizeof(dec_job) == 112, "dec_job layout");

// Work item of the persistent decoder: `ngroups` consecutive jobs with the same (BG, Z);
// group g runs on threads [g*n, (g+1)*n) of the workgroup, n = Z (one check row per
// thread) or Z/2 (rows2: rows p and p + Z/2 in the two 16-bit lanes of a register).
constexpr int kMaxGroups = 64;
struct alignas(16) dec_item {
  uint32_t first_job;
  uint16_t ngroups;
  uint8_t bg;
  uint8_t rows2; // 1: two check rows per thread (even Z)
  uint16_t lmax;  // layer bound of the item (max over its jobs)
  uint16_t scatter; // 1: every job is a fused new-data load with E <= P and Z % 8 == 0 (scatter load)
  uint32_t stage; // LDS bytes per group for the codeword LLRs (0: read them from HBM)
};
static_assert(sizeof(dec_item) == 16, "dec_item layout");

// Header of the device descriptor block of one decoder launch.
// Items [0, grid) go to workgroup w = item (the first round, largest items first); the
// rest are nq interleaved sub-queues (sub-queue q: items grid + q + k nq) that the
// workgroups w = q (mod nq) claim through counter q, kCounterStride words apart from
// ctr_off on (words from this header), zero at upload.
constexpr uint32_t kMaxSubQueues = 16;
constexpr uint32_t kCounterStride = 32; // one 128-byte line per counter
struct alignas(16) launch_header {
  uint32_t nq;         // sub-queues of the dynamic part (0: no dynamic part)
  uint32_t n_items;    // items of this queue: [first_item, first_item + n_items)
  uint32_t first_item;
  uint32_t ctr_off;
};

// Static scratch of the allocation-free decode entry point: one descriptor block (four
// headers, at most one item and one job per codeblock) of up to kNoAllocMaxCbs codeblocks.
constexpr unsigned kNoAllocMaxCbs = 4096;
constexpr size_t kScratchBytes = ((4 * sizeof(launch_header) + 4 * kMaxSubQueues * kCounterStride * 4 +
                                   kNoAllocMaxCbs * (sizeof(dec_item) + sizeof(dec_job))) + 255) & ~(size_t)255;

// One synthetic codeblock of the generator (128 bytes): message bits from seed_msg,
// noise from seed_noise, rate matching by `rm` (cw_off = LLR output offset).
struct alignas(16) gen_job {
  uint64_t msg_off;    // packed message bits output (K_b*Z bits)
  uint64_t seed_msg;
  uint64_t seed_noise;
  double n0;           // noise variance per complex symbol (Es = 1)
  uint32_t gt_off;
  uint16_t Z;
  uint8_t bg;
  uint8_t crc_idx;     // 0 none, 1 CRC24A, 2 CRC24B, 3 CRC16
  uint16_t F;
  uint8_t write_msg;
  uint8_t msg_input; // the message is read from msgs + msg_off (data, CRC and fillers)
  uint8_t want_cw;   // encoder: also the codeword (one bit per byte) at cwb_off
  uint8_t pad0;
  uint16_t pad1;
  uint64_t cwb_off;
  uint32_t pad2[2];
  rm_params rm;      // encoder: rm.cw_off is the packed rate-matched output
};
static_assert(sizeof(gen_job) == 128, "gen_job layout");

// One transport block of the TB generator (same layout as lh_gen_tb_desc).
struct alignas(8) gen_tb_job {
  uint64_t tb_off;
  uint64_t msg_off;
  uint64_t seed;
  uint32_t msg_stride;
  uint32_t tbs_bits;
  uint32_t nof_cbs;
  uint32_t segment_length;
  uint32_t nof_filler_bits;
  uint32_t pad;
};
static_assert(sizeof(gen_tb_job) == 48, "gen_tb_job layout");

// One codeblock of a learned-decoder training batch (48 bytes).
struct alignas(16) train_job {
  uint64_t in_off;  // int8 rate-dematched soft bits (n_in of them)
  uint64_t tgt_off; // packed message bits (the targets)
  uint64_t ws_off;  // workspace of this codeblock, in floats
  uint32_t n_in;
  uint32_t gt_off;
  uint16_t Z;
  uint16_t L;       // layers decoded
  uint16_t F;
  uint16_t pad0;
  uint32_t pad1[2];
};
static_assert(sizeof(train_job) == 48, "train_job layout");

// Transport-block descriptor as seen by the device (same layout as lh_tb_desc).
struct alignas(8) lh_tb_desc_dev {
  uint64_t tb_offset;
  uint64_t cb_bits_offset;
  uint32_t cb_bits_stride;
  uint32_t tbs_bits;
  uint32_t first_cb;
  uint32_t nof_cbs;
  uint32_t cb_data_bits;
  uint32_t pad;
};
static_assert(sizeof(lh_tb_desc_dev) == 40, "lh_tb_desc_dev layout");

// Completion signal of a TB finalize launch whose outputs live in host memory: every
// workgroup also copies its transport block's codeblock results to res_out, and the last
// workgroup to finish stores seq into *flag (system scope, after all the launch's
// writes), so the host sees the whole launch complete by reading one word.  The TB
// offsets of such a launch are multiples of 4 (the bytes go out as words).
struct tb_signal {
  int32_t* res_out;  // host-visible codeblock results, indexed like the decoder's
  uint32_t* counter; // device word, zero between launches (the last workgroup rewinds it)
  uint32_t* flag;    // host-visible completion word
  uint32_t seq;
  uint32_t pad;
};

// TB finalize inside the decoder launch (the serving runtime): every workgroup, after an
// item, counts its codeblocks off their transport blocks' tb_left; the workgroup that
// counts a transport block's last codeblock reassembles and checks it at once (as one
// workgroup of the TB finalize launch would, with the same signal).  tbs == nullptr: off.
struct fused_tb_args {
  const lh_tb_desc_dev* tbs; // the launch's transport blocks
  const uint32_t* cb_tb;     // per codeblock (decoder result index): its transport block
  uint32_t* tb_left;         // per transport block: codeblocks not yet counted (nof_cbs at upload)
  uint8_t* tb_out;           // TB bytes (host-visible)
  int32_t* status;           // TB verdicts (host-visible)
  const uint32_t* lut24a;    // CRC24A byte table, then x^(2^k) and the byte-shift digits
  tb_signal sig;
  uint32_t n_tbs;
  uint32_t pad;
};
// LDS the TB finalize of a fused launch uses at the start of the decoder's dynamic LDS:
// the CRC tables (352 words), the reduction words (32), one flag, the count and the list
// of transport blocks to finalize (up to kMaxGroups).
constexpr size_t kFusedTbLdsBytes = 4 * (352 + 32 + 1 + 1 + 64);

// Demodulation + descrambling: one block (a span of symbols with its scrambling
// sequence start) and one tile of 256 of its symbols per workgroup.
constexpr int kDemodJumps = 20;   // A^(32 * 2^j), j < 20: blocks up to 2^25 sequence bits
constexpr int kDemodFloats = 163; // slope/intercept tables + 1/sqrt(42), 1/sqrt(170), 1/sqrt(10)
constexpr int kDemodTableWords = 2 * kDemodJumps * 31 + kDemodFloats;
struct demod_job {
  uint64_t sym_off;     // first symbol (and noise variance)
  uint64_t llr_off;     // first LLR
  uint32_t nof_symbols;
  uint32_t modulation;  // 0 pi/2-BPSK, 1 BPSK, 2 QPSK, 4 16-QAM, 6 64-QAM, 8 256-QAM
  uint32_t pad[2];
};
struct demod_tile {
  uint32_t job;
  uint32_t tile;
  uint32_t x1, x2; // LFSR states at the tile's first sequence bit (Nc included)
};
// LFSR states at the first sequence bit of each of a block's tiles.
void demod_tile_states(uint32_t c_init, uint64_t pos, int qm, uint32_t nof_tiles, uint32_t* x1, uint32_t* x2);
void build_demod_tables(uint32_t* out);

// UL-SCH demultiplexing: one record per OFDM symbol of a codeword (TS38.212 6.2.7 field
// spacings of that symbol), one tile of kDemuxTileUnits units of `unit` bytes of one symbol
// per workgroup.  Field sets of a symbol with M REs (RE indices 0..M-1), each "every d-th
// element of a base set, n of them":
//   R  reserved (HARQ-ACK <= 2 bits): every dr-th of all REs
//   A  HARQ-ACK > 2 bits: every da-th of all REs
//   C1 CSI Part 1: every d1-th of the REs in neither A nor R
//   C2 CSI Part 2: every d2-th of the REs in neither A nor C1
//   S  HARQ-ACK <= 2 bits: every ds-th of R (S stays in UL-SCH / C2 with zero soft bits)
//   UL-SCH: the REs in none of A, C1, C2.
constexpr int kDemuxThreads = 256;
constexpr int kDemuxTileUnits = 256;
struct demux_sym {
  uint64_t llr_in;  // first soft bit of the symbol in d_llrs
  uint64_t sch_out; // first UL-SCH soft bit of the symbol in d_sch
  uint64_t ack_out; // and of its HARQ-ACK, CSI Part 1 and CSI Part 2 soft bits
  uint64_t c1_out;
  uint64_t c2_out;
  uint32_t M;       // REs of the symbol
  uint32_t x1, x2;  // LFSR states at the symbol's first soft bit (Nc included)
  uint32_t lcm_ar;  // A and R share the multiples of lcm(da, dr) below lim_ar
  uint32_t lim_ar;
  uint16_t nbpr;    // soft bits per RE (bits per symbol x layers)
  uint8_t qm;
  uint8_t unit;     // bytes per thread: 1, 2 or 4
  uint8_t ph_ack, ph_c1, ph_c2; // payload bits (1 or 2) of a field with placeholders, else 0
  uint8_t plain;    // no field but UL-SCH in the symbol
  uint16_t dr, nr, da, na, d1, n1, d2, n2, ds, ns;
};
struct demux_tile {
  uint32_t sym;
  uint32_t tile;
};
// Per-symbol records of one codeword; returns 0, or -1 with *why set.  syms may be null
// (validation and sizes only); it receives at most 14 records.
int demux_plan(const lh_ulsch_demux_desc& d, demux_sym* syms, unsigned* nof_syms, lh_ulsch_demux_sizes* sizes,
               const char** why);
// get_ulsch_information of a grant; returns 0, or -1 with *why set.
int ulsch_information(const lh_ulsch_config& c, lh_ulsch_info& o, const char** why);

inline int crc_index(uint32_t poly)
{
  return poly == 0x1864cfbu ? 1 : poly == 0x1800063u ? 2 : poly == 0x11021u ? 3 : (poly == 0 ? 0 : -1);
}

} // namespace lh
