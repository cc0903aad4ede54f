/*
 * CPU oracle for the PUSCH LDPC receive chain -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's CPU algorithms, used by tests/, the smoke
 * check and the cpu_baseline leg of bench.py as the checker.  Nothing in the product
 * (the HIP library under ocudu-dapp-platform_amd/) links or calls this code.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * reference repository root).  Parity pins (see tests/test_oracle_*.py):
 *   - LLR arithmetic: log_likelihood_ratio_test.cpp KATs + exhaustive 256x256 vs _ref.
 *   - CRC: bitwise long division closed form + _ref.
 *   - Base graph: nnz/degree invariants, H*c = 0 on the encoder output, _ref dump.
 *   - Rate dematcher: the reference test's golden restatement sweep, run against _ref.
 *   - Decoder: trajectory vs the reference generic decoder built from source (_ref),
 *     noise-free convergence in one iteration, short/zero input rules.
 *   - Encoder / rate matcher: H*c = 0, and bit-exact vs _ref.
 */
#ifndef LDPC_ORACLE_H
#define LDPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LLR_MAX 120
#define ORC_LLR_INF 127

/* ---- a1: log_likelihood_ratio arithmetic -------------------------------------- */
int8_t orc_llr_add(int8_t a, int8_t b);
int8_t orc_llr_promotion_sum(int8_t a, int8_t b);
int8_t orc_llr_quantize(float value, float range_limit);

/* ---- a9: CRC (MSB-first long division, zero init, no reflection, no final xor) --- */
#define ORC_CRC24A 0x1864cfbu
#define ORC_CRC24B 0x1800063u
#define ORC_CRC16 0x11021u
uint32_t orc_crc_packed(uint32_t poly, const uint8_t* bits_msb_first, unsigned nbits);
uint32_t orc_crc_unpacked(uint32_t poly, const uint8_t* bits, unsigned nbits);

/* ---- a11: lifted base graph ------------------------------------------------------ */
int orc_lifting_index(unsigned Z);    /* iLS 0..7 or -1 */
int orc_lifting_position(unsigned Z); /* 0..50 or -1 */
unsigned orc_lifting_size_at(unsigned pos);
int orc_bg_info_nodes(int bg);  /* K_b: 22 / 10 */
int orc_bg_check_nodes(int bg); /* M: 46 / 42 */
/* Shift of the lifted edge (row, col), or -1 if there is no edge. */
int orc_graph_shift(int bg, unsigned Z, unsigned row, unsigned col);

/* ---- a4..a10: layered normalized min-sum decoder ------------------------------- */
typedef struct {
  int bg;              /* 1 or 2 */
  unsigned Z;          /* lifting size */
  unsigned nof_filler_bits;
  unsigned nof_crc_bits; /* 16 or 24 */
  unsigned max_iterations;
} orc_dec_cfg;

/* Returns the number of iterations on success, 0 for "no value" (std::nullopt).
 * crc_poly == 0 means no CRC calculator.  out_packed holds K_b*Z bits, MSB first.
 * If stats_iters != NULL it receives the number of iterations actually run. */
unsigned orc_ldpc_decode(const orc_dec_cfg* cfg, int force_decoding, int early_stop_syndrome, uint32_t crc_poly,
                         const int8_t* llrs, unsigned nof_llrs, uint8_t* out_packed);
/* The same decoder with learned per-(iteration, layer) normalisation: tab holds 8 rows of
 * M words A | B << 16, B signed (row min(it, 7) serves iteration it); the 0.8 scaling
 * becomes min(120, max(0, floor((A m - B + 128) / 256))). */
unsigned orc_ldpc_decode_learned(const orc_dec_cfg* cfg, const uint32_t* tab, uint32_t crc_poly, const int8_t* llrs,
                                 unsigned nof_llrs, uint8_t* out_packed);

/* Loss of the learned decoder's floating-point training model (float and double
 * arithmetic): T iterations of L layers, tables alpha/beta [T][M], iteration weights w,
 * logit scale tau; msg = packed message bits (targets). */
double orc_nms_forward_f(int bg, unsigned Z, unsigned F, unsigned L, unsigned T, const int8_t* in, unsigned n_in,
                         const uint8_t* msg, const double* alpha, const double* beta, const double* w, double tau);
double orc_nms_forward_d(int bg, unsigned Z, unsigned F, unsigned L, unsigned T, const int8_t* in, unsigned n_in,
                         const uint8_t* msg, const double* alpha, const double* beta, const double* w, double tau);

/* ---- a3: rate dematcher ---------------------------------------------------------- */
typedef struct {
  unsigned rv;        /* 0..3 */
  unsigned Qm;        /* bits per symbol: 1,2,4,6,8 */
  unsigned Nref;      /* 0 = no limit */
  unsigned nof_filler_bits;
} orc_rm_cfg;

/* out has block_length = N_short*Z entries (BG and Z inferred as in the reference). */
int orc_rate_dematch(int8_t* out, unsigned block_length, const int8_t* in, unsigned E, int new_data,
                     const orc_rm_cfg* cfg);

/* ---- a2: receive segmentation ---------------------------------------------------- */
typedef struct {
  unsigned nof_segments;
  unsigned Z;
  unsigned segment_length;  /* K_b*Z */
  unsigned nof_filler_bits; /* per CB */
  unsigned nof_crc_bits;    /* per CB CRC bits: TB CRC when C==1, else 24 */
  unsigned nof_tb_crc_bits;
  unsigned full_length;     /* N_short*Z */
  unsigned rm_length[256];
  unsigned cw_offset[256];
} orc_segmentation;

int orc_segment_rx(orc_segmentation* seg, int bg, unsigned tbs_bits, unsigned nof_ch_symbols, unsigned Qm,
                   unsigned nof_layers);
unsigned orc_nof_codeblocks(unsigned tbs_bits, int bg);
unsigned orc_compute_lifting_size(unsigned tbs_bits, int bg, unsigned nof_segments);

/* ---- a14: data generation (encoder, rate matcher, mapper, channel, demapper) ------ */
/* msg: K_b*Z bits (one per byte, fillers must be 0).  cw: (K_b+M)*Z bits, full codeword
 * before the 2Z shortening.  Returns 0 on success. */
int orc_ldpc_encode(int bg, unsigned Z, const uint8_t* msg, uint8_t* cw);
/* Checks H*c = 0 over all M rows; returns the number of unsatisfied lifted checks. */
unsigned orc_syndrome_weight(int bg, unsigned Z, const uint8_t* cw);
/* Rate matching TS38.212 5.4.2: d = transmitted codeword (N_short*Z bits, fillers
 * flagged by filler_start/nof_filler), output E bits, interleaved by Qm. */
int orc_rate_match(uint8_t* out, unsigned E, const uint8_t* d, unsigned N, int bg, unsigned Z, unsigned rv,
                   unsigned Qm, unsigned Nref, unsigned nof_filler_bits);
/* Map bits to symbols (TS38.211 5.1), add AWGN of total variance n0 (seeded), max-log
 * demap with the reference's range limits, quantize to int8. */
void orc_modulate_awgn_demap(int8_t* llrs, const uint8_t* bits, unsigned nbits, unsigned Qm, double n0,
                             uint64_t seed);
void orc_modulate(float* iq, const uint8_t* bits, unsigned nbits, unsigned Qm);
void orc_demap(int8_t* llrs, const float* iq, unsigned nsym, unsigned Qm, float n0);

/* Deterministic generator used by the data generators (splitmix64). */
uint64_t orc_rand_u64(uint64_t* state);
double orc_rand_normal(uint64_t* state);

/* ---- a13: TB reassembly ----------------------------------------------------------- */
/* Concatenates the data bits of each CB (first K_b*Z - F - crc bits, capped by the TB
 * size + TB CRC) into tb_packed (TBS+CRC bits) and returns the TB CRC remainder. */
uint32_t orc_tb_concat_crc(uint8_t* tb_packed, unsigned tbs_bits, const uint8_t* cb_bits_packed,
                           unsigned cb_bits_stride_bytes, const orc_segmentation* seg);

/* Synthetic codeblock (the device generator's statement): message packed, E LLRs. */
int orc_gen_codeblock(int bg, unsigned Z, unsigned F, uint32_t crc_poly, unsigned E, unsigned rv, unsigned Qm,
                      unsigned Nref, double n0, uint64_t seed_msg, uint64_t seed_noise, uint8_t* msg_packed,
                      int8_t* llrs);

/* f1: PUSCH soft demodulation and descrambling. */
void orc_scrambling_sequence(uint8_t* c, uint32_t c_init, unsigned offset, unsigned n);
void orc_demodulate_soft(int8_t* llrs, const float* iq, const float* nv, unsigned nsym, int modulation);
void orc_pusch_demodulate(int8_t* llrs, const float* iq, const float* nv, unsigned nsym, int modulation,
                          uint32_t c_init, unsigned seq_offset);

/* f1: UL-SCH demultiplexing (ulsch_demultiplex::configuration,
 * include/ocudu/phy/upper/channel_processors/pusch/ulsch_demultiplex.h:35-66, with CSI
 * Part 2 given up front). */
typedef struct {
  int modulation;                       /* 0 pi/2-BPSK, 1, 2, 4, 6, 8 bits per symbol */
  unsigned nof_layers;                  /* 1..4 */
  unsigned nof_prb;
  unsigned start_symbol_index;
  unsigned nof_symbols;
  unsigned dmrs_symbol_mask;            /* bit l: OFDM symbol l carries DM-RS */
  unsigned dmrs_type;                   /* 1 or 2 */
  unsigned nof_cdm_groups_without_data; /* 1..2 (type 1), 1..3 (type 2) */
  unsigned nof_harq_ack_rvd;            /* G_rvd^ACK */
  unsigned nof_harq_ack_bits;           /* O_ACK */
  unsigned nof_enc_harq_ack_bits;       /* G^ACK */
  unsigned nof_csi_part1_bits;
  unsigned nof_enc_csi_part1_bits;
  unsigned nof_csi_part2_bits;
  unsigned nof_enc_csi_part2_bits;
} orc_demux_config;

typedef struct {
  uint32_t nof_llrs; /* codeword soft bits in */
  uint32_t nof_sch;
  uint32_t nof_harq_ack;
  uint32_t nof_csi_part1;
  uint32_t nof_csi_part2;
} orc_demux_sizes;

int orc_ulsch_demultiplex(const orc_demux_config* cfg, const int8_t* llrs, uint32_t c_init, int8_t* sch,
                          int8_t* harq_ack, int8_t* csi1, int8_t* csi2, orc_demux_sizes* sizes);
int orc_ulsch_positions(const orc_demux_config* cfg, int32_t* sch, int32_t* harq_ack, int32_t* csi1,
                        int32_t* csi2);

/* UL-SCH bit accounting: ulsch_configuration / ulsch_information / sch_information
 * (include/ocudu/ran/pusch/ulsch_info.h, include/ocudu/ran/sch/sch_segmentation.h). */
typedef struct {
  unsigned base_graph;
  unsigned tb_crc_size;
  unsigned nof_cb;
  unsigned lifting_size;
  unsigned nof_bits_per_cb;
  unsigned nof_filler_bits_per_cb;
} orc_sch_info_t;

typedef struct {
  unsigned tbs_bits;
  float target_code_rate; /* normalised R */
  int modulation;         /* bits per symbol, 0 pi/2-BPSK */
  unsigned nof_harq_ack_bits;
  unsigned nof_csi_part1_bits;
  unsigned nof_csi_part2_bits;
  float alpha_scaling;
  float beta_offset_harq_ack;
  float beta_offset_csi_part1;
  float beta_offset_csi_part2;
  unsigned nof_rb;
  unsigned start_symbol_index;
  unsigned nof_symbols;
  unsigned dmrs_type;
  unsigned dmrs_symbol_mask;
  unsigned nof_cdm_groups_without_data;
  unsigned nof_layers;
  int contains_dc;
} orc_ulsch_config;

typedef struct {
  orc_sch_info_t sch;
  unsigned nof_ul_sch_bits;
  unsigned nof_harq_ack_bits;
  unsigned nof_harq_ack_rvd;
  unsigned nof_csi_part1_bits;
  unsigned nof_csi_part2_bits;
  unsigned nof_harq_ack_re;
  unsigned nof_csi_part1_re;
  unsigned nof_csi_part2_re;
  unsigned nof_dc_overlap_bits;
} orc_ulsch_info;

void orc_sch_info(unsigned tbs, float R, orc_sch_info_t* out);
int orc_ulsch_information(const orc_ulsch_config* cfg, orc_ulsch_info* info);

#ifdef __cplusplus
}
#endif

#endif
