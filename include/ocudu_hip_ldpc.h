/*
 * ocudu_hip_ldpc -- C ABI of the MI355X (gfx950) PUSCH LDPC receive chain.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes (0 = success,
 * negative = error, message in lh_last_error()).  No exceptions cross this boundary.
 * Device pointers ("d_*") are HIP device allocations; "stream" is a hipStream_t (NULL =
 * the null stream).  A context is single-threaded: one caller at a time, like the
 * reference's decoder instances (pusch_codeblock_decoder_cuda_batch.h decode_mutex_).
 *
 * Every entry point replaces one reference interface (paths relative to the reference
 * repository root); the C++ adapters in ocudu-dapp-platform_amd/host implement the
 * reference's class interfaces on top of these functions.
 */
#ifndef OCUDU_HIP_LDPC_H
#define OCUDU_HIP_LDPC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LH_OK 0
#define LH_ERR_INVALID (-1)  /* invalid argument (a contract break in the reference) */
#define LH_ERR_HIP (-2)      /* HIP runtime error */
#define LH_ERR_CAPACITY (-3) /* batch exceeds the context capacity */

/* CRC generator polynomials, include/ocudu/phy/upper/channel_coding/crc_calculator.h:17-24. */
#define LH_CRC24A 0x1864cfbu
#define LH_CRC24B 0x1800063u
#define LH_CRC16 0x11021u

/* Decoder flags: ldpc_decoder_factory_configuration{force_decoding, early_stop_syndrome}
 * (include/ocudu/phy/upper/channel_coding/channel_coding_factories.h:35-54). */
#define LH_FORCE_DECODING 1u
#define LH_EARLY_STOP_SYNDROME 2u

typedef struct lh_ctx lh_ctx;

/* Context: device tables, staging for descriptors, capacity.  Replaces the device-side
 * state of pusch_codeblock_decoder_cuda_batch (constructed from pusch_decoder_capacity,
 * lib/phy/upper/channel_coding/ldpc/cuda/pusch_decoder_capacity.h:14-45).
 * max_codeblocks bounds one batch call. */
int lh_ctx_create(lh_ctx** ctx, int device, unsigned max_codeblocks);
int lh_ctx_destroy(lh_ctx* ctx);
/* Message of the last error on the calling thread (never NULL). */
const char* lh_last_error(void);

/* ldpc_decoder::configuration, include/ocudu/phy/upper/channel_coding/ldpc/ldpc_decoder.h:23-34. */
typedef struct lh_decoder_config {
  uint32_t base_graph;      /* 1 = BG1, 2 = BG2 */
  uint32_t lifting_size;    /* one of the 51 NR lifting sizes */
  uint32_t nof_filler_bits; /* F */
  uint32_t nof_crc_bits;    /* 16 or 24 */
  uint32_t max_iterations;  /* 1..255 */
} lh_decoder_config;

/* One ldpc_decoder::decode() call of a batch. */
typedef struct lh_decode_desc {
  uint64_t llr_offset;  /* byte offset of this codeblock's LLRs in d_llrs */
  uint64_t bits_offset; /* byte offset of its packed output (K_b*Z bits, MSB first) */
  uint32_t nof_llrs;    /* input length: (K_b+2)*Z .. N_short*Z */
  uint32_t crc_poly;    /* LH_CRC24A / LH_CRC24B / LH_CRC16, or 0 = no CRC calculator */
  lh_decoder_config cfg;
} lh_decode_desc;

/* Batched ldpc_decoder::decode(output, input, crc, cfg)
 * (include/ocudu/phy/upper/channel_coding/ldpc/ldpc_decoder.h:52-57, behaviour of
 * lib/phy/upper/channel_coding/ldpc/ldpc_decoder_impl.cpp:41-145 with the generic
 * arithmetic of ldpc_decoder_generic.cpp).  d_results[i] = number of iterations, or 0
 * for std::nullopt.  Bit-exact to the reference generic decoder.  flags: LH_FORCE_*.
 * descs is host memory; d_llrs, d_bits, d_results are device memory. */
int lh_ldpc_decode(lh_ctx* ctx, const lh_decode_desc* descs, unsigned n, const int8_t* d_llrs, uint8_t* d_bits,
                   int32_t* d_results, uint32_t flags, void* stream);

/* codeblock_metadata of one codeblock (include/ocudu/phy/upper/codeblock_metadata.h:24-58)
 * plus its buffers. */
typedef struct lh_rm_desc {
  uint64_t cw_offset;   /* cb_specific.cw_offset: its E LLRs in the codeword buffer */
  uint64_t rm_offset;   /* offset of its N_short*Z soft-bit (HARQ) buffer */
  uint32_t rm_length;   /* cb_specific.rm_length (E), multiple of modulation_order */
  uint32_t base_graph;  /* tb_common.base_graph: 1 / 2 */
  uint32_t lifting_size;
  uint32_t rv;          /* 0..3 */
  uint32_t modulation_order; /* bits per symbol: 1, 2, 4, 6, 8 */
  uint32_t Nref;        /* 0 = no limited buffer */
  uint32_t nof_filler_bits;
} lh_rm_desc;

/* Batched ldpc_rate_dematcher::rate_dematch(output, input, new_data, cfg)
 * (include/ocudu/phy/upper/channel_coding/ldpc/ldpc_rate_dematcher.h:36-39, behaviour of
 * lib/phy/upper/channel_coding/ldpc/ldpc_rate_dematcher_impl.cpp:29-250): deinterleave,
 * place from k0 skipping fillers, fillers := +127, copy or saturated HARQ combine.
 * d_rm is read (when !new_data) and written in place. */
int lh_rate_dematch(lh_ctx* ctx, const lh_rm_desc* descs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                    int new_data, void* stream);

/* One codeblock of a PUSCH batch: described_rx_codeblock = (LLR view, codeblock_metadata)
 * (codeblock_metadata.h:144) plus its HARQ buffer and output placement. */
typedef struct lh_pusch_cb_desc {
  lh_rm_desc rm;        /* dematching parameters and buffers */
  uint64_t bits_offset; /* packed message bits of this codeblock */
  uint32_t nof_crc_bits;/* cb_specific.nof_crc_bits */
  uint32_t crc_poly;    /* CB CRC: 24B when C > 1, else the TB CRC (24A or 16) */
} lh_pusch_cb_desc;

/* pusch_resident_codeblock_decoder::decode_batch(codeblocks, cb_data, rm_buffers,
 * new_data, crc_poly, use_early_stop, nof_ldpc_iterations)
 * (include/ocudu/phy/upper/channel_processors/pusch/pusch_resident_codeblock_decoder.h:62-149)
 * with the per-codeblock semantics of pusch_codeblock_decoder::decode
 * (lib/phy/upper/channel_processors/pusch/pusch_codeblock_decoder.cpp:17-58): rate
 * dematch into the HARQ buffer d_rm (combined soft bits are written back), then decode;
 * with use_early_stop the CRC stops the iterations, otherwise the syndrome gates a final
 * CRC.  d_results[i] = iterations or 0 (CRC failed).  Mixed BG/Z/E/rate in one batch. */
int lh_pusch_decode_batch(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                          uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                          unsigned nof_ldpc_iterations, void* stream);

/* Learned min-sum normalisation table of one base graph (the lh_set_scale_table of the
 * hot-path design; the reference's fixed counterpart is the 0.8 scaling factor of
 * ldpc_decoder_impl.h:185 applied by ldpc_decoder_generic.cpp:52-61).  Entry (t, l),
 * t < iters, l < layers (46 for BG1, 42 for BG2): the check-to-variable magnitude of
 * layer l in iteration t is min(120, max(0, floor((alpha_q * m - beta_q + 128) / 256)))
 * for the min-sum magnitude m, i.e. round(alpha m - beta) with alpha = alpha_q / 256
 * (alpha_q <= 511) and beta = beta_q / 256.  Iterations from `iters` on use the last row.
 * Do not call while a learned decode on this context is in flight. */
int lh_set_scale_table(lh_ctx* ctx, unsigned base_graph, const uint16_t* alpha_q, const int16_t* beta_q,
                       unsigned iters, unsigned layers);

/* lh_pusch_decode_batch with the learned decoder, then the plain decoder as fallback:
 * every codeblock is decoded with the learned tables (at most nof_ldpc_iterations); those
 * whose CRC did not pass are decoded again by the plain (reference-exact) decoder from the
 * soft bits the first pass wrote to their HARQ buffers (at most nof_fallback_iterations,
 * 0 = no fallback).  d_results[i] = iterations of the pass that decoded it, or 0. */
int lh_pusch_decode_batch_learned(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw,
                                  int8_t* d_rm, uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                                  unsigned nof_ldpc_iterations, unsigned nof_fallback_iterations, void* stream);

/* lh_pusch_decode_batch without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * as lh_pusch_decode_batch; the launch's descriptor block (work-queue headers and
 * counters, work items, codeblock jobs) goes to a static device array of the library
 * (one per GPU, held until the process ends) instead of the context's staging buffers,
 * and the host packs it in workspaces the context reserved at creation.  At most 4096
 * codeblocks per call (LH_ERR_CAPACITY beyond).  Every call on one GPU shares that array:
 * calls on one stream are ordered, but calls that overlap on one GPU (other streams,
 * other contexts of the same device) race. */
int lh_pusch_decode_batch_noalloc(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw,
                                  int8_t* d_rm, uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                                  unsigned nof_ldpc_iterations, void* stream);

/* ---- Training of the learned decoder (no counterpart in the reference: its decoders
 * use a fixed scaling factor).  The model is the layered normalised-offset min-sum
 * decoder unrolled over `iters` iterations in FP32, with per-(iteration, layer) scale
 * alpha and offset beta; the loss is sum_t w[t] * mean BCE of the message bits with
 * logits tau * soft after iteration t. ---- */

/* One codeblock of a training batch (one base graph per batch). */
typedef struct lh_train_desc {
  uint64_t llr_offset;   /* int8 rate-dematched soft bits in d_llrs (nof_llrs of them) */
  uint64_t msg_offset;   /* packed message bits (the targets) in d_msgs */
  uint32_t nof_llrs;
  uint32_t lifting_size;
  uint32_t nof_filler_bits;
  uint32_t nof_layers;   /* layers decoded, 4..46 (BG1) / 4..42 (BG2) */
} lh_train_desc;

/* Workspace bytes lh_nms_train_step needs for this batch. */
size_t lh_nms_workspace_bytes(unsigned base_graph, const lh_train_desc* descs, unsigned n, unsigned iters);

/* Forward (and, with backward != 0, backward) pass over a batch: d_loss[i] = loss of
 * codeblock i; with backward, d_grad_alpha / d_grad_beta ([iters][M] floats) receive the
 * gradient of the summed loss added to their contents.  d_alpha / d_beta: [iters][M]
 * device floats (M = 46 / 42); w: iters host floats; d_workspace of
 * lh_nms_workspace_bytes bytes. */
int lh_nms_train_step(lh_ctx* ctx, unsigned base_graph, const lh_train_desc* descs, unsigned n, unsigned iters,
                      const float* d_alpha, const float* d_beta, const float* w, float tau, const int8_t* d_llrs,
                      const uint8_t* d_msgs, void* d_workspace, size_t workspace_bytes, float* d_loss,
                      float* d_grad_alpha, float* d_grad_beta, int backward, void* stream);

/* One Adam step on d_params (n floats: n_alpha scales, then offsets) with gradients
 * d_grads and moments d_m, d_v; step counts from 1.  Scales are kept in [0, 511/256],
 * offsets in [0, 64] (a non-negative offset keeps the model continuous in its inputs). */
int lh_nms_adam(lh_ctx* ctx, float* d_params, const float* d_grads, float* d_m, float* d_v, unsigned n,
                unsigned n_alpha, float lr, float beta1, float beta2, float eps, unsigned step, void* stream);

/* Device-side CRC of packed bit buffers: crc_calculator::calculate(bit_buffer)
 * (include/ocudu/phy/upper/channel_coding/crc_calculator.h:59-81, behaviour of
 * lib/phy/upper/channel_coding/crc_calculator_generic_impl.cpp:69-91).
 * Buffer i is d_bits + offsets[i] with nbits[i] bits; d_out[i] = remainder. */
int lh_crc_calculate(lh_ctx* ctx, uint32_t poly, const uint64_t* offsets, const uint32_t* nbits, unsigned n,
                     const uint8_t* d_bits, uint32_t* d_out, void* stream);

/* One transport block of lh_tb_finalize: its codeblocks are consecutive entries of the
 * decoder outputs (first_cb .. first_cb + nof_cbs - 1), their packed message bits
 * cb_bits_stride bytes apart. */
typedef struct lh_tb_desc {
  uint64_t tb_offset;       /* output: TB bytes (tbs_bits / 8) */
  uint64_t cb_bits_offset;  /* packed bits of the first codeblock */
  uint32_t cb_bits_stride;  /* bytes between consecutive codeblocks' bits */
  uint32_t tbs_bits;        /* transport block size, multiple of 8 */
  uint32_t first_cb;        /* index of the first codeblock's decoder result */
  uint32_t nof_cbs;
  uint32_t cb_data_bits;    /* data bits per codeblock: K_b*Z - F - CB CRC bits (C > 1) */
  uint32_t pad;
} lh_tb_desc;

#define LH_TB_CRC_OK 1       /* TB CRC passed (C == 1: the codeblock CRC passed) */
#define LH_TB_CB_FAILED 0    /* at least one codeblock CRC failed: TB not assembled */
#define LH_TB_CRC_FAILED (-1)/* all codeblock CRCs passed, TB CRC failed: reset the CB CRC flags */

/* TB reassembly and TB CRC: pusch_decoder_impl::join_and_notify + concatenate_codeblocks
 * (lib/phy/upper/channel_processors/pusch/pusch_decoder_impl.cpp:612-781).  With one
 * codeblock the TB CRC is the codeblock CRC and the first tbs_bits are copied; with
 * several, each codeblock contributes min(remaining, cb_data_bits) bits, the 24-bit TB
 * checksum follows the data of the last codeblock, and CRC24A of the TB must match it.
 * d_tb_status[i] = LH_TB_CRC_OK / LH_TB_CB_FAILED / LH_TB_CRC_FAILED. */
int lh_tb_finalize(lh_ctx* ctx, const lh_tb_desc* tbs, unsigned n, const uint8_t* d_cb_bits,
                   const int32_t* d_cb_results, uint8_t* d_tb, int32_t* d_tb_status, void* stream);

/* ---- Slot graphs: one slot's batch captured once as a HIP graph and replayed. ----
 * The descriptors of the batch are packed and uploaded once; the graph rewinds the work
 * queue, runs the persistent decoder (learned + plain fallback when `learned`), and
 * reassembles the transport blocks (lh_tb_finalize) when n_tbs > 0.  Each launch decodes
 * whatever LLRs d_cw holds at that time into the same buffers, so a serving loop refills
 * d_cw per slot and replays (the hipGraph per slot of the hot-path design; the
 * reference's CUDA backend launches its kernels per call,
 * lib/phy/upper/channel_coding/ldpc/cuda/pusch_codeblock_decoder_cuda_batch.h). */
typedef struct lh_slot_graph lh_slot_graph;
int lh_slot_graph_build(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                        uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                        unsigned nof_ldpc_iterations, int learned, unsigned nof_fallback_iterations,
                        const lh_tb_desc* tbs, unsigned n_tbs, uint8_t* d_tb, int32_t* d_tb_status,
                        lh_slot_graph** graph);
int lh_slot_graph_launch(lh_slot_graph* graph, void* stream);
int lh_slot_graph_destroy(lh_slot_graph* graph);

/* Launch tuning of one (base graph, lifting size), as an autotuner found it: codeblocks
 * packed per workgroup (0 = as many as fit the workgroup) and LDS staging of the codeword
 * LLRs (0 = read them from HBM, the default).  Applies to later decode calls and graph
 * builds. */
int lh_tune_set(lh_ctx* ctx, unsigned base_graph, unsigned lifting_size, unsigned groups, int stage_llrs);

/* Row mode of the decoder for one (base graph, lifting size): 0 one check row per thread,
 * 1 two rows per thread (even lifting size), -1 decided per launch (default: two rows per
 * thread when every class of the launch has a compute unit's resident codeblocks limited
 * by waves rather than LDS and the launch fills every compute unit at one row per thread).
 * Applies to later decode calls and graph builds; results do not depend on it. */
int lh_tune_set_rows(lh_ctx* ctx, unsigned base_graph, unsigned lifting_size, int row_pairs);

/* One synthetic codeblock of lh_gen_codeblocks. */
typedef struct lh_gen_desc {
  uint64_t llr_offset;  /* its rm_length int8 LLRs in d_llrs */
  uint64_t msg_offset;  /* its packed message (K_b*Z bits, MSB first) in d_msgs */
  uint64_t seed_msg;    /* data bits: splitmix64 stream */
  uint64_t seed_noise;  /* AWGN: splitmix64 stream, Box-Muller */
  double n0;            /* noise variance per complex symbol with Es = 1 (0 = noise-free) */
  uint32_t rm_length;   /* E, multiple of modulation_order */
  uint32_t base_graph;
  uint32_t lifting_size;
  uint32_t rv;
  uint32_t modulation_order; /* 2, 4, 6, 8 */
  uint32_t Nref;
  uint32_t nof_filler_bits;
  uint32_t crc_poly;    /* CRC attached to the data bits: LH_CRC24A/B, LH_CRC16, or 0 */
  uint32_t flags;       /* LH_GEN_MSG_INPUT: take the whole message from d_msgs + msg_offset */
  uint32_t pad;
} lh_gen_desc;

#define LH_GEN_MSG_INPUT 1u

/* Device data generator: the transmit chain the reference's chain test builds from
 * ldpc_encoder::encode (include/ocudu/phy/upper/channel_coding/ldpc/ldpc_encoder.h:41),
 * ldpc_rate_matcher::rate_match (ldpc_rate_matcher.h:29), modulation_mapper::modulate
 * (include/ocudu/phy/upper/channel_modulation/modulation_mapper.h:34), an AWGN channel and
 * demodulation_mapper::demodulate_soft (demodulation_mapper.h:48) with the reference's
 * quantiser.  Codeblock i: message = data bits (seed_msg) + CRC + F zero fillers, written
 * packed to d_msgs + msg_offset (when d_msgs is not NULL); LLRs of its rate-matched,
 * mapped, noisy, max-log demapped codeword to d_llrs + llr_offset. */
int lh_gen_codeblocks(lh_ctx* ctx, const lh_gen_desc* descs, unsigned n, int8_t* d_llrs, uint8_t* d_msgs,
                      void* stream);

/* One codeblock of the LDPC encoder + rate matcher (transmit side). */
typedef struct lh_encode_desc {
  uint64_t msg_offset;  /* its message in d_msgs: K_b*Z bits packed MSB first (data, CRC, zero fillers) */
  uint64_t out_offset;  /* its rate-matched bits in d_out: rm_length bits packed MSB first (the last byte's
                           bits past rm_length are 0) */
  uint64_t cw_offset;   /* its codeword in d_codewords: N_short*Z bits, one per byte */
  uint32_t rm_length;   /* E, multiple of modulation_order */
  uint32_t base_graph;
  uint32_t lifting_size;
  uint32_t rv;
  uint32_t modulation_order; /* 1, 2, 4, 6, 8 */
  uint32_t Nref;        /* 0: unlimited buffer */
  uint32_t nof_filler_bits;
  uint32_t pad;
} lh_encode_desc;

/* LDPC encoding and rate matching of a batch of codeblocks, one workgroup each.  Replaces
 * ldpc_encoder::encode (include/ocudu/phy/upper/channel_coding/ldpc/ldpc_encoder.h:41) with
 * ldpc_encoder_buffer::write_codeblock (ldpc_encoder_buffer.h:31) for the codeword (when
 * d_codewords is not NULL), followed by ldpc_rate_matcher::rate_match (ldpc_rate_matcher.h:29)
 * for the rate-matched, bit-interleaved output (TS38.212 5.3.2, 5.4.2). */
int lh_ldpc_encode_rate_match(lh_ctx* ctx, const lh_encode_desc* descs, unsigned n, const uint8_t* d_msgs,
                              uint8_t* d_out, uint8_t* d_codewords, void* stream);

/* Modulation of a demodulated block: the reference's modulation_scheme values
 * (include/ocudu/ran/sch/modulation_scheme.h:21-35), i.e. bits per symbol, 0 for pi/2-BPSK. */
#define LH_MOD_PI_2_BPSK 0
#define LH_MOD_BPSK 1
#define LH_MOD_QPSK 2
#define LH_MOD_QAM16 4
#define LH_MOD_QAM64 6
#define LH_MOD_QAM256 8

/* One block of lh_pusch_demodulate: a span of equalized symbols of one codeword, as one
 * demodulation_mapper::demodulate_soft call of the PUSCH demodulator. */
typedef struct lh_demod_desc {
  uint64_t symbol_offset; /* first symbol in d_symbols (complex float, re then im) and its noise
                             variance in d_noise_vars */
  uint64_t llr_offset;    /* first LLR in d_llrs: nof_symbols * bits per symbol LLRs */
  uint32_t nof_symbols;
  uint32_t modulation;    /* LH_MOD_* */
  uint32_t c_init;        /* scrambling initialisation: RNTI * 2^15 + n_ID (TS38.211 6.3.1.1) */
  uint32_t seq_offset;    /* position of the block's first LLR in the codeword's scrambling
                             sequence (blocks after the first continue the sequence) */
} lh_demod_desc;

/* Soft demodulation and descrambling of a batch of blocks: demodulation_mapper::demodulate_soft
 * (include/ocudu/phy/upper/channel_modulation/demodulation_mapper.h:48) followed by the
 * PUSCH descrambling of pusch_demodulator_impl.cpp:400-414 (pseudo_random_generator::init
 * with c_init and generate, pseudo_random_generator.h:36,88, then revert_scrambling:
 * an LLR is negated where the sequence bit is 1).  LLRs are bit-exact with the reference's
 * scalar demappers for finite symbols.  Blocks up to 2^25 sequence bits
 * (seq_offset + nof_symbols * bits per symbol). */
int lh_pusch_demodulate(lh_ctx* ctx, const lh_demod_desc* descs, unsigned n, const float* d_symbols,
                        const float* d_noise_vars, int8_t* d_llrs, void* stream);

/* One codeword of lh_ulsch_demultiplex: ulsch_demultiplex::configuration
 * (include/ocudu/phy/upper/channel_processors/pusch/ulsch_demultiplex.h:35-66) plus the CSI
 * Part 2 sizes of set_csi_part2 (:75) and the codeword's scrambling initialisation. */
typedef struct lh_ulsch_demux_desc {
  uint64_t llr_offset;       /* the descrambled codeword (lh_pusch_demodulate's output) in d_llrs */
  uint64_t sch_offset;       /* UL-SCH soft bits written to d_sch */
  uint64_t harq_ack_offset;  /* HARQ-ACK soft bits written to d_harq_ack */
  uint64_t csi_part1_offset; /* CSI Part 1 soft bits written to d_csi_part1 */
  uint64_t csi_part2_offset; /* CSI Part 2 soft bits written to d_csi_part2 */
  uint32_t modulation;       /* LH_MOD_* */
  uint32_t nof_layers;       /* 1..4 */
  uint32_t nof_prb;          /* 1..275 */
  uint32_t start_symbol_index;
  uint32_t nof_symbols;      /* start_symbol_index + nof_symbols <= 14 */
  uint32_t dmrs_symbol_mask; /* bit l set: OFDM symbol l of the slot carries DM-RS */
  uint32_t dmrs_type;        /* 1 or 2 */
  uint32_t nof_cdm_groups_without_data;
  uint32_t nof_harq_ack_rvd;       /* G_rvd^ACK, bits */
  uint32_t nof_harq_ack_bits;      /* O_ACK */
  uint32_t nof_enc_harq_ack_bits;  /* G^ACK */
  uint32_t nof_csi_part1_bits;     /* O_CSI-1 */
  uint32_t nof_enc_csi_part1_bits; /* G^CSI-1 */
  uint32_t nof_csi_part2_bits;     /* O_CSI-2 (0: no CSI Part 2) */
  uint32_t nof_enc_csi_part2_bits; /* G^CSI-2 */
  uint32_t c_init;                 /* the codeword's scrambling initialisation (lh_demod_desc.c_init) */
} lh_ulsch_demux_desc;

typedef struct lh_ulsch_demux_sizes {
  uint32_t nof_llrs;      /* codeword soft bits read */
  uint32_t nof_sch;       /* UL-SCH soft bits: the rate dematcher's codeword */
  uint32_t nof_harq_ack;
  uint32_t nof_csi_part1;
  uint32_t nof_csi_part2;
} lh_ulsch_demux_sizes;

/* ulsch_configuration (include/ocudu/ran/pusch/ulsch_info.h:21-60). */
typedef struct lh_ulsch_config {
  uint32_t tbs_bits;           /* 0: no UL-SCH (UCI only) */
  float target_code_rate;      /* normalised R; up to 0.25 (0 included) selects BG2 for the
                                  codeblock sizes, as get_ldpc_base_graph does */
  uint32_t modulation;         /* LH_MOD_* */
  uint32_t nof_harq_ack_bits;
  uint32_t nof_csi_part1_bits;
  uint32_t nof_csi_part2_bits;
  float alpha_scaling;
  float beta_offset_harq_ack;
  float beta_offset_csi_part1;
  float beta_offset_csi_part2;
  uint32_t nof_rb;
  uint32_t start_symbol_index;
  uint32_t nof_symbols;
  uint32_t dmrs_type;          /* 1 or 2 */
  uint32_t dmrs_symbol_mask;   /* bit l: OFDM symbol l carries DM-RS */
  uint32_t nof_cdm_groups_without_data;
  uint32_t nof_layers;
  int32_t contains_dc;
} lh_ulsch_config;

/* ulsch_information with its sch_information (include/ocudu/ran/sch/sch_segmentation.h:17-31). */
typedef struct lh_ulsch_info {
  uint32_t sch_base_graph;     /* 0 without a TB */
  uint32_t sch_tb_crc_size;
  uint32_t sch_nof_cb;
  uint32_t sch_lifting_size;
  uint32_t sch_nof_bits_per_cb;
  uint32_t sch_nof_filler_bits_per_cb;
  uint32_t nof_ul_sch_bits;
  uint32_t nof_harq_ack_bits;  /* G^ACK */
  uint32_t nof_harq_ack_rvd;   /* G_rvd^ACK */
  uint32_t nof_csi_part1_bits; /* G^CSI-1 */
  uint32_t nof_csi_part2_bits; /* G^CSI-2 */
  uint32_t nof_harq_ack_re;
  uint32_t nof_csi_part1_re;
  uint32_t nof_csi_part2_re;
  uint32_t nof_dc_overlap_bits;
  uint32_t pad;
} lh_ulsch_info;

/* get_ulsch_information (lib/ran/pusch/ulsch_info.cpp:148-303, with get_sch_segmentation_info,
 * lib/ran/sch/sch_segmentation.cpp:11-42): the UL-SCH / HARQ-ACK / CSI bit accounting of a
 * grant, in the reference's float arithmetic (host only, no context).  Rejects what the
 * reference asserts on (CDM groups, DM-RS symbols outside the allocation). */
int lh_ulsch_information(const lh_ulsch_config* cfg, lh_ulsch_info* info);

/* Sizes of one codeword's demultiplexed fields (host only, no context).  Rejects what the
 * reference rejects with an assertion: a field whose encoded bits are not a multiple of
 * the bits per RE, that does not fit its REs, or whose payload and encoded sizes are not
 * both zero or both non-zero; an invalid DM-RS configuration. */
int lh_ulsch_demux_get_sizes(const lh_ulsch_demux_desc* desc, lh_ulsch_demux_sizes* sizes);

/* UL-SCH demultiplexing of TS38.212 6.2.7 on the device: ulsch_demultiplex_impl
 * (lib/phy/upper/channel_processors/pusch/ulsch_demultiplex_impl.cpp:178-573) over whole
 * codewords, CSI Part 2 known before the first OFDM symbol (set_csi_part2 right after
 * demultiplex; the reference's resident index builder
 * pusch_demodulator_gpu_impl.cpp:471-625 allocates the same).  Per OFDM symbol: reserved
 * REs for a HARQ-ACK of up to 2 bits, HARQ-ACK of more than 2 bits, CSI Part 1, CSI Part 2;
 * the UL-SCH soft bits are the rest, in RE order, with the REs of a 1- or 2-bit HARQ-ACK
 * zeroed; UCI of 1 or 2 bits has its placeholder bits re-scrambled as
 * on_uci_placeholder_1bit/_2bit (:86-176) do.  The outputs' sizes are those of
 * lh_ulsch_demux_get_sizes; an output pointer may be NULL when no codeword has that field. */
int lh_ulsch_demultiplex(lh_ctx* ctx, const lh_ulsch_demux_desc* descs, unsigned n, const int8_t* d_llrs,
                         int8_t* d_sch, int8_t* d_harq_ack, int8_t* d_csi_part1, int8_t* d_csi_part2,
                         void* stream);

/* One transport block of lh_gen_transport_blocks. */
typedef struct lh_gen_tb_desc {
  uint64_t tb_offset;      /* TB bytes (tbs_bits / 8) written to d_tb */
  uint64_t msg_offset;     /* message of its first codeblock in d_msgs (packed, MSB first) */
  uint64_t seed;           /* TB data: splitmix64 stream */
  uint32_t msg_stride;     /* bytes between consecutive codeblock messages */
  uint32_t tbs_bits;       /* multiple of 8 */
  uint32_t nof_cbs;        /* segmentation (lh_segment_rx) */
  uint32_t segment_length; /* K_b * Z */
  uint32_t nof_filler_bits;
  uint32_t pad;
} lh_gen_tb_desc;

/* Device transport-block generator: random TB bytes, the TB CRC (CRC16 up to 3824 bits,
 * else CRC24A) and the codeblock messages of the TX segmentation
 * (ldpc_segmenter_tx::new_transmission + ldpc_segmenter_buffer::read_codeblock,
 * include/ocudu/phy/upper/channel_coding/ldpc/ldpc_segmenter_tx.h:31,
 * ldpc_segmenter_buffer.h), ready for lh_gen_codeblocks with LH_GEN_MSG_INPUT. */
int lh_gen_transport_blocks(lh_ctx* ctx, const lh_gen_tb_desc* descs, unsigned n, uint8_t* d_tb, uint8_t* d_msgs,
                            void* stream);

/* Codeblock segmentation of a received transport block (host only, no context):
 * ldpc_segmenter_rx::segment (lib/phy/upper/channel_coding/ldpc/ldpc_segmenter_rx_impl.cpp:33-85)
 * with compute_nof_codeblocks / compute_lifting_size (include/ocudu/ran/sch/sch_segmentation.h:52-110)
 * and compute_rm_length (ldpc_segmenter_helpers.h:63-76). */
typedef struct lh_segmentation {
  uint32_t nof_cbs;
  uint32_t lifting_size;
  uint32_t segment_length;  /* K_b * Z */
  uint32_t nof_filler_bits;
  uint32_t nof_tb_crc_bits; /* 16 or 24 */
  uint32_t nof_cb_crc_bits; /* the codeblock CRC: 24 when C > 1, else the TB CRC's */
  uint32_t cb_data_bits;    /* data bits per codeblock (lh_tb_desc.cb_data_bits) */
  uint32_t zero_pad;
} lh_segmentation;

/* rm_lengths / cw_offsets receive nof_cbs entries (at most max_cbs). */
int lh_segment_rx(unsigned base_graph, unsigned tbs_bits, unsigned modulation_order, unsigned nof_layers,
                  unsigned nof_ch_symbols, lh_segmentation* seg, uint32_t* rm_lengths, uint32_t* cw_offsets,
                  unsigned max_cbs);

/* ---- Slot-deadline serving runtime (the resident batch path of the reference:
 * pusch_resident_codeblock_decoder::decode_resident_softbits,
 * include/ocudu/phy/upper/channel_processors/pusch/pusch_resident_codeblock_decoder.h:125-149,
 * HARQ soft buffers kept on the device as rx_buffer_pool does on the host). ----
 * Transport-block requests are admitted when their HARQ buffers fit the HBM budget (a
 * retransmission combines into its process's buffers), batched continuously into slot
 * launches (decode + TB reassembly) on up to max_slots_in_flight streams whose work
 * overlaps on the device, and completed asynchronously.  A TB whose CRC passes releases its
 * HARQ buffers; a failed one keeps them for a retransmission (lh_server_harq_release drops
 * them).  As the reference's rx_buffer, a HARQ process keeps each codeblock's message bits
 * and CRC state: a codeblock whose CRC passed in an earlier transmission only combines the
 * new LLRs and is not decoded again (pusch_decoder_impl.cpp:412-431), and all are decoded
 * again after a TB CRC failure with every codeblock passed (:657-669).  As rx_buffer_pool
 * (lib/phy/upper/rx_buffer_pool_impl.cpp:79-125), a process expires expire_timeout_slots
 * after the slot of its last reservation (lh_server_run_slot), unless it is in use, which
 * renews its expiry.  Single-threaded: one caller drives submit / step / poll / run_slot. */
typedef struct lh_server lh_server;

typedef struct lh_server_config {
  size_t harq_budget_bytes;        /* device memory for HARQ processes (per codeblock N_short*Z soft bits + its
                                      message bits, K_b*Z/8 bytes) */
  unsigned max_cbs_per_slot;       /* codeblocks per slot launch (<= the context capacity) */
  unsigned max_slots_in_flight;    /* 1..8 */
  unsigned nof_ldpc_iterations;
  unsigned nof_fallback_iterations;/* learned decoder: plain fallback iterations (0 = none) */
  int learned;                     /* decode with the learned tables */
  unsigned expire_timeout_slots;   /* HARQ process lifetime in slots after its last reservation (0 = no
                                      expiry): rx_buffer_pool_config::expire_timeout_slots */
} lh_server_config;

typedef struct lh_tb_request {
  uint64_t id;              /* caller's tag */
  const int8_t* d_llrs;     /* codeword LLRs on the device: nof_ch_symbols * modulation_order */
  uint32_t harq_process;    /* HARQ buffer key */
  int32_t new_data;         /* 1: new transmission; 0: retransmission combined into the buffers */
  uint32_t base_graph;
  uint32_t tbs_bits;
  uint32_t modulation_order;
  uint32_t nof_layers;
  uint32_t nof_ch_symbols;
  uint32_t rv;
  uint32_t Nref;
  uint32_t slot;            /* slot of the transmission (a monotonic counter, as lh_server_run_slot's): its HARQ
                               process expires expire_timeout_slots after it, or after the slot it is admitted in
                               when it waited for HARQ memory past its own slot */
} lh_tb_request;

#define LH_TB_REJECTED (-2) /* a retransmission without matching HARQ buffers */

typedef struct lh_tb_result {
  uint64_t id;
  uint32_t harq_process;
  int32_t tb_crc_ok;        /* LH_TB_CRC_OK / LH_TB_CB_FAILED / LH_TB_CRC_FAILED / LH_TB_REJECTED */
  uint32_t nof_cbs;
  uint32_t nof_cbs_ok;
  uint32_t max_iterations;  /* over the codeblocks that passed */
  uint32_t pad;
  double latency_us;        /* submit -> completion seen by poll */
} lh_tb_result;

typedef struct lh_server_stats_t {
  uint64_t harq_free_bytes;
  uint32_t harq_processes;
  uint32_t pending;         /* waiting for HARQ memory or for their process */
  uint32_t ready;           /* admitted, not launched */
  uint32_t slots_in_flight;
} lh_server_stats_t;

int lh_server_create(lh_ctx* ctx, const lh_server_config* cfg, lh_server** server);
int lh_server_destroy(lh_server* server);
/* Queues a request (its LLRs must stay valid until its result is polled); tb_out (host,
 * tbs_bits / 8 bytes, may be NULL) receives the TB when its CRC passes. */
int lh_server_submit(lh_server* server, const lh_tb_request* req, uint8_t* tb_out);
/* n requests at once (a slot's arrivals); tb_outs may be NULL. */
int lh_server_submit_batch(lh_server* server, const lh_tb_request* reqs, unsigned n, uint8_t* const* tb_outs);
/* Launches one slot of the admitted requests when a slot is free (non-blocking). */
int lh_server_step(lh_server* server);
/* Retires completed slots and returns up to max results; wait != 0 blocks until one
 * in-flight slot completes when no result is ready. */
int lh_server_poll(lh_server* server, lh_tb_result* out, unsigned max, unsigned* n_out, int wait);
int lh_server_harq_release(lh_server* server, uint32_t harq_process);
int lh_server_stats(lh_server* server, lh_server_stats_t* stats);
/* Housekeeping of the HARQ processes at `slot` (rx_buffer_pool::run_slot,
 * include/ocudu/phy/upper/rx_buffer_pool.h:68, lib/phy/upper/rx_buffer_pool_impl.cpp:88-125):
 * a process in use (admitted, not yet completed) has its expiry renewed to slot +
 * expire_timeout_slots; any other whose expiry is at or before `slot` is released, so a
 * later retransmission for it is rejected.  Slots compare modulo 2^32. */
int lh_server_run_slot(lh_server* server, uint32_t slot);

/* ---- The slot loop, natively: the serving side of the PUSCH processor over a trace of
 * transport blocks (each first transmitted in its slot, retransmitted with its second
 * request harq_rtt_slots later when the first fails), driven by one host thread without
 * an interpreter between the slot boundaries and the completions. ---- */
#define LH_TB_NOT_SENT (-3) /* result of a retransmission that was not needed or fell past the run */

typedef struct lh_serve_tb {
  lh_tb_request tx[2]; /* new transmission, then its retransmission (same harq_process); ids are set by the loop
                          to 2 i + tx and slots to the slot each is sent in */
  uint8_t* tb_out[2];  /* host TB bytes of each transmission (may be NULL) */
  uint32_t slot;       /* slot of the new transmission */
  uint32_t pad;
} lh_serve_tb;

typedef struct lh_serve_config {
  double slot_us;          /* slot duration */
  double timeout_s;        /* the loop fails with LH_ERR_CAPACITY past this wall time */
  unsigned nof_slots;      /* slots with new transmissions: 0 .. nof_slots - 1 */
  unsigned harq_rtt_slots; /* slots from a failed transmission to its retransmission */
  int realtime;            /* 1: slot k is released k * slot_us after the start; 0: as soon as the slots whose
                              retransmissions fall in it have completed (deterministic) */
  unsigned pad;
} lh_serve_config;

/* Runs the trace through the server: at each slot boundary lh_server_run_slot(k), then the
 * slot's new transmissions and the retransmissions due in it are submitted and stepped;
 * completions are polled continuously.  A retransmission is due max(slot + rtt, current
 * slot) after its failed transmission; one that would fall past the run, and a failed
 * retransmission, release their HARQ process.  results: 2 n_tbs entries, index 2 i + tx
 * (LH_TB_NOT_SENT where nothing was sent); slot_latency_us: nof_slots + harq_rtt_slots + 2
 * entries, submission of a slot to the completion of its last transmission (-1: nothing
 * sent in it); wall_s: the loop's wall time. */
int lh_server_serve(lh_server* server, const lh_serve_tb* tbs, unsigned n_tbs, const lh_serve_config* cfg,
                    lh_tb_result* results, double* slot_latency_us, double* wall_s);

#ifdef __cplusplus
}
#endif

#endif
