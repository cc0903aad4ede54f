// C entry points over the reference's own CPU LDPC implementation -- TEST
// INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile.ref together with the reference
// sources where they lie under the reference tree; the output goes to oracle/_ref/
// (git-ignored).  Used by tests/ to pin the oracle and by bench.py's cpu_baseline leg
// to time the reference AVX-512/AVX2 decoders on the host cores.
#include "crc_calculator_generic_impl.h"
#include "ldpc_decoder_avx2.h"
#include "ldpc_decoder_avx512.h"
#include "ldpc_decoder_generic.h"
#include "ldpc_encoder_generic.h"
#include "ldpc_graph_impl.h"
#include "ldpc_luts_impl.h"
#include "ldpc_rate_dematcher_impl.h"
#include "ldpc_rate_matcher_impl.h"
#include "ldpc_segmenter_rx_impl.h"
#include "ldpc_segmenter_tx_impl.h"
#include "ocudu/adt/bit_buffer.h"
#include "ocudu/phy/upper/log_likelihood_ratio.h"
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

using namespace ocudu;

namespace {

ldpc_base_graph_type to_bg(int bg)
{
  return bg == 1 ? ldpc_base_graph_type::BG1 : ldpc_base_graph_type::BG2;
}

modulation_scheme to_mod(unsigned qm)
{
  switch (qm) {
    case 1:
      return modulation_scheme::BPSK;
    case 2:
      return modulation_scheme::QPSK;
    case 4:
      return modulation_scheme::QAM16;
    case 6:
      return modulation_scheme::QAM64;
    default:
      return modulation_scheme::QAM256;
  }
}

crc_generator_poly to_poly(uint32_t p)
{
  return static_cast<crc_generator_poly>(p);
}

std::unique_ptr<ldpc_decoder> make_decoder(int kind, bool force, bool es)
{
  if (kind == 2 && __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw")) {
    return std::make_unique<ldpc_decoder_avx512>(force, es);
  }
  if (kind >= 1 && __builtin_cpu_supports("avx2")) {
    return std::make_unique<ldpc_decoder_avx2>(force, es);
  }
  return std::make_unique<ldpc_decoder_generic>(force, es);
}

} // namespace

extern "C" {

// 0 = generic, 1 = avx2, 2 = avx512: the implementation make_decoder() really builds.
int ref_decoder_kind(int kind)
{
  if (kind == 2 && __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw")) {
    return 2;
  }
  if (kind >= 1 && __builtin_cpu_supports("avx2")) {
    return 1;
  }
  return 0;
}

void* ref_decoder_create(int kind, int force_decoding, int early_stop_syndrome)
{
  return make_decoder(kind, force_decoding != 0, early_stop_syndrome != 0).release();
}

void ref_decoder_destroy(void* d)
{
  delete static_cast<ldpc_decoder*>(d);
}

// Returns the decoder's optional<unsigned>: the value, or 0 for nullopt.
unsigned ref_decoder_decode(void*          dec,
                            const int8_t*  llrs,
                            unsigned       nof_llrs,
                            int            bg,
                            unsigned       Z,
                            unsigned       nof_filler_bits,
                            unsigned       nof_crc_bits,
                            unsigned       max_iterations,
                            uint32_t       crc_poly,
                            uint8_t*       out_packed)
{
  auto*                       d = static_cast<ldpc_decoder*>(dec);
  ldpc_decoder::configuration cfg;
  cfg.base_graph      = to_bg(bg);
  cfg.lifting_size    = static_cast<ldpc::lifting_size_t>(Z);
  cfg.nof_filler_bits = nof_filler_bits;
  cfg.nof_crc_bits    = nof_crc_bits;
  cfg.max_iterations  = max_iterations;
  unsigned                    K = (bg == 1 ? 22 : 10) * Z;
  dynamic_bit_buffer          out(K);
  std::memcpy(out.get_buffer().data(), out_packed, (K + 7) / 8);
  span<const log_likelihood_ratio> in(reinterpret_cast<const log_likelihood_ratio*>(llrs), nof_llrs);
  std::unique_ptr<crc_calculator>  crc;
  if (crc_poly != 0) {
    crc = std::make_unique<crc_calculator_generic_impl>(to_poly(crc_poly));
  }
  std::optional<unsigned> r = d->decode(out, in, crc.get(), cfg);
  std::memcpy(out_packed, out.get_buffer().data(), (K + 7) / 8);
  return r.has_value() ? r.value() : 0;
}

int ref_rate_dematch(int8_t*       out,
                     unsigned      block_length,
                     const int8_t* in,
                     unsigned      E,
                     int           new_data,
                     int           bg,
                     unsigned      Z,
                     unsigned      rv,
                     unsigned      Qm,
                     unsigned      Nref,
                     unsigned      nof_filler_bits)
{
  static thread_local ldpc_rate_dematcher_impl dm;
  codeblock_metadata                         md{};
  md.tb_common.base_graph       = to_bg(bg);
  md.tb_common.lifting_size     = static_cast<ldpc::lifting_size_t>(Z);
  md.tb_common.rv               = rv;
  md.tb_common.mod              = to_mod(Qm);
  md.tb_common.Nref             = Nref;
  md.cb_specific.nof_filler_bits = nof_filler_bits;
  md.cb_specific.rm_length      = E;
  md.cb_specific.full_length    = block_length;
  dm.rate_dematch(span<log_likelihood_ratio>(reinterpret_cast<log_likelihood_ratio*>(out), block_length),
                  span<const log_likelihood_ratio>(reinterpret_cast<const log_likelihood_ratio*>(in), E),
                  new_data != 0,
                  md);
  return 0;
}

// msg: K_b*Z bits one per byte (fillers = 0).  Writes N_short*Z transmitted bits (one
// per byte) and then rate-matches them to E bits (one per byte) into rm_out if non-null.
int ref_encode_rate_match(const uint8_t* msg,
                          int            bg,
                          unsigned       Z,
                          uint8_t*       cw_out,
                          uint8_t*       rm_out,
                          unsigned       E,
                          unsigned       rv,
                          unsigned       Qm,
                          unsigned       Nref,
                          unsigned       nof_filler_bits)
{
  static thread_local ldpc_encoder_generic     enc;
  static thread_local ldpc_rate_matcher_impl   rm;
  unsigned                                     K = (bg == 1 ? 22 : 10) * Z;
  dynamic_bit_buffer                           in(K);
  for (unsigned i = 0; i != K; ++i) {
    in.insert(msg[i] & 1, i, 1);
  }
  ldpc_encoder::configuration ecfg;
  ecfg.base_graph   = to_bg(bg);
  ecfg.lifting_size = static_cast<ldpc::lifting_size_t>(Z);
  ecfg.Nref         = Nref;
  const ldpc_encoder_buffer& eb = enc.encode(in, ecfg);
  unsigned                   N  = eb.get_codeblock_length();
  if (cw_out != nullptr) {
    eb.write_codeblock(span<uint8_t>(cw_out, N), 0);
  }
  if (rm_out != nullptr) {
    codeblock_metadata md{};
    md.tb_common.base_graph        = to_bg(bg);
    md.tb_common.lifting_size      = static_cast<ldpc::lifting_size_t>(Z);
    md.tb_common.rv                = rv;
    md.tb_common.mod               = to_mod(Qm);
    md.tb_common.Nref              = Nref;
    md.cb_specific.nof_filler_bits = nof_filler_bits;
    md.cb_specific.rm_length       = E;
    md.cb_specific.full_length     = N;
    dynamic_bit_buffer out(E);
    rm.rate_match(out, eb, md);
    for (unsigned i = 0; i != E; ++i) {
      rm_out[i] = out.extract(i, 1);
    }
  }
  return static_cast<int>(N);
}

uint32_t ref_crc(uint32_t poly, const uint8_t* packed, unsigned nbits)
{
  crc_calculator_generic_impl crc(to_poly(poly));
  dynamic_bit_buffer          b(nbits);
  if (nbits > 0) {
    std::memcpy(b.get_buffer().data(), packed, (nbits + 7) / 8);
  }
  return crc.calculate(b);
}

// Shift of the lifted edge (row, col) or -1 without an edge.
int ref_graph_shift(int bg, unsigned Z, unsigned row, unsigned col)
{
  auto g = ldpc::get_graph(to_bg(bg), static_cast<ldpc::lifting_size_t>(Z));
  uint16_t v = g[row][col];
  return v == ldpc::NO_EDGE ? -1 : static_cast<int>(v);
}

int8_t ref_llr_add(int8_t a, int8_t b)
{
  return (log_likelihood_ratio(a) + log_likelihood_ratio(b)).to_value_type();
}

int8_t ref_llr_promotion_sum(int8_t a, int8_t b)
{
  return log_likelihood_ratio::promotion_sum(a, b).to_value_type();
}

int8_t ref_llr_quantize(float v, float r)
{
  return log_likelihood_ratio::quantize(v, r).to_value_type();
}

// Throughput of the reference decoder: nthreads worker threads, one decoder instance
// each (the codeblock decoder pool of the PUSCH decoder), decode CBs round-robin from
// the given set until at least min_seconds elapsed.  Returns seconds; *cbs_done and
// *crc_ok receive the counts.
double ref_decode_throughput(int            kind,
                             int            nthreads,
                             const int8_t*  llrs,
                             unsigned       n_cb,
                             unsigned       llr_stride,
                             int            bg,
                             unsigned       Z,
                             unsigned       nof_filler_bits,
                             unsigned       max_iterations,
                             uint32_t       crc_poly,
                             double         min_seconds,
                             unsigned long* cbs_done,
                             unsigned long* crc_ok,
                             unsigned long* iters_sum)
{
  std::atomic<unsigned long> done{0}, ok{0}, its{0};
  auto                       t0 = std::chrono::steady_clock::now();
  std::vector<std::thread>   pool;
  for (int t = 0; t < nthreads; ++t) {
    pool.emplace_back([&, t]() {
      auto                        dec = make_decoder(kind, false, false);
      crc_calculator_generic_impl crc(to_poly(crc_poly));
      ldpc_decoder::configuration cfg;
      cfg.base_graph      = to_bg(bg);
      cfg.lifting_size    = static_cast<ldpc::lifting_size_t>(Z);
      cfg.nof_filler_bits = nof_filler_bits;
      cfg.nof_crc_bits    = 24;
      cfg.max_iterations  = max_iterations;
      dynamic_bit_buffer out((bg == 1 ? 22 : 10) * Z);
      unsigned long      my_done = 0, my_ok = 0, my_its = 0;
      unsigned           i       = static_cast<unsigned>(t) % n_cb;
      while (true) {
        span<const log_likelihood_ratio> in(
            reinterpret_cast<const log_likelihood_ratio*>(llrs + static_cast<size_t>(i) * llr_stride), llr_stride);
        auto r = dec->decode(out, in, &crc, cfg);
        ++my_done;
        if (r.has_value()) {
          ++my_ok;
          my_its += r.value();
        } else {
          my_its += max_iterations;
        }
        i = (i + static_cast<unsigned>(nthreads)) % n_cb;
        if ((my_done % 8) == 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= min_seconds) {
          break;
        }
      }
      done += my_done;
      ok += my_ok;
      its += my_its;
    });
  }
  for (auto& th : pool) {
    th.join();
  }
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *cbs_done   = done.load();
  *crc_ok     = ok.load();
  *iters_sum  = its.load();
  return secs;
}

} // extern "C"

namespace {

// Codeblock metadata as 8 words: lifting size, segment length, full length, rm length,
// filler bits, cw offset, CRC bits, info bits.
void put_meta(unsigned* m, const codeblock_metadata& md, unsigned seg_len, unsigned info_bits)
{
  m[0] = static_cast<unsigned>(md.tb_common.lifting_size);
  m[1] = seg_len;
  m[2] = md.cb_specific.full_length;
  m[3] = md.cb_specific.rm_length;
  m[4] = md.cb_specific.nof_filler_bits;
  m[5] = md.cb_specific.cw_offset;
  m[6] = md.cb_specific.nof_crc_bits;
  m[7] = info_bits;
}

segmenter_config seg_config(int bg, unsigned tbs_bytes, unsigned qm, unsigned nof_layers, unsigned nof_ch_symbols,
                            unsigned rv)
{
  segmenter_config cfg;
  cfg.transport_block_size = units::bytes(tbs_bytes);
  cfg.base_graph           = to_bg(bg);
  cfg.rv                   = rv;
  cfg.mod                  = to_mod(qm);
  cfg.Nref                 = 0;
  cfg.nof_layers           = nof_layers;
  cfg.nof_ch_symbols       = nof_ch_symbols;
  return cfg;
}

} // namespace

extern "C" {

// ldpc_segmenter_tx_impl::new_transmission + read_codeblock on a transport block: returns
// the number of codeblocks (or -1 when more than max_cbs), writes 8 metadata words per
// codeblock and, when cb_bits is not null, the codeblock messages (segment length bits
// each, one bit per byte, cb_stride bytes apart).
int ref_segment_tx(int            bg,
                   const uint8_t* tb,
                   unsigned       tbs_bytes,
                   unsigned       qm,
                   unsigned       nof_layers,
                   unsigned       nof_ch_symbols,
                   unsigned       rv,
                   unsigned       max_cbs,
                   unsigned*      meta,
                   uint8_t*       cb_bits,
                   unsigned       cb_stride)
{
  ldpc_segmenter_tx_impl::sch_crc crcs{std::make_unique<crc_calculator_generic_impl>(crc_generator_poly::CRC16),
                                       std::make_unique<crc_calculator_generic_impl>(crc_generator_poly::CRC24A),
                                       std::make_unique<crc_calculator_generic_impl>(crc_generator_poly::CRC24B)};
  ldpc_segmenter_tx_impl       seg(crcs);
  const ldpc_segmenter_buffer& buf = seg.new_transmission(seg_config(bg, tbs_bytes, qm, nof_layers, nof_ch_symbols, rv));
  unsigned                     C   = buf.get_nof_codeblocks();
  if (C > max_cbs) {
    return -1;
  }
  unsigned           seg_len = buf.get_segment_length().value();
  dynamic_bit_buffer cb(seg_len);
  for (unsigned c = 0; c != C; ++c) {
    put_meta(meta + 8 * c, buf.get_cb_metadata(c), seg_len, buf.get_cb_info_bits(c).value());
    if (cb_bits != nullptr) {
      buf.read_codeblock(cb, span<const uint8_t>(tb, tbs_bytes), c);
      for (unsigned i = 0; i != seg_len; ++i) {
        cb_bits[static_cast<size_t>(c) * cb_stride + i] = static_cast<uint8_t>(cb.extract(i, 1));
      }
    }
  }
  return static_cast<int>(C);
}

// ldpc_segmenter_rx_impl::segment on a codeword of nof_ch_symbols * qm LLRs: returns the
// number of codeblocks (or -1 when more than max_cbs) and 8 metadata words per codeblock
// (the info-bits word is 0).
int ref_segment_rx(int      bg,
                   unsigned tbs_bytes,
                   unsigned qm,
                   unsigned nof_layers,
                   unsigned nof_ch_symbols,
                   unsigned rv,
                   unsigned max_cbs,
                   unsigned* meta)
{
  std::vector<log_likelihood_ratio> cw(static_cast<size_t>(nof_ch_symbols) * qm);
  static_vector<described_rx_codeblock, MAX_NOF_SEGMENTS> cbs;
  ldpc_segmenter_rx_impl seg;
  seg.segment(cbs, cw, seg_config(bg, tbs_bytes, qm, nof_layers, nof_ch_symbols, rv));
  if (cbs.size() > max_cbs) {
    return -1;
  }
  unsigned lifting = 0;
  for (unsigned c = 0; c != cbs.size(); ++c) {
    const codeblock_metadata& md = cbs[c].second;
    lifting                      = static_cast<unsigned>(md.tb_common.lifting_size);
    put_meta(meta + 8 * c, md, (bg == 1 ? 22 : 10) * lifting, 0);
  }
  return static_cast<int>(cbs.size());
}

} // extern "C"
