// extern "C" access to the reference's own soft demappers, modulation mapper,
// pseudo-random generator and UL-SCH demultiplexer, built from their sources
// (oracle/Makefile.ref) -- TEST INFRASTRUCTURE ONLY: the checker of the oracle's
// modulation, demodulation, scrambling and demultiplexing restatements.  Built without AVX/NEON flags, so
// the reference's scalar (generic) demapper path runs.
#include "demodulation_mapper_impl.h"
#include "modulation_mapper_lut_impl.h"
#include "pseudo_random_generator_impl.h"
#include "ulsch_demultiplex_impl.h"
#include "ocudu/phy/upper/channel_processors/pusch/pusch_decoder_buffer.h"
#include "ocudu/adt/bit_buffer.h"

#include <algorithm>
#include <vector>

using namespace ocudu;

extern "C" {

// demodulation_mapper::demodulate_soft over nsym symbols (re, im interleaved).
void ref_demodulate_soft(int8_t* llrs, const float* iq, const float* nv, unsigned nsym, int modulation)
{
  static demodulation_mapper_impl demapper;
  const unsigned qm = modulation == 0 ? 1u : (unsigned)modulation;
  span<log_likelihood_ratio> out(reinterpret_cast<log_likelihood_ratio*>(llrs), (size_t)nsym * qm);
  span<const cf_t> sym(reinterpret_cast<const cf_t*>(iq), nsym);
  span<const float> n(nv, nsym);
  demapper.demodulate_soft(out, sym, n, static_cast<modulation_scheme>(modulation));
}

// modulation_mapper::modulate (the reference's LUT mapper) of nbits bits, one per byte,
// into nbits / Qm symbols (re, im interleaved).
void ref_modulate(float* iq, const uint8_t* bits, unsigned nbits, int modulation)
{
  static modulation_mapper_lut_impl mapper;
  const unsigned qm = modulation == 0 ? 1u : (unsigned)modulation;
  dynamic_bit_buffer in(nbits);
  for (unsigned i = 0; i < nbits; ++i) {
    in.insert(bits[i] & 1u, i, 1);
  }
  span<cf_t> sym(reinterpret_cast<cf_t*>(iq), nbits / qm);
  mapper.modulate(sym, in, static_cast<modulation_scheme>(modulation));
}

// pseudo_random_generator: init(c_init), advance(offset), generate(n bits); one bit per byte.
void ref_scrambling_sequence(uint8_t* c, uint32_t c_init, unsigned offset, unsigned n)
{
  pseudo_random_generator_impl prg;
  prg.init(c_init);
  prg.advance(offset);
  dynamic_bit_buffer seq(n);
  prg.generate(seq);
  for (unsigned i = 0; i < n; ++i) {
    c[i] = (uint8_t)seq.extract(i, 1);
  }
}

} // extern "C"

// ---- UL-SCH demultiplexer (ulsch_demultiplex_impl.cpp built from source) -------------

namespace {

// Collects the soft bits handed to one decoder buffer.
class collecting_buffer : public pusch_decoder_buffer
{
public:
  std::vector<int8_t> bits;
  bool ended = false;
  std::vector<log_likelihood_ratio> view;

  span<log_likelihood_ratio> get_next_block_view(unsigned block_size) override
  {
    view.resize(block_size);
    return view;
  }
  void on_new_softbits(span<const log_likelihood_ratio> softbits) override
  {
    for (const log_likelihood_ratio& v : softbits) {
      bits.push_back(static_cast<int8_t>(v));
    }
  }
  void on_end_softbits() override { ended = true; }
};

} // namespace

extern "C" {

// ulsch_demultiplex::demultiplex, then set_csi_part2 before any soft bits (when CSI Part 2
// is present), then the codeword in blocks of `block` soft bits (the last one shorter) with
// the sequence of pseudo_random_generator(c_init).  cfg: modulation, layers, PRBs, start
// symbol, symbols, DM-RS symbol mask, DM-RS type, CDM groups without data, G_rvd, O_ACK,
// G_ACK, O_CSI1, G_CSI1, O_CSI2, G_CSI2.  Writes the four outputs and their sizes
// (sizes[0..3]); returns 0, -1 when a buffer did not receive its end, -2 for a first
// symbol without data REs.
int ref_ulsch_demultiplex(const unsigned* cfg, const int8_t* llrs, unsigned nof_llrs, uint32_t c_init,
                          unsigned block, int8_t* sch, int8_t* ack, int8_t* csi1, int8_t* csi2, unsigned* sizes)
{
  ulsch_demultiplex::configuration c;
  c.modulation = static_cast<modulation_scheme>(cfg[0]);
  c.nof_layers = cfg[1];
  c.nof_prb = cfg[2];
  c.start_symbol_index = cfg[3];
  c.nof_symbols = cfg[4];
  c.dmrs_symbol_mask = symbol_slot_mask(14);
  for (unsigned l = 0; l < 14; ++l) {
    if ((cfg[5] >> l) & 1u) {
      c.dmrs_symbol_mask.set(l);
    }
  }
  c.dmrs = cfg[6] == 1 ? dmrs_config_type::type1 : dmrs_config_type::type2;
  c.nof_cdm_groups_without_data = cfg[7];
  c.nof_harq_ack_rvd = cfg[8];
  c.nof_harq_ack_bits = cfg[9];
  c.nof_enc_harq_ack_bits = cfg[10];
  c.nof_csi_part1_bits = cfg[11];
  c.nof_enc_csi_part1_bits = cfg[12];

  // A first symbol without data REs (DM-RS with every CDM group empty of data) leaves the
  // reference's codeword buffer at a zero block size forever (on_new_block,
  // ulsch_demultiplex_impl.cpp:262-300): not fed to it.
  const bool first_dmrs = (cfg[5] >> cfg[3]) & 1u;
  if (first_dmrs && cfg[7] * (cfg[6] == 1 ? 6u : 4u) == 12u) {
    return -2;
  }
  static ulsch_demultiplex_impl demux;
  collecting_buffer b_sch, b_ack, b_c1, b_c2;
  pusch_codeword_buffer& cw = demux.demultiplex(b_sch, b_ack, b_c1, c);
  if (cfg[13] != 0) {
    demux.set_csi_part2(b_c2, cfg[13], cfg[14]);
  }
  pseudo_random_generator_impl prg;
  prg.init(c_init);
  for (unsigned off = 0; off < nof_llrs;) {
    const unsigned n = std::min(block, nof_llrs - off);
    span<log_likelihood_ratio> dst = cw.get_next_block_view(n);
    const unsigned m = std::min(n, (unsigned)dst.size());
    for (unsigned i = 0; i < m; ++i) {
      dst[i] = log_likelihood_ratio(llrs[off + i]);
    }
    dynamic_bit_buffer seq(m);
    prg.generate(seq);
    cw.on_new_block(dst.first(m), seq);
    off += m;
  }
  cw.on_end_codeword();
  const std::vector<int8_t>* outs[4] = {&b_sch.bits, &b_ack.bits, &b_c1.bits, &b_c2.bits};
  int8_t* dsts[4] = {sch, ack, csi1, csi2};
  for (int k = 0; k < 4; ++k) {
    sizes[k] = (unsigned)outs[k]->size();
    if (dsts[k] != nullptr) {
      std::copy(outs[k]->begin(), outs[k]->end(), dsts[k]);
    }
  }
  if (!b_sch.ended || (cfg[10] && !b_ack.ended) || (cfg[12] && !b_c1.ended) || (cfg[14] && !b_c2.ended)) {
    return -1;
  }
  return 0;
}

} // extern "C"

// ---- UL-SCH bit accounting (lib/ran/pusch/ulsch_info.cpp built from source) ------------

#include "ocudu/ran/pusch/ulsch_info.h"

extern "C" {

// get_ulsch_information.  w: tbs bits, modulation, O_ACK, O_CSI1, O_CSI2, PRBs, start
// symbol, symbols, DM-RS type, DM-RS mask, CDM groups without data, layers, contains_dc;
// f: target code rate (normalised), alpha, beta ACK, beta CSI1, beta CSI2.  out: UL-SCH
// bits, HARQ-ACK bits, reserved bits, CSI1 bits, CSI2 bits, HARQ-ACK REs, CSI1 REs,
// CSI2 REs, DC overlap bits, then (with a TB) the SCH base graph, codeblocks, bits per
// codeblock, lifting size, filler bits per codeblock, TB CRC bits.
void ref_ulsch_information(const unsigned* w, const float* f, unsigned* out)
{
  ulsch_configuration c;
  c.tbs = units::bits(w[0]);
  c.mcs_descr.modulation = static_cast<modulation_scheme>(w[1]);
  c.mcs_descr.target_code_rate = f[0] * 1024.0F;
  c.nof_harq_ack_bits = units::bits(w[2]);
  c.nof_csi_part1_bits = units::bits(w[3]);
  c.nof_csi_part2_bits = units::bits(w[4]);
  c.alpha_scaling = f[1];
  c.beta_offset_harq_ack = f[2];
  c.beta_offset_csi_part1 = f[3];
  c.beta_offset_csi_part2 = f[4];
  c.nof_rb = w[5];
  c.start_symbol_index = w[6];
  c.nof_symbols = w[7];
  c.dmrs_type = w[8] == 1 ? dmrs_config_type::type1 : dmrs_config_type::type2;
  c.dmrs_symbol_mask = bounded_bitset<MAX_NSYMB_PER_SLOT>(MAX_NSYMB_PER_SLOT);
  for (unsigned l = 0; l < MAX_NSYMB_PER_SLOT; ++l) {
    if ((w[9] >> l) & 1u) {
      c.dmrs_symbol_mask.set(l);
    }
  }
  c.nof_cdm_groups_without_data = w[10];
  c.nof_layers = w[11];
  c.contains_dc = w[12] != 0;
  const ulsch_information i = get_ulsch_information(c);
  out[0] = i.nof_ul_sch_bits.value();
  out[1] = i.nof_harq_ack_bits.value();
  out[2] = i.nof_harq_ack_rvd.value();
  out[3] = i.nof_csi_part1_bits.value();
  out[4] = i.nof_csi_part2_bits.value();
  out[5] = i.nof_harq_ack_re;
  out[6] = i.nof_csi_part1_re;
  out[7] = i.nof_csi_part2_re;
  out[8] = i.nof_dc_overlap_bits.value();
  if (i.sch.has_value()) {
    out[9] = i.sch->base_graph == ldpc_base_graph_type::BG1 ? 1 : 2;
    out[10] = i.sch->nof_cb;
    out[11] = i.sch->nof_bits_per_cb.value();
    out[12] = i.sch->lifting_size;
    out[13] = i.sch->nof_filler_bits_per_cb.value();
    out[14] = i.sch->tb_crc_size.value();
  }
}

} // extern "C"
