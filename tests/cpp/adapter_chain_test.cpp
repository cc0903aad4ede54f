// Chain test of the HIP adapters through the reference's class interfaces (mirrors the
// reference's ldpc_chain_test, tests/unittests/phy/upper/channel_coding/ldpc/ldpc_chain_test.cpp:262-367,
// and adds trajectory parity): for BG x Z x rate x modulation, random message + CRC,
// encode, rate match, map to LLRs (noise-free +-10 and AWGN), then
//   * ldpc_rate_dematcher "hip"  == oracle dematcher (new data and HARQ combine),
//   * ldpc_decoder "hip"         == oracle decoder (bits, iterations/nullopt),
//   * pusch batch decode_batch   == oracle per codeblock, HARQ buffers written back;
//   * decode_resident_softbits   == oracle TB chain (TX segmentation statement, RX
//     segmentation, dematch, decode, TB CRC) for C = 1 and C > 1 TBs, clean and noisy;
//     export_resident_softbits == oracle dematched buffers, once per decode; invalid
//     grants rejected without touching the output; timing statistics; the algorithm
//     selection ("learned" without tables decodes as the plain decoder).
// The oracle (oracle/liboracle.so) is the checker.  Exit code 0 = pass.
#include "../../oracle/ldpc_oracle.h"
#include "hip_ldpc_adapters.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <random>
#include <vector>

using namespace ocudu;

namespace {

int failures = 0;

#define EXPECT(cond, ...)                                                                                            \
  do {                                                                                                               \
    if (!(cond)) {                                                                                                   \
      ++failures;                                                                                                    \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                               \
      std::printf(__VA_ARGS__);                                                                                      \
      std::printf("\n");                                                                                             \
    }                                                                                                                \
  } while (0)

class oracle_crc : public crc_calculator
{
public:
  explicit oracle_crc(crc_generator_poly p) : poly(p) {}
  crc_calculator_checksum_t calculate_byte(span<const uint8_t> d) const override
  {
    return orc_crc_packed(static_cast<uint32_t>(poly), d.data(), static_cast<unsigned>(8 * d.size()));
  }
  crc_calculator_checksum_t calculate_bit(span<const uint8_t> d) const override
  {
    return orc_crc_unpacked(static_cast<uint32_t>(poly), d.data(), static_cast<unsigned>(d.size()));
  }
  crc_calculator_checksum_t calculate(const bit_buffer& d) const override
  {
    return orc_crc_packed(static_cast<uint32_t>(poly), d.get_buffer().data(), static_cast<unsigned>(d.size()));
  }
  crc_generator_poly get_generator_poly() const override { return poly; }

private:
  crc_generator_poly poly;
};

struct cb_case {
  int bg;
  unsigned Z, E, Qm, rv, F;
  double snr;
  bool clean;
};

std::vector<int8_t> make_llrs(const cb_case& c, std::mt19937_64& rng, std::vector<uint8_t>& msg)
{
  const unsigned K = (c.bg == 1 ? 22 : 10) * c.Z;
  const unsigned M = c.bg == 1 ? 46 : 42;
  const unsigned N = (c.bg == 1 ? 66 : 50) * c.Z;
  const unsigned crc_bits = K > 48 ? 24 : 16;
  const uint32_t poly = crc_bits == 24 ? ORC_CRC24B : ORC_CRC16;
  msg.assign(K, 0);
  const unsigned kd = K - c.F - crc_bits;
  for (unsigned i = 0; i != kd; ++i) {
    msg[i] = rng() & 1;
  }
  uint32_t r = orc_crc_unpacked(poly, msg.data(), kd);
  for (unsigned i = 0; i != crc_bits; ++i) {
    msg[kd + i] = (r >> (crc_bits - 1 - i)) & 1;
  }
  std::vector<uint8_t> cw((K / c.Z + M) * c.Z);
  orc_ldpc_encode(c.bg, c.Z, msg.data(), cw.data());
  std::vector<uint8_t> e(c.E);
  orc_rate_match(e.data(), c.E, cw.data() + 2 * c.Z, N, c.bg, c.Z, c.rv, c.Qm, 0, c.F);
  std::vector<int8_t> llr(c.E);
  if (c.clean) {
    for (unsigned i = 0; i != c.E; ++i) {
      llr[i] = e[i] ? -10 : 10;
    }
  } else {
    orc_modulate_awgn_demap(llr.data(), e.data(), c.E, c.Qm, std::pow(10.0, -c.snr / 10.0), rng());
  }
  return llr;
}

modulation_scheme mod_of(unsigned qm)
{
  return qm == 2 ? modulation_scheme::QPSK
                 : qm == 4 ? modulation_scheme::QAM16 : qm == 6 ? modulation_scheme::QAM64 : modulation_scheme::QAM256;
}

void append_crc(std::vector<uint8_t>& bits, uint32_t poly, unsigned order)
{
  const uint32_t r = orc_crc_unpacked(poly, bits.data(), static_cast<unsigned>(bits.size()));
  for (unsigned i = 0; i != order; ++i) {
    bits.push_back((r >> (order - 1 - i)) & 1);
  }
}

// One transport block through the resident path against the oracle TB chain.
void resident_case(pusch_resident_codeblock_decoder& batch, int bg, unsigned tbs, unsigned Qm, unsigned layers,
                   unsigned nsym, double snr, std::mt19937_64& rng, bool expect_pass)
{
  orc_segmentation seg;
  if (orc_segment_rx(&seg, bg, tbs, nsym, Qm, layers) != 0) {
    EXPECT(false, "segmentation of TB %u", tbs);
    return;
  }
  const unsigned C = seg.nof_segments, Z = seg.Z, K = seg.segment_length, F = seg.nof_filler_bits;
  const unsigned N = (bg == 1 ? 66 : 50) * Z, M = bg == 1 ? 46 : 42;
  // TX segmentation statement (TS38.212 5.1-5.2): TB + TB CRC, split, CB CRC24B, fillers.
  std::vector<uint8_t> tb_bits(tbs);
  for (auto& b : tb_bits) {
    b = rng() & 1;
  }
  std::vector<uint8_t> b = tb_bits;
  append_crc(b, tbs > 3824 ? ORC_CRC24A : ORC_CRC16, tbs > 3824 ? 24 : 16);
  std::vector<std::vector<uint8_t>> msgs(C);
  const unsigned D = K - F - (C > 1 ? 24 : 0);
  b.resize(static_cast<size_t>(C) * D, 0);
  for (unsigned c = 0; c != C; ++c) {
    msgs[c].assign(b.begin() + static_cast<size_t>(c) * D, b.begin() + static_cast<size_t>(c + 1) * D);
    if (C > 1) {
      append_crc(msgs[c], ORC_CRC24B, 24);
    }
    msgs[c].resize(K, 0);
  }
  size_t total = 0;
  for (unsigned c = 0; c != C; ++c) {
    total += seg.rm_length[c];
  }
  std::vector<int8_t> llr(total);
  std::vector<codeblock_metadata::cb_specific_metadata> cbs(C);
  std::vector<std::vector<int8_t>> want_rm(C);
  const uint32_t cb_poly = C > 1 ? ORC_CRC24B : (tbs > 3824 ? ORC_CRC24A : ORC_CRC16);
  const unsigned cb_crc = C > 1 ? 24 : (tbs > 3824 ? 24 : 16);
  unsigned passed = 0;
  for (unsigned c = 0; c != C; ++c) {
    std::vector<uint8_t> cw((K / Z + M) * Z), e(seg.rm_length[c]);
    orc_ldpc_encode(bg, Z, msgs[c].data(), cw.data());
    orc_rate_match(e.data(), seg.rm_length[c], cw.data() + 2 * Z, N, bg, Z, 0, Qm, 0, F);
    orc_modulate_awgn_demap(llr.data() + seg.cw_offset[c], e.data(), seg.rm_length[c], Qm, std::pow(10.0, -snr / 10.0),
                            rng());
    cbs[c].full_length = N;
    cbs[c].rm_length = seg.rm_length[c];
    cbs[c].nof_filler_bits = F;
    cbs[c].cw_offset = seg.cw_offset[c];
    cbs[c].nof_crc_bits = cb_crc;
    orc_rm_cfg rc{0, Qm, 0, F};
    want_rm[c].assign(N, 0);
    orc_rate_dematch(want_rm[c].data(), N, llr.data() + seg.cw_offset[c], seg.rm_length[c], 1, &rc);
    orc_dec_cfg oc{bg, Z, F, cb_crc, 6};
    std::vector<uint8_t> bits((K + 7) / 8);
    passed += orc_ldpc_decode(&oc, 0, 0, cb_poly, want_rm[c].data(), N, bits.data()) > 0 ? 1 : 0;
  }
  codeblock_metadata::tb_common_metadata common;
  common.base_graph = bg == 1 ? ldpc_base_graph_type::BG1 : ldpc_base_graph_type::BG2;
  common.lifting_size = static_cast<ldpc::lifting_size_t>(Z);
  common.rv = 0;
  common.mod = mod_of(Qm);
  common.Nref = 0;
  common.cw_length = static_cast<unsigned>(total);
  void* d_llr = nullptr;
  if (hipMalloc(&d_llr, total) != hipSuccess || hipMemcpy(d_llr, llr.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
    EXPECT(false, "TB %u: device LLR buffer", tbs);
    return;
  }
  std::vector<uint8_t> out(tbs / 8, 0xA5);
  const auto poly = static_cast<crc_generator_poly>(cb_poly);
  // Rejections first: nothing is touched and nothing can be exported.
  auto rej = batch.decode_resident_softbits(d_llr, total - 1, nullptr, common, cbs, out, poly, 6, nullptr);
  pusch_resident_codeblock_decoder::scrambling_config scr;
  auto rej2 = batch.decode_resident_softbits(d_llr, total, &scr, common, cbs, out, poly, 6, nullptr);
  auto rej3 = batch.decode_resident_softbits(nullptr, total, nullptr, common, cbs, out, poly, 6, nullptr);
  EXPECT(!rej.tb_crc_ok && !rej2.tb_crc_ok && !rej3.tb_crc_ok && rej.nof_iterations == 0, "TB %u: rejection", tbs);
  EXPECT(std::all_of(out.begin(), out.end(), [](uint8_t x) { return x == 0xA5; }), "TB %u: rejected output touched", tbs);
  batch.set_timing_enabled(true);
  auto r = batch.decode_resident_softbits(d_llr, total, nullptr, common, cbs, out, poly, 6, nullptr);
  const auto& st = batch.get_last_timing_stats();
  EXPECT(st.timing_enabled && st.nof_cbs == C && st.ldpc_decode_us > 0.0f && st.total_e2e_us >= st.ldpc_decode_us,
         "TB %u: timing statistics", tbs);
  batch.set_timing_enabled(false);
  // Oracle verdict: all codeblocks pass and the TB CRC holds (the data is what was sent).
  const bool want_ok = passed == C;
  EXPECT(r.all_cb_crcs_ok == want_ok && r.tb_crc_ok == want_ok, "TB %u: verdict %d/%d vs oracle %u of %u CBs", tbs,
         r.tb_crc_ok, r.all_cb_crcs_ok, passed, C);
  EXPECT(want_ok == expect_pass, "TB %u: case does not exercise what it should", tbs);
  if (want_ok) {
    std::vector<uint8_t> pk(tbs / 8, 0);
    for (unsigned i = 0; i != tbs; ++i) {
      pk[i / 8] |= tb_bits[i] << (7 - i % 8);
    }
    EXPECT(out == pk, "TB %u: TB bytes", tbs);
  }
  // Export: the oracle's dematched buffers, exactly once.
  std::vector<std::vector<log_likelihood_ratio>> host(C, std::vector<log_likelihood_ratio>(N, log_likelihood_ratio(77)));
  std::vector<span<log_likelihood_ratio>> spans;
  for (auto& h : host) {
    spans.emplace_back(h.data(), h.size());
  }
  EXPECT(batch.export_resident_softbits(spans), "TB %u: export", tbs);
  for (unsigned c = 0; c != C; ++c) {
    EXPECT(std::memcmp(host[c].data(), want_rm[c].data(), N) == 0, "TB %u: exported soft bits of CB %u", tbs, c);
  }
  EXPECT(!batch.export_resident_softbits(spans), "TB %u: second export", tbs);
  // "learned" without tables: the plain decoder, same verdict.
  batch.set_ldpc_decoder_algorithm("learned");
  auto r2 = batch.decode_resident_softbits(d_llr, total, nullptr, common, cbs, out, poly, 6, nullptr);
  batch.set_ldpc_decoder_algorithm("min_sum");
  EXPECT(r2.tb_crc_ok == r.tb_crc_ok && r2.nof_iterations == r.nof_iterations, "TB %u: algorithm fallback", tbs);
  EXPECT(hipFree(d_llr) == hipSuccess, "TB %u: hipFree", tbs);
}

} // namespace

int main()
{
  hip_backend_config hcfg;
  auto dec_factory = create_ldpc_decoder_factory_hip(hcfg, {false, false});
  auto dm_factory = create_ldpc_rate_dematcher_factory_hip(hcfg);
  auto batch = create_pusch_batch_decoder_hip(hcfg);
  auto dec = dec_factory->create();
  auto dm = dm_factory->create();
  std::printf("decoder type: %s\n", get_ldpc_decoder_type_str_hip());

  std::mt19937_64 rng(2026);
  std::vector<cb_case> cases;
  for (int bg = 1; bg <= 2; ++bg) {
    for (unsigned Z : {8u, 36u, 104u, 208u, 384u}) {
      for (double rate : {0.2, 0.4, 0.6, 0.9}) {
        for (unsigned Qm : {2u, 6u}) {
          const unsigned K = (bg == 1 ? 22 : 10) * Z;
          unsigned E = static_cast<unsigned>(K / rate) / Qm * Qm;
          const unsigned F = Z >= 36 ? 16 : 0;
          const bool clean = (cases.size() % 2) == 0;
          const double snr = Qm == 2 ? 3.0 : 13.0;
          cases.push_back({bg, Z, E, Qm, static_cast<unsigned>(cases.size() % 4), F, snr, clean});
        }
      }
    }
  }

  std::vector<std::vector<int8_t>> llrs(cases.size());
  std::vector<std::vector<uint8_t>> msgs(cases.size());
  std::vector<std::vector<log_likelihood_ratio>> harq(cases.size());
  std::vector<std::vector<uint8_t>> bits(cases.size());
  std::vector<described_rx_codeblock> described;
  std::vector<codeblock_metadata> mds(cases.size());
  std::vector<int> expect_iters(cases.size());
  std::vector<std::vector<uint8_t>> expect_bits(cases.size());
  std::vector<std::vector<int8_t>> expect_rm(cases.size());

  for (size_t k = 0; k != cases.size(); ++k) {
    const cb_case& c = cases[k];
    llrs[k] = make_llrs(c, rng, msgs[k]);
    const unsigned K = (c.bg == 1 ? 22 : 10) * c.Z;
    const unsigned N = (c.bg == 1 ? 66 : 50) * c.Z;
    const unsigned crc_bits = K > 48 ? 24 : 16;
    codeblock_metadata& md = mds[k];
    md.tb_common.base_graph = c.bg == 1 ? ldpc_base_graph_type::BG1 : ldpc_base_graph_type::BG2;
    md.tb_common.lifting_size = c.Z;
    md.tb_common.rv = c.rv;
    md.tb_common.mod = mod_of(c.Qm);
    md.cb_specific.full_length = N;
    md.cb_specific.rm_length = c.E;
    md.cb_specific.nof_filler_bits = c.F;
    md.cb_specific.nof_crc_bits = crc_bits;

    // Oracle expectations.
    orc_rm_cfg rc{c.rv, c.Qm, 0, c.F};
    expect_rm[k].assign(N, 0);
    orc_rate_dematch(expect_rm[k].data(), N, llrs[k].data(), c.E, 1, &rc);
    orc_dec_cfg oc{c.bg, c.Z, c.F, crc_bits, 6};
    expect_bits[k].assign((K + 7) / 8, 0);
    expect_iters[k] = static_cast<int>(orc_ldpc_decode(&oc, 0, 0, crc_bits == 24 ? ORC_CRC24B : ORC_CRC16,
                                                       expect_rm[k].data(), N, expect_bits[k].data()));

    // Rate dematcher "hip".
    std::vector<log_likelihood_ratio> rm(N, log_likelihood_ratio(5));
    dm->rate_dematch(rm, span<const log_likelihood_ratio>(reinterpret_cast<const log_likelihood_ratio*>(llrs[k].data()), c.E),
                     true, md);
    EXPECT(std::memcmp(rm.data(), expect_rm[k].data(), N) == 0, "dematch case %zu", k);
    // HARQ combine with itself (rv repeated) against the oracle.
    std::vector<int8_t> comb = expect_rm[k];
    orc_rate_dematch(comb.data(), N, llrs[k].data(), c.E, 0, &rc);
    dm->rate_dematch(rm, span<const log_likelihood_ratio>(reinterpret_cast<const log_likelihood_ratio*>(llrs[k].data()), c.E),
                     false, md);
    EXPECT(std::memcmp(rm.data(), comb.data(), N) == 0, "combine case %zu", k);

    // Decoder "hip".
    std::vector<uint8_t> out((K + 7) / 8, 0);
    bit_buffer ob(out.data(), K);
    oracle_crc crc(crc_bits == 24 ? crc_generator_poly::CRC24B : crc_generator_poly::CRC16);
    ldpc_decoder::configuration cfg;
    cfg.base_graph = md.tb_common.base_graph;
    cfg.lifting_size = c.Z;
    cfg.nof_filler_bits = c.F;
    cfg.nof_crc_bits = crc_bits;
    cfg.max_iterations = 6;
    std::optional<unsigned> r = dec->decode(
        ob, span<const log_likelihood_ratio>(reinterpret_cast<const log_likelihood_ratio*>(expect_rm[k].data()), N), &crc, cfg);
    EXPECT((r ? static_cast<int>(*r) : 0) == expect_iters[k], "decode case %zu iters %d vs %d", k, r ? (int)*r : 0,
           expect_iters[k]);
    EXPECT(out == expect_bits[k], "decode case %zu bits", k);
    if (c.clean && c.rv == 0) {
      // Noise-free and self-decodable (rv0): decodes to the transmitted message (chain test).
      std::vector<uint8_t> pk((K + 7) / 8, 0);
      for (unsigned i = 0; i != K; ++i) {
        pk[i / 8] |= msgs[k][i] << (7 - i % 8);
      }
      EXPECT(r.has_value() && out == pk, "clean case %zu not recovered", k);
    }

    harq[k].assign(N, log_likelihood_ratio(0));
    bits[k].assign((K + 7) / 8, 0);
  }

  // PUSCH batch backend over all cases at once (mixed BG/Z/E/Qm).
  std::vector<bit_buffer> bbs;
  std::vector<span<log_likelihood_ratio>> rms;
  for (size_t k = 0; k != cases.size(); ++k) {
    described.push_back({span<const log_likelihood_ratio>(reinterpret_cast<const log_likelihood_ratio*>(llrs[k].data()),
                                                          cases[k].E),
                         mds[k]});
    const unsigned K = (cases[k].bg == 1 ? 22 : 10) * cases[k].Z;
    bbs.emplace_back(bits[k].data(), K);
    rms.emplace_back(harq[k].data(), harq[k].size());
  }
  // All cases use CRC24B or CRC16 per size; batch calls are made per CRC polynomial.
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<described_rx_codeblock> d;
    std::vector<bit_buffer> b;
    std::vector<span<log_likelihood_ratio>> h;
    std::vector<size_t> idx;
    for (size_t k = 0; k != cases.size(); ++k) {
      const unsigned K = (cases[k].bg == 1 ? 22 : 10) * cases[k].Z;
      if ((K > 48) == (pass == 0)) {
        d.push_back(described[k]);
        b.push_back(bbs[k]);
        h.push_back(rms[k]);
        idx.push_back(k);
      }
    }
    if (d.empty()) {
      continue;
    }
    auto res = batch->decode_batch(d, b, h, true, pass == 0 ? crc_generator_poly::CRC24B : crc_generator_poly::CRC16,
                                   true, 6);
    for (size_t j = 0; j != idx.size(); ++j) {
      const size_t k = idx[j];
      EXPECT(res[j].crc_ok == (expect_iters[k] > 0), "batch case %zu crc", k);
      EXPECT(!res[j].crc_ok || (int)res[j].nof_iterations == expect_iters[k], "batch case %zu iters", k);
      EXPECT(bits[k] == expect_bits[k], "batch case %zu bits", k);
      EXPECT(std::memcmp(harq[k].data(), expect_rm[k].data(), harq[k].size()) == 0, "batch case %zu harq", k);
    }
  }
  // Resident path: C = 1 (CRC16, CRC24A) and C > 1 TBs, BG1 and BG2, passing and failing.
  resident_case(*batch, 2, 1000, 2, 1, 1440, 8.0, rng, true);
  resident_case(*batch, 1, 8192, 2, 1, 4608, 12.0, rng, true);
  resident_case(*batch, 1, 42016, 4, 2, 17280, 14.0, rng, true);
  resident_case(*batch, 2, 3368, 4, 2, 2880, 12.0, rng, true);
  resident_case(*batch, 1, 42016, 4, 2, 17280, 3.0, rng, false);
  std::printf("%zu cases, %d failures\n", cases.size(), failures);
  return failures == 0 ? 0 : 1;
}
