// Chain test of the HIP adapters through the reference's class interfaces (mirrors the
// reference's ldpc_chain_test, tests/unittests/phy/upper/channel_coding/ldpc/ldpc_chain_test.cpp:262-367,
// and adds trajectory parity): for BG x Z x rate x modulation, random message + CRC,
// encode, rate match, map to LLRs (noise-free +-10 and AWGN), then
//   * ldpc_rate_dematcher "hip"  == oracle dematcher (new data and HARQ combine),
//   * ldpc_decoder "hip"         == oracle decoder (bits, iterations/nullopt),
//   * pusch batch decode_batch   == oracle per codeblock, HARQ buffers written back.
// The oracle (oracle/liboracle.so) is the checker.  Exit code 0 = pass.
#include "../../oracle/ldpc_oracle.h"
#include "hip_ldpc_adapters.h"

#include <cmath>
#include <cstdio>
#include <cstring>
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
  uint32_t calculate(const bit_buffer& d) const override
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
  std::printf("%zu cases, %d failures\n", cases.size(), failures);
  return failures == 0 ? 0 : 1;
}
