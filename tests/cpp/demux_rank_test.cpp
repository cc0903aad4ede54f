// Host-side check of the device UL-SCH demultiplexer's arithmetic (lh::demux_plan and the
// closed-form lh::demux_classify of csrc/lh_demux_rank.h) against the CPU oracle's
// bitset restatement, with no GPU: the kernel's per-unit work (load, classify,
// placeholders, stores) is replayed here thread by thread, with the placeholder sequence
// bits taken from the oracle's Gold sequence.
//
// stdin: one codeword per line, 16 integers: modulation, layers, PRBs, start symbol,
// symbols, DM-RS mask, DM-RS type, CDM groups without data, G_rvd, O_ACK, G_ACK, O_CSI1,
// G_CSI1, O_CSI2, G_CSI2, c_init.  stdout: "ok <n>" or the first mismatch; exit status 1
// on a mismatch.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lh_demux_rank.h"

extern "C" {
#include "../../oracle/ldpc_oracle.h"
}

using namespace lh;

namespace {

uint32_t placeholder(uint32_t v, uint32_t r, uint32_t u, uint32_t qm, uint32_t payload, const uint8_t* c)
{
  if (payload == 0) {
    return v;
  }
  uint32_t out = 0;
  for (uint32_t i = 0; i < u; ++i) {
    const uint32_t pos = r + i, k = pos % qm;
    uint32_t neg = 0;
    if (k == 1 && payload == 1) {
      neg = c[pos] ^ c[pos - 1];
    } else if (k >= 2) {
      neg = c[pos];
    }
    const int8_t b = (int8_t)(v >> (8 * i));
    out |= (uint32_t)(uint8_t)(neg ? (int8_t)-b : b) << (8 * i);
  }
  return out;
}

uint32_t load(const int8_t* p, uint32_t u)
{
  uint32_t v = 0;
  for (uint32_t i = 0; i < u; ++i) {
    v |= (uint32_t)(uint8_t)p[i] << (8 * i);
  }
  return v;
}

void store(int8_t* p, uint32_t v, uint32_t u)
{
  for (uint32_t i = 0; i < u; ++i) {
    p[i] = (int8_t)(v >> (8 * i));
  }
}

} // namespace

int main()
{
  unsigned w[16];
  int n = 0;
  while (true) {
    int got = 0;
    for (int k = 0; k < 16; ++k) {
      got += scanf("%u", &w[k]) == 1;
    }
    if (got != 16) {
      break;
    }
    lh_ulsch_demux_desc d = {};
    d.modulation = w[0];
    d.nof_layers = w[1];
    d.nof_prb = w[2];
    d.start_symbol_index = w[3];
    d.nof_symbols = w[4];
    d.dmrs_symbol_mask = w[5];
    d.dmrs_type = w[6];
    d.nof_cdm_groups_without_data = w[7];
    d.nof_harq_ack_rvd = w[8];
    d.nof_harq_ack_bits = w[9];
    d.nof_enc_harq_ack_bits = w[10];
    d.nof_csi_part1_bits = w[11];
    d.nof_enc_csi_part1_bits = w[12];
    d.nof_csi_part2_bits = w[13];
    d.nof_enc_csi_part2_bits = w[14];
    d.c_init = w[15];
    orc_demux_config oc = {(int)w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], w[9], w[10], w[11], w[12],
                           w[13], w[14]};
    demux_sym syms[14];
    unsigned ns = 0;
    lh_ulsch_demux_sizes sz;
    const char* why = nullptr;
    const int rc_plan = demux_plan(d, syms, &ns, &sz, &why);
    orc_demux_sizes osz;
    const int rc_orc = orc_ulsch_demultiplex(&oc, nullptr, 0, nullptr, nullptr, nullptr, nullptr, &osz);
    if ((rc_plan == 0) != (rc_orc == 0)) {
      printf("case %d: plan %d (%s) vs oracle %d\n", n, rc_plan, why ? why : "", rc_orc);
      return 1;
    }
    if (rc_plan != 0) {
      ++n;
      continue;
    }
    if (sz.nof_llrs != osz.nof_llrs || sz.nof_sch != osz.nof_sch || sz.nof_harq_ack != osz.nof_harq_ack ||
        sz.nof_csi_part1 != osz.nof_csi_part1 || sz.nof_csi_part2 != osz.nof_csi_part2) {
      printf("case %d: sizes differ\n", n);
      return 1;
    }
    std::vector<int8_t> in(sz.nof_llrs + 4);
    uint64_t st = 0x9e3779b97f4a7c15ull * (uint64_t)(n + 1);
    for (auto& v : in) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      v = (int8_t)((int)((st >> 33) % 241) - 120);
    }
    std::vector<uint8_t> c(sz.nof_llrs + 1);
    orc_scrambling_sequence(c.data(), d.c_init, 0, sz.nof_llrs);
    std::vector<int8_t> o[4], e[4];
    const uint32_t lens[4] = {sz.nof_sch, sz.nof_harq_ack, sz.nof_csi_part1, sz.nof_csi_part2};
    for (int k = 0; k < 4; ++k) {
      o[k].assign(lens[k] + 4, 99);
      e[k].assign(lens[k] + 4, 99);
    }
    orc_ulsch_demultiplex(&oc, in.data(), d.c_init, e[0].data(), e[1].data(), e[2].data(), e[3].data(), nullptr);
    for (unsigned k = 0; k < ns; ++k) {
      const demux_sym& s = syms[k];
      const uint32_t u = s.unit, nbpr = s.nbpr;
      const uint8_t* cs = c.data() + s.llr_in;
      for (uint32_t r = 0; r < s.M * nbpr; r += u) {
        const uint32_t v = load(in.data() + s.llr_in + r, u);
        if (s.plain) {
          store(o[0].data() + s.sch_out + r, v, u);
          continue;
        }
        const uint32_t x = r / nbpr, b = r - x * nbpr;
        const demux_re f = demux_classify(s, x);
        const uint32_t vz = f.punctured ? 0u : v;
        if (f.ack != kDemuxNone) {
          store(o[1].data() + s.ack_out + (uint64_t)f.ack * nbpr + b, placeholder(v, r, u, s.qm, s.ph_ack, cs), u);
        }
        if (f.c1 != kDemuxNone) {
          store(o[2].data() + s.c1_out + (uint64_t)f.c1 * nbpr + b, placeholder(v, r, u, s.qm, s.ph_c1, cs), u);
        }
        if (f.c2 != kDemuxNone) {
          store(o[3].data() + s.c2_out + (uint64_t)f.c2 * nbpr + b, placeholder(vz, r, u, s.qm, s.ph_c2, cs), u);
        }
        if (f.sch != kDemuxNone) {
          store(o[0].data() + s.sch_out + (uint64_t)f.sch * nbpr + b, vz, u);
        }
      }
    }
    for (int k = 0; k < 4; ++k) {
      if (o[k] != e[k]) {
        size_t i = 0;
        while (o[k][i] == e[k][i]) {
          ++i;
        }
        printf("case %d: field %d differs at %zu of %u (%d vs oracle %d)\n", n, k, i, lens[k], o[k][i], e[k][i]);
        return 1;
      }
    }
    ++n;
  }
  printf("ok %d\n", n);
  return 0;
}
