// Host side of the UL-SCH demultiplexer: the per-OFDM-symbol field spacings of TS38.212
// 6.2.7 as ulsch_demultiplex_impl::configure_current_ofdm_symbol computes them
// (lib/phy/upper/channel_processors/pusch/ulsch_demultiplex_impl.cpp:313-449), with CSI
// Part 2 known from the first symbol.  Only scalars per symbol: the device kernel turns
// them into each RE's field and rank with closed forms, so no RE index list is built or
// uploaded.
#include <hip/hip_runtime.h>

#include "ocudu_hip_ldpc.h"
#include "lh_internal.h"

namespace lh {

namespace {

// Spreading of one field over `avail` REs with `remainder` REs still to place
// (ulsch_demultiplex_impl.cpp:343-349, and the same lines of every later step): all
// available REs while they do not cover the remainder, else every (avail / remainder)-th.
void spread(unsigned remainder, unsigned avail, unsigned& d, unsigned& m)
{
  d = 1;
  m = avail;
  if (remainder < avail) {
    d = avail / remainder;
    m = remainder;
  }
}

unsigned gcd(unsigned a, unsigned b)
{
  while (b != 0) {
    const unsigned t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// Placeholder payload bits of a field: 1 or 2 when its soft bits carry placeholders
// (on_uci_placeholder_1bit/_2bit, :86-176, do nothing for one bit per symbol).
uint8_t placeholder_bits(unsigned payload_bits, unsigned qm)
{
  return (qm > 1 && (payload_bits == 1 || payload_bits == 2)) ? (uint8_t)payload_bits : 0;
}

} // namespace

int demux_plan(const lh_ulsch_demux_desc& d, demux_sym* syms, unsigned* nof_syms, lh_ulsch_demux_sizes* sizes,
               const char** why)
{
  const unsigned mod = d.modulation;
  if (mod != 0 && mod != 1 && mod != 2 && mod != 4 && mod != 6 && mod != 8) {
    *why = "modulation is not 0, 1, 2, 4, 6 or 8";
    return -1;
  }
  const unsigned qm = mod == 0 ? 1u : mod;
  if (d.nof_layers < 1 || d.nof_layers > 4 || d.nof_prb < 1 || d.nof_prb > 275) {
    *why = "layers not in 1..4 or PRBs not in 1..275";
    return -1;
  }
  const unsigned nbpr = qm * d.nof_layers;
  if (d.nof_symbols == 0 || d.start_symbol_index + d.nof_symbols > 14) {
    *why = "OFDM symbols beyond the slot";
    return -1;
  }
  const unsigned mask = d.dmrs_symbol_mask & 0x3fffu;
  if ((d.dmrs_type != 1 && d.dmrs_type != 2) || d.nof_cdm_groups_without_data < 1 ||
      d.nof_cdm_groups_without_data > (d.dmrs_type == 1 ? 2u : 3u) || mask == 0) {
    *why = "invalid DM-RS configuration";
    return -1;
  }
  // get_ulsch_demultiplex_l1 / _l1_csi (:15-35): the first symbol without DM-RS after the
  // first DM-RS symbol, and the first symbol without DM-RS.
  unsigned l1 = (unsigned)__builtin_ctz(mask);
  while (l1 < 14 && ((mask >> l1) & 1u)) {
    ++l1;
  }
  unsigned l1_csi = 0;
  while (l1_csi < 14 && ((mask >> l1_csi) & 1u)) {
    ++l1_csi;
  }
  if (l1 >= 14) {
    *why = "no OFDM symbol without DM-RS after the first DM-RS symbol";
    return -1;
  }
  const unsigned G_rvd = d.nof_harq_ack_rvd, G_ack = d.nof_enc_harq_ack_bits, G_c1 = d.nof_enc_csi_part1_bits,
                 G_c2 = d.nof_enc_csi_part2_bits;
  if ((d.nof_harq_ack_bits == 0) != (G_ack == 0) || (d.nof_csi_part1_bits == 0) != (G_c1 == 0) ||
      (d.nof_csi_part2_bits == 0) != (G_c2 == 0)) {
    *why = "a UCI field with payload bits but no encoded bits, or the reverse";
    return -1;
  }
  if (G_ack % nbpr != 0 || G_c1 % nbpr != 0 || G_c2 % nbpr != 0) {
    *why = "encoded UCI bits not a multiple of the soft bits per RE";
    return -1;
  }
  // get_ulsch_demultiplex_nof_re_prb_dmrs (:37-53).
  const unsigned M_dmrs = (12u - d.nof_cdm_groups_without_data * (d.dmrs_type == 1 ? 6u : 4u)) * d.nof_prb;
  const bool ack_small = d.nof_harq_ack_bits <= 2;
  unsigned unit = 4;
  for (const uint64_t off : {d.llr_offset, d.sch_offset, d.harq_ack_offset, d.csi_part1_offset, d.csi_part2_offset,
                             (uint64_t)nbpr}) {
    while (off % unit != 0) {
      unit >>= 1;
    }
  }
  unsigned m_rvd = 0, m_ack = 0, m_c1 = 0, m_c2 = 0, n = 0;
  uint64_t q = 0, o_sch = 0, o_ack = 0, o_c1 = 0, o_c2 = 0;
  const bool need_seq = placeholder_bits(d.nof_harq_ack_bits, qm) || placeholder_bits(d.nof_csi_part1_bits, qm) ||
                        placeholder_bits(d.nof_csi_part2_bits, qm);
  for (unsigned l = d.start_symbol_index; l < d.start_symbol_index + d.nof_symbols; ++l) {
    const bool dm = (mask >> l) & 1u;
    const unsigned M = dm ? M_dmrs : d.nof_prb * 12u;
    if (M == 0) {
      continue;
    }
    demux_sym s = {};
    s.M = M;
    s.nbpr = (uint16_t)nbpr;
    s.qm = (uint8_t)qm;
    s.unit = (uint8_t)unit;
    unsigned M_uci = dm ? 0 : M, dd = 0, m = 0;
    // Step 1 (:340-356): reserved REs, from all REs.
    const unsigned rem_rvd = (G_rvd - m_rvd) / nbpr;
    if (l >= l1 && M_uci > 0 && rem_rvd > 0) {
      spread(rem_rvd, M_uci, dd, m);
      s.dr = (uint16_t)dd;
      s.nr = (uint16_t)m;
      m_rvd += m * nbpr;
    }
    // Step 2 (:358-380): HARQ-ACK of more than 2 bits, from all REs.
    const unsigned rem_ack = (G_ack - m_ack) / nbpr;
    if (l >= l1 && M_uci > 0 && !ack_small && rem_ack > 0) {
      spread(rem_ack, M_uci, dd, m);
      s.da = (uint16_t)dd;
      s.na = (uint16_t)m;
      M_uci -= m;
      m_ack += m * nbpr;
    }
    // Step 3 (:382-407): CSI Part 1 on the REs in neither A nor R, spread over the count
    // M_uci - M_rvd (common REs of A and R ignored, as the reference counts).
    const unsigned rem_c1 = (G_c1 - m_c1) / nbpr;
    const unsigned M_rvd = s.nr;
    if (l >= l1_csi && M_uci != M_rvd && rem_c1 > 0) {
      if (M_uci < M_rvd) {
        *why = "CSI Part 1 in a symbol whose reserved REs exceed its UCI REs";
        return -1;
      }
      spread(rem_c1, M_uci - M_rvd, dd, m);
      s.d1 = (uint16_t)dd;
      s.n1 = (uint16_t)m;
      M_uci -= m;
      m_c1 += m * nbpr;
    }
    // Step 3bis (:431-449): CSI Part 2 on the UCI REs left.
    const unsigned rem_c2 = (G_c2 - m_c2) / nbpr;
    if (l >= l1_csi && M_uci > 0 && rem_c2 > 0) {
      spread(rem_c2, M_uci, dd, m);
      s.d2 = (uint16_t)dd;
      s.n2 = (uint16_t)m;
      m_c2 += m * nbpr;
    }
    // Step 5 (:413-426): HARQ-ACK of up to 2 bits on the reserved REs.
    if (M_rvd > 0 && ack_small && rem_ack > 0) {
      spread(rem_ack, M_rvd, dd, m);
      s.ds = (uint16_t)dd;
      s.ns = (uint16_t)m;
      m_ack += m * nbpr;
    }
    if (s.na > 0 && s.nr > 0) {
      const uint64_t lcm = (uint64_t)s.da / gcd(s.da, s.dr) * s.dr;
      s.lcm_ar = lcm > 0xffffffffu ? 0xffffffffu : (uint32_t)lcm;
      const unsigned la = (unsigned)s.na * s.da, lr = (unsigned)s.nr * s.dr;
      s.lim_ar = la < lr ? la : lr;
    }
    s.plain = (s.nr | s.na | s.n1 | s.n2) == 0;
    const unsigned n_ack = s.na + s.ns, n_sch = M - s.na - s.n1 - s.n2;
    s.llr_in = d.llr_offset + q;
    s.sch_out = d.sch_offset + o_sch;
    s.ack_out = d.harq_ack_offset + o_ack;
    s.c1_out = d.csi_part1_offset + o_c1;
    s.c2_out = d.csi_part2_offset + o_c2;
    if (!s.plain) {
      s.ph_ack = placeholder_bits(d.nof_harq_ack_bits, qm);
      s.ph_c1 = placeholder_bits(d.nof_csi_part1_bits, qm);
      s.ph_c2 = placeholder_bits(d.nof_csi_part2_bits, qm);
      if (need_seq && syms != nullptr) {
        demod_tile_states(d.c_init, q, (int)qm, 1, &s.x1, &s.x2);
      }
    }
    if (syms != nullptr) {
      syms[n] = s;
    }
    ++n;
    q += (uint64_t)M * nbpr;
    o_sch += (uint64_t)n_sch * nbpr;
    o_ack += (uint64_t)n_ack * nbpr;
    o_c1 += (uint64_t)s.n1 * nbpr;
    o_c2 += (uint64_t)s.n2 * nbpr;
  }
  // on_end_codeword (:302-308): every UCI field complete.
  if (m_ack != G_ack || m_c1 != G_c1 || m_c2 != G_c2) {
    *why = "a UCI field does not fit its REs";
    return -1;
  }
  if (nof_syms != nullptr) {
    *nof_syms = n;
  }
  if (sizes != nullptr) {
    sizes->nof_llrs = (uint32_t)q;
    sizes->nof_sch = (uint32_t)o_sch;
    sizes->nof_harq_ack = (uint32_t)o_ack;
    sizes->nof_csi_part1 = (uint32_t)o_c1;
    sizes->nof_csi_part2 = (uint32_t)o_c2;
  }
  return 0;
}

} // namespace lh
