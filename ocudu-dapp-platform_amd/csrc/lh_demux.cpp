// Host side of the UL-SCH demultiplexer: the per-OFDM-symbol field spacings of TS38.212
// 6.2.7 as ulsch_demultiplex_impl::configure_current_ofdm_symbol computes them
// (lib/phy/upper/channel_processors/pusch/ulsch_demultiplex_impl.cpp:313-449), with CSI
// Part 2 known from the first symbol.  Only scalars per symbol: the device kernel turns
// them into each RE's field and rank with closed forms, so no RE index list is built or
// uploaded.
#include <hip/hip_runtime.h>

#include <cmath>

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


// ---- UL-SCH bit accounting ------------------------------------------------------------

namespace {

// get_uci_crc_size (include/ocudu/ran/uci/uci_info.h:36-47).
unsigned uci_crc_size(unsigned payload)
{
  return payload < 12 ? 0u : (payload < 20 ? 6u : 11u);
}

// The reference converts std::ceil of a float expression to unsigned.
unsigned ceil_u(float v)
{
  return (unsigned)std::ceil(v);
}

// calculate_nof_re_harq_ack (lib/ran/pusch/ulsch_info.cpp:12-31) and its UCI-only form (:33-52).
unsigned harq_ack_res(bool with_sch, unsigned bits, float beta, unsigned nre_uci, unsigned sum_cb, float R, unsigned qm,
                      float alpha, unsigned nre_l0)
{
  if (bits == 0) {
    return 0;
  }
  const float payload = (float)(bits + uci_crc_size(bits));
  const unsigned left = with_sch ? ceil_u(payload * beta * (float)nre_uci / (float)sum_cb)
                                 : ceil_u(payload * beta / (R * (float)qm));
  const unsigned right = ceil_u(alpha * (float)nre_l0);
  return left < right ? left : right;
}

// The NR lifting sizes (TS38.212 Table 5.3.2-1), ascending.
constexpr unsigned kZs[kNofLiftingSizes] = {2,   3,   4,   5,   6,   7,   8,   9,   10,  11,  12,  13,  14,
                                            15,  16,  18,  20,  22,  24,  26,  28,  30,  32,  36,  40,  44,
                                            48,  52,  56,  60,  64,  72,  80,  88,  96,  104, 112, 120, 128,
                                            144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

// get_sch_segmentation_info (lib/ran/sch/sch_segmentation.cpp:11-42): base graph by
// get_ldpc_base_graph (include/ocudu/ran/sch/ldpc_base_graph.h:20-30), codeblocks, lifting
// size by compute_lifting_size (sch_segmentation.h:49-95).
void sch_info(unsigned tbs, float R, lh_ulsch_info& o)
{
  const unsigned crc = tbs <= 3824u ? 16u : 24u;
  const bool bg2 = tbs <= 292u || R <= 0.25F || (tbs <= 3824u && R <= 0.67F);
  const unsigned B = tbs + crc;
  const unsigned k_cb = bg2 ? 3840u : 8448u;
  const unsigned C = B <= k_cb ? 1u : (B + k_cb - 25u) / (k_cb - 24u);
  const unsigned payload_per_cb = B / C + (C > 1 ? 24u : 0u);
  unsigned ref = 22;
  if (bg2) {
    ref = B > 640u ? 10u : (B > 560u ? 9u : (B > 192u ? 8u : 6u));
  }
  const unsigned out_bits = B + (C > 1 ? 24u * C : 0u);
  unsigned Z = 0;
  for (unsigned i = 0; i < (unsigned)kNofLiftingSizes && Z == 0; ++i) {
    if (kZs[i] * ref * C >= out_bits) {
      Z = kZs[i];
    }
  }
  o.sch_base_graph = bg2 ? 2u : 1u;
  o.sch_tb_crc_size = crc;
  o.sch_nof_cb = C;
  o.sch_lifting_size = Z;
  o.sch_nof_bits_per_cb = (bg2 ? 10u : 22u) * Z;
  o.sch_nof_filler_bits_per_cb = o.sch_nof_bits_per_cb - payload_per_cb;
}

} // namespace

int ulsch_information(const lh_ulsch_config& c, lh_ulsch_info& o, const char** why)
{
  o = lh_ulsch_info{};
  const unsigned mask = c.dmrs_symbol_mask & 0x3fffu;
  const unsigned mod = c.modulation;
  if (mod != 0 && mod != 1 && mod != 2 && mod != 4 && mod != 6 && mod != 8) {
    *why = "modulation is not 0, 1, 2, 4, 6 or 8";
    return -1;
  }
  if ((c.dmrs_type != 1 && c.dmrs_type != 2) || c.nof_cdm_groups_without_data < 1 ||
      c.nof_cdm_groups_without_data > (c.dmrs_type == 1 ? 2u : 3u)) {
    *why = "invalid DM-RS type or CDM groups without data";
    return -1;
  }
  if (mask == 0 || c.nof_symbols == 0 || c.start_symbol_index + c.nof_symbols > 14 ||
      (unsigned)__builtin_ctz(mask) < c.start_symbol_index ||
      31u - (unsigned)__builtin_clz(mask) >= c.start_symbol_index + c.nof_symbols) {
    *why = "DM-RS symbols outside the allocation";
    return -1;
  }
  if (c.tbs_bits == 0 && !(c.target_code_rate > 0.0F)) {
    *why = "a grant without UL-SCH needs a code rate";
    return -1;
  }
  const bool with_sch = c.tbs_bits > 0;
  if (with_sch) {
    sch_info(c.tbs_bits, c.target_code_rate, o);
  }
  const unsigned nsym_dmrs = (unsigned)__builtin_popcount(mask);
  const unsigned re_dmrs_rb = nsym_dmrs * c.nof_cdm_groups_without_data * (c.dmrs_type == 1 ? 6u : 4u);
  const unsigned nre_total = c.nof_rb * (c.nof_symbols * 12u - re_dmrs_rb);
  const unsigned nre_uci = (c.nof_symbols - nsym_dmrs) * c.nof_rb * 12u;
  const unsigned qm = mod == 0 ? 1u : mod;
  unsigned nre_l0 = 0;
  for (unsigned l = (unsigned)__builtin_ctz(mask); l < c.start_symbol_index + c.nof_symbols; ++l) {
    nre_l0 += ((mask >> l) & 1u) ? 0u : c.nof_rb * 12u;
  }
  const unsigned sum_cb = o.sch_nof_cb * o.sch_nof_bits_per_cb;
  const float R = c.target_code_rate;
  const unsigned O_ack = c.nof_harq_ack_bits, O_c1 = c.nof_csi_part1_bits, O_c2 = c.nof_csi_part2_bits;
  o.nof_harq_ack_re =
      harq_ack_res(with_sch, O_ack, c.beta_offset_harq_ack, nre_uci, sum_cb, R, qm, c.alpha_scaling, nre_l0);
  // Reservation sized for two bits when fewer are sent (:225-252).
  unsigned rvd_re = 0;
  if (O_ack < 2) {
    rvd_re = harq_ack_res(with_sch, 2, c.beta_offset_harq_ack, nre_uci, sum_cb, R, qm, c.alpha_scaling, nre_l0);
  } else if (O_ack == 2) {
    rvd_re = o.nof_harq_ack_re;
  }
  // CSI Part 1 (:56-104) against the reserved REs when HARQ-ACK has at most 2 bits.
  const unsigned ack_c1 = O_ack <= 2 ? rvd_re : o.nof_harq_ack_re;
  if (O_c1 != 0) {
    const float payload = (float)(O_c1 + uci_crc_size(O_c1));
    unsigned left, right;
    if (with_sch) {
      left = ceil_u(payload * c.beta_offset_csi_part1 * (float)nre_uci / (float)sum_cb);
      right = ceil_u(c.alpha_scaling * (float)nre_uci) - ack_c1;
    } else if (O_c2 == 0) {
      left = right = nre_uci - ack_c1;
    } else {
      left = ceil_u(payload * c.beta_offset_csi_part1 / (R * (float)qm));
      right = nre_uci - ack_c1;
    }
    o.nof_csi_part1_re = left < right ? left : right;
  }
  // CSI Part 2 (:106-146), HARQ-ACK REs not counted when at most 2 bits.
  const unsigned ack_c2 = O_ack <= 2 ? 0u : o.nof_harq_ack_re;
  if (O_c2 != 0) {
    if (with_sch) {
      const float payload = (float)(O_c2 + uci_crc_size(O_c2));
      const unsigned left = ceil_u(payload * c.beta_offset_csi_part2 * (float)nre_uci / (float)sum_cb);
      const unsigned right = ceil_u(c.alpha_scaling * (float)nre_uci) - ack_c2 - o.nof_csi_part1_re;
      o.nof_csi_part2_re = left < right ? left : right;
    } else {
      o.nof_csi_part2_re = nre_uci - ack_c2 - o.nof_csi_part1_re;
    }
  }
  const unsigned ack_sch = O_ack > 2 ? o.nof_harq_ack_re : 0u;
  const unsigned nre_sch = with_sch ? nre_total - ack_sch - o.nof_csi_part1_re - o.nof_csi_part2_re : 0u;
  const unsigned bpr = c.nof_layers * qm;
  o.nof_ul_sch_bits = nre_sch * bpr;
  o.nof_harq_ack_bits = o.nof_harq_ack_re * bpr;
  o.nof_harq_ack_rvd = rvd_re * bpr;
  o.nof_csi_part1_bits = o.nof_csi_part1_re * bpr;
  o.nof_csi_part2_bits = o.nof_csi_part2_re * bpr;
  o.nof_dc_overlap_bits = c.contains_dc ? c.nof_symbols * qm : 0u;
  return 0;
}

} // namespace lh
