// Receive-side codeblock segmentation of a transport block (host code of the C ABI).
// TS38.212 5.2.2 as the reference computes it: TB CRC size and number of codeblocks
// (include/ocudu/ran/sch/sch_segmentation.h:40-66), lifting size (:69-110), information bits
// per codeblock and rate-matched lengths (lib/phy/upper/channel_coding/ldpc/
// ldpc_segmenter_rx_impl.cpp:33-85, ldpc_segmenter_helpers.h:63-100).
#include "ocudu_hip_ldpc.h"

#include <cstdarg>
#include <cstdio>

#include "lh_internal.h"

namespace {

const unsigned kSizes[] = {2,  3,  4,  5,  6,  7,  8,   9,   10,  11,  12,  13,  14,  15,  16,  18,  20,
                           22, 24, 26, 28, 30, 32, 36,  40,  44,  48,  52,  56,  60,  64,  72,  80,  88,
                           96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

unsigned tb_crc_bits(unsigned tbs) { return tbs <= 3824 ? 16 : 24; }

unsigned nof_codeblocks(unsigned tbs, unsigned bg)
{
  const unsigned b = tbs + tb_crc_bits(tbs);
  const unsigned kcb = bg == 1 ? 8448 : 3840;
  return b <= kcb ? 1 : (b + (kcb - 24) - 1) / (kcb - 24);
}

unsigned lifting_size(unsigned tbs, unsigned bg, unsigned C)
{
  const unsigned b = tbs + tb_crc_bits(tbs);
  unsigned kb = 22;
  if (bg == 2) {
    kb = b > 640 ? 10 : (b > 560 ? 9 : (b > 192 ? 8 : 6));
  }
  const unsigned bout = b + (C > 1 ? 24 * C : 0);
  for (unsigned z : kSizes) {
    if (z * kb * C >= bout) {
      return z;
    }
  }
  return 0;
}

} // namespace

extern "C" int lh_segment_rx(unsigned base_graph, unsigned tbs_bits, unsigned modulation_order, unsigned nof_layers,
                             unsigned nof_ch_symbols, lh_segmentation* seg, uint32_t* rm_lengths,
                             uint32_t* cw_offsets, unsigned max_cbs)
{
  if (seg == nullptr || (base_graph != 1 && base_graph != 2) || tbs_bits == 0 || tbs_bits % 8 != 0 ||
      tbs_bits + 24 > 1277992 || nof_layers < 1 || nof_layers > 4 || nof_ch_symbols % nof_layers != 0 ||
      (modulation_order != 1 && modulation_order != 2 && modulation_order != 4 && modulation_order != 6 &&
       modulation_order != 8)) {
    return LH_ERR_INVALID;
  }
  const unsigned C = nof_codeblocks(tbs_bits, base_graph);
  if (C > max_cbs || (C > 0 && (rm_lengths == nullptr || cw_offsets == nullptr))) {
    return LH_ERR_CAPACITY;
  }
  const unsigned Z = lifting_size(tbs_bits, base_graph, C);
  if (Z == 0) {
    return LH_ERR_INVALID;
  }
  const unsigned L = tb_crc_bits(tbs_bits);
  const unsigned K = (base_graph == 1 ? 22u : 10u) * Z;
  const unsigned cb_crc = C > 1 ? 24u : 0u;
  const unsigned bout = tbs_bits + L + (C > 1 ? 24 * C : 0);
  const unsigned info = (bout + C - 1) / C - cb_crc; // information bits per codeblock
  if (info + cb_crc > K) {
    return LH_ERR_INVALID;
  }
  seg->nof_cbs = C;
  seg->lifting_size = Z;
  seg->segment_length = K;
  seg->nof_filler_bits = K - info - cb_crc;
  seg->nof_tb_crc_bits = L;
  seg->nof_cb_crc_bits = C > 1 ? 24 : L;
  seg->cb_data_bits = C > 1 ? info : tbs_bits;
  seg->zero_pad = (info + cb_crc) * C - bout;
  const unsigned per_layer = nof_ch_symbols / nof_layers;
  const unsigned n_short = C - per_layer % C;
  unsigned off = 0;
  for (unsigned c = 0; c < C; ++c) {
    const unsigned t = c < n_short ? per_layer / C : (per_layer + C - 1) / C;
    rm_lengths[c] = t * nof_layers * modulation_order;
    cw_offsets[c] = off;
    off += rm_lengths[c];
  }
  return LH_OK;
}
