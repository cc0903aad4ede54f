// HIP backend for the reference's channel-coding plug-in points: an ldpc_decoder and an
// ldpc_rate_dematcher selectable as factory type "hip" (beside "generic"/"avx2"/"avx512"
// of create_ldpc_decoder_factory_sw, lib/phy/upper/channel_coding/channel_coding_factories.cpp:84-148,181-216),
// and a pusch_resident_codeblock_decoder batch backend (where the CUDA batch decoder plugs
// in, lib/phy/upper/channel_processors/pusch/processor_factories.cpp:188-291).
// All of them call the C ABI of include/ocudu_hip_ldpc.h.
#pragma once

#include "ocudu_mirror.h"

#include <cstdint>
#include <string>
#include <vector>

namespace ocudu {

// A learned min-sum normalisation table (lh_set_scale_table): iters x layers entries.
struct hip_learned_table {
  std::vector<uint16_t> alpha_q;
  std::vector<int16_t> beta_q;
  unsigned iters = 0;
};

struct hip_backend_config {
  int device = 0;
  // Capacity of one batch (pusch_decoder_capacity codeblocks).
  unsigned max_codeblocks = 512;
  // Tables of the "learned" decoder algorithm (BG1, BG2); empty: that base graph decodes
  // with the plain decoder.
  hip_learned_table learned[2];
};

std::shared_ptr<ldpc_decoder_factory>
create_ldpc_decoder_factory_hip(const hip_backend_config& hcfg,
                                ldpc_decoder_factory::ldpc_decoder_factory_configuration cfg = {false, false});

std::shared_ptr<ldpc_rate_dematcher_factory> create_ldpc_rate_dematcher_factory_hip(const hip_backend_config& hcfg);

std::unique_ptr<pusch_resident_codeblock_decoder> create_pusch_batch_decoder_hip(const hip_backend_config& hcfg);

// Implementation string reported by the factories (get_ldpc_decoder_type_str).
const char* get_ldpc_decoder_type_str_hip();

} // namespace ocudu
