// The reference interfaces that the HIP backend plugs into.  Inside the reference tree
// (OCUDU_HIP_REFERENCE_HEADERS defined, the reference's include/ on the include path) they
// are the reference's own headers:
//   span / bit_buffer          include/ocudu/adt/span.h, include/ocudu/adt/bit_buffer.h
//   log_likelihood_ratio       include/ocudu/phy/upper/log_likelihood_ratio.h
//   crc_calculator             include/ocudu/phy/upper/channel_coding/crc_calculator.h:17-81
//   codeblock_metadata         include/ocudu/phy/upper/codeblock_metadata.h:24-144
//   ldpc_decoder               include/ocudu/phy/upper/channel_coding/ldpc/ldpc_decoder.h:19-59
//   ldpc_rate_dematcher        include/ocudu/phy/upper/channel_coding/ldpc/ldpc_rate_dematcher.h:17-40
//   *_factory                  include/ocudu/phy/upper/channel_coding/channel_coding_factories.h:26-100
//   pusch_resident_codeblock_decoder
//                              include/ocudu/phy/upper/channel_processors/pusch/pusch_resident_codeblock_decoder.h:20-149
// Without it, the restatement below (same names, signatures and semantics) lets the
// adapters and their chain test build stand-alone.
#pragma once

#ifdef OCUDU_HIP_REFERENCE_HEADERS

#include "ocudu/phy/upper/channel_coding/channel_coding_factories.h"
#include "ocudu/phy/upper/channel_coding/crc_calculator.h"
#include "ocudu/phy/upper/channel_coding/ldpc/ldpc_decoder.h"
#include "ocudu/phy/upper/channel_coding/ldpc/ldpc_rate_dematcher.h"
#include "ocudu/phy/upper/channel_processors/pusch/pusch_resident_codeblock_decoder.h"

#else

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <string>
#include <utility>
#include <vector>

namespace ocudu {

template <typename T>
class span
{
public:
  span() = default;
  span(T* p, size_t n) : ptr(p), len(n) {}
  template <typename U>
  span(std::vector<U>& v) : ptr(v.data()), len(v.size())
  {
  }
  template <typename U>
  span(const std::vector<U>& v) : ptr(v.data()), len(v.size())
  {
  }
  template <typename U>
  span(const span<U>& o) : ptr(o.data()), len(o.size())
  {
  }
  T* data() const { return ptr; }
  size_t size() const { return len; }
  bool empty() const { return len == 0; }
  T& operator[](size_t i) const { return ptr[i]; }
  T* begin() const { return ptr; }
  T* end() const { return ptr + len; }
  span first(size_t n) const { return span(ptr, n); }
  span subspan(size_t o, size_t n) const { return span(ptr + o, n); }

private:
  T* ptr = nullptr;
  size_t len = 0;
};

class log_likelihood_ratio
{
public:
  using value_type = int8_t;
  constexpr log_likelihood_ratio() = default;
  constexpr log_likelihood_ratio(int v) : value(static_cast<value_type>(v)) {}
  constexpr value_type to_value_type() const { return value; }
  constexpr int to_int() const { return value; }
  constexpr bool operator==(log_likelihood_ratio o) const { return value == o.value; }
  constexpr bool operator!=(log_likelihood_ratio o) const { return value != o.value; }

private:
  value_type value = 0;
};
static_assert(sizeof(log_likelihood_ratio) == 1, "LLRs are bytes");

// Packed bit view, MSB-first within each byte.
class bit_buffer
{
public:
  bit_buffer() = default;
  bit_buffer(uint8_t* d, size_t nbits) : bytes(d), nbits_(nbits) {}
  size_t size() const { return nbits_; }
  span<uint8_t> get_buffer() const { return span<uint8_t>(bytes, (nbits_ + 7) / 8); }
  uint8_t extract(size_t i) const { return (bytes[i / 8] >> (7 - i % 8)) & 1u; }
  void insert(uint8_t b, size_t i)
  {
    const uint8_t m = static_cast<uint8_t>(0x80u >> (i % 8));
    bytes[i / 8] = b ? (bytes[i / 8] | m) : (bytes[i / 8] & static_cast<uint8_t>(~m));
  }
  bit_buffer first(size_t n) const { return bit_buffer(bytes, n); }

private:
  uint8_t* bytes = nullptr;
  size_t nbits_ = 0;
};

enum class crc_generator_poly : uint32_t {
  CRC24A = 0x1864cfb,
  CRC24B = 0x1800063,
  CRC24C = 0x1b2b117,
  CRC16 = 0x11021,
  CRC11 = 0xe21,
  CRC6 = 0x61,
};

using crc_calculator_checksum_t = unsigned;

class crc_calculator
{
public:
  virtual ~crc_calculator() = default;
  // 8 packed bits per byte, MSB first.
  virtual crc_calculator_checksum_t calculate_byte(span<const uint8_t> data) const = 0;
  // One bit per byte.
  virtual crc_calculator_checksum_t calculate_bit(span<const uint8_t> data) const = 0;
  virtual crc_calculator_checksum_t calculate(const bit_buffer& data) const = 0;
  virtual crc_generator_poly get_generator_poly() const = 0;
};

enum class ldpc_base_graph_type { BG1 = 1, BG2 = 2 };

namespace ldpc {
using lifting_size_t = unsigned;
}

enum class modulation_scheme { PI_2_BPSK = 0, BPSK = 1, QPSK = 2, QAM16 = 4, QAM64 = 6, QAM256 = 8 };

inline unsigned get_bits_per_symbol(modulation_scheme m)
{
  return m == modulation_scheme::PI_2_BPSK ? 1u : static_cast<unsigned>(m);
}

struct codeblock_metadata {
  struct tb_common_metadata {
    ldpc_base_graph_type base_graph = ldpc_base_graph_type::BG1;
    ldpc::lifting_size_t lifting_size = 2;
    unsigned rv = 0;
    modulation_scheme mod = modulation_scheme::BPSK;
    unsigned Nref = 0;
    unsigned cw_length = 0;
  };
  struct cb_specific_metadata {
    unsigned full_length = 0;
    unsigned rm_length = 0;
    unsigned nof_filler_bits = 0;
    unsigned cw_offset = 0;
    unsigned nof_crc_bits = 16;
  };
  tb_common_metadata tb_common;
  cb_specific_metadata cb_specific;
};

using described_rx_codeblock = std::pair<span<const log_likelihood_ratio>, codeblock_metadata>;

class ldpc_decoder
{
public:
  virtual ~ldpc_decoder() = default;
  struct configuration {
    ldpc_base_graph_type base_graph = ldpc_base_graph_type::BG1;
    ldpc::lifting_size_t lifting_size = 2;
    unsigned nof_filler_bits = 0;
    unsigned nof_crc_bits = 16;
    unsigned max_iterations = 6;
  };
  virtual std::optional<unsigned>
  decode(bit_buffer& output, span<const log_likelihood_ratio> input, crc_calculator* crc, const configuration& cfg) = 0;
};

class ldpc_rate_dematcher
{
public:
  virtual ~ldpc_rate_dematcher() = default;
  virtual void rate_dematch(span<log_likelihood_ratio>       output,
                            span<const log_likelihood_ratio> input,
                            bool                             new_data,
                            const codeblock_metadata&        cfg) = 0;
};

class ldpc_decoder_factory
{
public:
  struct ldpc_decoder_factory_configuration {
    bool force_decoding;
    bool early_stop_syndrome;
  };
  virtual ~ldpc_decoder_factory() = default;
  virtual std::unique_ptr<ldpc_decoder> create() = 0;
};

class ldpc_rate_dematcher_factory
{
public:
  virtual ~ldpc_rate_dematcher_factory() = default;
  virtual std::unique_ptr<ldpc_rate_dematcher> create() = 0;
};

struct pusch_resident_codeblock_decoder_timing_stats {
  float h2d_transfer_us = 0.0f;
  float deinterleave_us = 0.0f;
  float rate_dematch_us = 0.0f;
  float ldpc_decode_us = 0.0f;
  float crc_check_us = 0.0f;
  float d2h_transfer_us = 0.0f;
  float decoder_config_us = 0.0f;
  float rate_match_config_us = 0.0f;
  float completion_wait_us = 0.0f;
  float tb_output_copy_us = 0.0f;
  float avg_iters_query_us = 0.0f;
  float cpu_crc_us = 0.0f;
  float total_backend_us = 0.0f;
  float total_e2e_us = 0.0f;
  unsigned nof_cbs = 0;
  bool timing_enabled = false;
  float extract_bits_us = 0.0f;
  float grid_staging_us = 0.0f;
  float demod_sync_us = 0.0f;
};

class pusch_resident_codeblock_decoder
{
public:
  struct codeblock_result {
    bool crc_ok = false;
    unsigned nof_iterations = 0;
  };
  struct scrambling_config {
    uint16_t n_rnti = 0;
    uint16_t n_id = 0;
    uint8_t q = 0;
    uint8_t n_s = 0;
    unsigned bit_offset = 0;
  };
  struct tb_decode_result {
    bool tb_crc_ok = false;
    bool all_cb_crcs_ok = false;
    unsigned nof_iterations = 0;
  };
  virtual ~pusch_resident_codeblock_decoder() = default;
  virtual std::vector<codeblock_result> decode_batch(span<const described_rx_codeblock> codeblock_llrs,
                                                     span<bit_buffer>                   cb_data_buffers,
                                                     span<span<log_likelihood_ratio>>   rm_buffers,
                                                     bool                               new_data,
                                                     crc_generator_poly                 crc_poly,
                                                     bool                               use_early_stop,
                                                     unsigned                           nof_ldpc_iterations) = 0;
  virtual void rate_dematch_only(span<const log_likelihood_ratio> cb_llrs,
                                 span<log_likelihood_ratio>       rm_buffer,
                                 bool                             new_data,
                                 const codeblock_metadata&        metadata) = 0;
  virtual bool export_resident_softbits(span<span<log_likelihood_ratio>> rm_buffers) = 0;
  virtual tb_decode_result decode_resident_softbits(void*                                                resident_llrs,
                                                    size_t                                               nof_softbits,
                                                    const scrambling_config*                             scrambling_cfg,
                                                    const codeblock_metadata::tb_common_metadata&        tb_common,
                                                    span<const codeblock_metadata::cb_specific_metadata> cb_specific,
                                                    span<uint8_t>                                        tb_output,
                                                    crc_generator_poly                                   cb_crc_poly,
                                                    unsigned nof_ldpc_iterations,
                                                    void*    execution_context,
                                                    int      buffer_index = -1) = 0;
  virtual const pusch_resident_codeblock_decoder_timing_stats& get_last_timing_stats() const = 0;
  virtual void set_demod_timing(float grid_staging_us, float demod_sync_us) = 0;
  virtual void set_timing_enabled(bool enable) = 0;
  virtual void set_ldpc_decoder_algorithm(const std::string& algorithm) = 0;
};

} // namespace ocudu

#endif // OCUDU_HIP_REFERENCE_HEADERS
