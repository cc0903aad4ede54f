// Adapters from the reference's class interfaces to the C ABI (include/ocudu_hip_ldpc.h).
// Per-call semantics follow the reference: spans are host memory valid for the call,
// results are host-visible on return, a decoder instance is not thread-safe (the shared
// device context serialises callers with a mutex, like the CUDA batch decoder's
// decode_mutex_).
#include "hip_ldpc_adapters.h"

#include "ocudu_hip_ldpc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

namespace ocudu {
namespace {

void check_hip(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

void check_lh(int rc, const char* what)
{
  if (rc != LH_OK) {
    throw std::runtime_error(std::string(what) + ": " + lh_last_error());
  }
}

// Device buffer that only grows.
struct dev_buf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t n)
  {
    if (n <= bytes) {
      return;
    }
    if (ptr) {
      (void)hipFree(ptr);
    }
    check_hip(hipMalloc(&ptr, n), "hipMalloc");
    bytes = n;
  }
  ~dev_buf()
  {
    if (ptr) {
      (void)hipFree(ptr);
    }
  }
};

// One HIP context + stream shared by the products of a factory.
class hip_device
{
public:
  explicit hip_device(const hip_backend_config& cfg) : device(cfg.device), max_codeblocks(cfg.max_codeblocks)
  {
    check_hip(hipSetDevice(device), "hipSetDevice");
    check_lh(lh_ctx_create(&ctx, device, cfg.max_codeblocks), "lh_ctx_create");
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    for (unsigned bg = 1; bg <= 2; ++bg) {
      const hip_learned_table& t = cfg.learned[bg - 1];
      if (t.iters > 0) {
        check_lh(lh_set_scale_table(ctx, bg, t.alpha_q.data(), t.beta_q.data(), t.iters, bg == 1 ? 46 : 42),
                 "lh_set_scale_table");
        learned[bg - 1] = true;
      }
    }
  }
  ~hip_device()
  {
    (void)hipStreamDestroy(stream);
    lh_ctx_destroy(ctx);
  }
  std::mutex mutex;
  int device;
  unsigned max_codeblocks;
  bool learned[2] = {false, false};
  lh_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  dev_buf llr, rm, bits, res, tb, tb_status;

  void h2d(void* dst, const void* src, size_t n) { check_hip(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream), "H2D"); }
  void d2h(void* dst, const void* src, size_t n) { check_hip(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, stream), "D2H"); }
  void sync() { check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
};

unsigned bg_k(ldpc_base_graph_type bg)
{
  return bg == ldpc_base_graph_type::BG1 ? 22 : 10;
}

class ldpc_decoder_hip : public ldpc_decoder
{
public:
  ldpc_decoder_hip(std::shared_ptr<hip_device> d, ldpc_decoder_factory::ldpc_decoder_factory_configuration c) :
    dev(std::move(d)), flags((c.force_decoding ? LH_FORCE_DECODING : 0u) | (c.early_stop_syndrome ? LH_EARLY_STOP_SYNDROME : 0u))
  {
  }

  std::optional<unsigned>
  decode(bit_buffer& output, span<const log_likelihood_ratio> input, crc_calculator* crc, const configuration& cfg) override
  {
    const unsigned K = bg_k(cfg.base_graph) * static_cast<unsigned>(cfg.lifting_size);
    if (output.size() != K) {
      throw std::invalid_argument("output size differs from the message length");
    }
    std::lock_guard<std::mutex> lock(dev->mutex);
    (void)hipSetDevice(dev->device);
    const size_t nbytes = (K + 7) / 8;
    dev->llr.reserve(input.size() + 16);
    dev->bits.reserve(nbytes + 16);
    dev->res.reserve(16);
    dev->h2d(dev->llr.ptr, input.data(), input.size());
    // The decoder leaves the output untouched when it does not decode (CRC given).
    dev->h2d(dev->bits.ptr, output.get_buffer().data(), nbytes);
    lh_decode_desc d{};
    d.llr_offset = 0;
    d.bits_offset = 0;
    d.nof_llrs = static_cast<uint32_t>(input.size());
    d.crc_poly = crc ? static_cast<uint32_t>(crc->get_generator_poly()) : 0u;
    d.cfg = {static_cast<uint32_t>(cfg.base_graph), static_cast<uint32_t>(cfg.lifting_size), cfg.nof_filler_bits,
             cfg.nof_crc_bits, cfg.max_iterations};
    check_lh(lh_ldpc_decode(dev->ctx, &d, 1, static_cast<const int8_t*>(dev->llr.ptr), static_cast<uint8_t*>(dev->bits.ptr),
                            static_cast<int32_t*>(dev->res.ptr), flags, dev->stream),
             "lh_ldpc_decode");
    int32_t r = 0;
    dev->d2h(output.get_buffer().data(), dev->bits.ptr, nbytes);
    dev->d2h(&r, dev->res.ptr, 4);
    dev->sync();
    if (r < 0) {
      throw std::runtime_error("decoder layer bound violated");
    }
    return r > 0 ? std::optional<unsigned>(static_cast<unsigned>(r)) : std::nullopt;
  }

private:
  std::shared_ptr<hip_device> dev;
  uint32_t flags;
};

class ldpc_rate_dematcher_hip : public ldpc_rate_dematcher
{
public:
  explicit ldpc_rate_dematcher_hip(std::shared_ptr<hip_device> d) : dev(std::move(d)) {}

  void rate_dematch(span<log_likelihood_ratio>       output,
                    span<const log_likelihood_ratio> input,
                    bool                             new_data,
                    const codeblock_metadata&        cfg) override
  {
    // The base graph follows from the buffer length, as in the reference
    // (ldpc_rate_dematcher_impl.cpp:62-80).
    const unsigned N = static_cast<unsigned>(output.size());
    const unsigned bg = (N % 66 == 0) ? 1 : 2;
    const unsigned Z = N / (bg == 1 ? 66 : 50);
    std::lock_guard<std::mutex> lock(dev->mutex);
    (void)hipSetDevice(dev->device);
    dev->llr.reserve(input.size() + 16);
    dev->rm.reserve(N + 16);
    dev->h2d(dev->llr.ptr, input.data(), input.size());
    if (!new_data) {
      dev->h2d(dev->rm.ptr, output.data(), N);
    }
    lh_rm_desc d{};
    d.cw_offset = 0;
    d.rm_offset = 0;
    d.rm_length = static_cast<uint32_t>(input.size());
    d.base_graph = bg;
    d.lifting_size = Z;
    d.rv = cfg.tb_common.rv;
    d.modulation_order = get_bits_per_symbol(cfg.tb_common.mod);
    d.Nref = cfg.tb_common.Nref;
    d.nof_filler_bits = cfg.cb_specific.nof_filler_bits;
    check_lh(lh_rate_dematch(dev->ctx, &d, 1, static_cast<const int8_t*>(dev->llr.ptr), static_cast<int8_t*>(dev->rm.ptr),
                             new_data ? 1 : 0, dev->stream),
             "lh_rate_dematch");
    dev->d2h(output.data(), dev->rm.ptr, N);
    dev->sync();
  }

private:
  std::shared_ptr<hip_device> dev;
};

// Event-timed sections of one call (hipEvents on the call's stream), read back only when
// timing is enabled.
class call_timer
{
public:
  call_timer()
  {
    for (auto& e : ev) {
      check_hip(hipEventCreate(&e), "hipEventCreate");
    }
  }
  ~call_timer()
  {
    for (auto& e : ev) {
      (void)hipEventDestroy(e);
    }
  }
  void mark(int i, hipStream_t s, bool on)
  {
    if (on) {
      (void)hipEventRecord(ev[i], s);
    }
  }
  float between(int a, int b) const
  {
    float ms = 0.0f;
    return hipEventElapsedTime(&ms, ev[a], ev[b]) == hipSuccess ? 1000.0f * ms : 0.0f;
  }

private:
  hipEvent_t ev[4] = {};
};

float us_since(std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<float, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

class pusch_batch_decoder_hip : public pusch_resident_codeblock_decoder
{
public:
  explicit pusch_batch_decoder_hip(std::shared_ptr<hip_device> d) : dev(std::move(d)), dematcher(dev) {}

  std::vector<codeblock_result> decode_batch(span<const described_rx_codeblock> cbs,
                                             span<bit_buffer>                   cb_data,
                                             span<span<log_likelihood_ratio>>   rm_buffers,
                                             bool                               new_data,
                                             crc_generator_poly                 crc_poly,
                                             bool                               use_early_stop,
                                             unsigned                           nof_ldpc_iterations) override
  {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = cbs.size();
    std::vector<codeblock_result> out(n);
    resident.valid = false;
    // Over-capacity or inconsistent batches: every codeblock fails, nothing is touched
    // (pusch_codeblock_decoder_cuda_batch.cpp:915-926).
    if (n == 0 || n > dev->max_codeblocks || cb_data.size() != n || rm_buffers.size() != n) {
      return out;
    }
    std::vector<lh_pusch_cb_desc> descs(n);
    size_t cw_total = 0, rm_total = 0, bits_total = 0;
    for (size_t i = 0; i != n; ++i) {
      const codeblock_metadata& md = cbs[i].second;
      if (cbs[i].first.size() != md.cb_specific.rm_length || cbs[i].first.empty()) {
        return out;
      }
      lh_pusch_cb_desc& d = descs[i];
      d.rm.cw_offset = cw_total;
      d.rm.rm_offset = rm_total;
      d.rm.rm_length = static_cast<uint32_t>(cbs[i].first.size());
      d.rm.base_graph = static_cast<uint32_t>(md.tb_common.base_graph);
      d.rm.lifting_size = static_cast<uint32_t>(md.tb_common.lifting_size);
      d.rm.rv = md.tb_common.rv;
      d.rm.modulation_order = get_bits_per_symbol(md.tb_common.mod);
      d.rm.Nref = md.tb_common.Nref;
      d.rm.nof_filler_bits = md.cb_specific.nof_filler_bits;
      d.bits_offset = bits_total;
      d.nof_crc_bits = md.cb_specific.nof_crc_bits;
      d.crc_poly = static_cast<uint32_t>(crc_poly);
      cw_total += cbs[i].first.size();
      rm_total += rm_buffers[i].size();
      bits_total += (cb_data[i].size() + 7) / 8;
    }
    std::lock_guard<std::mutex> lock(dev->mutex);
    (void)hipSetDevice(dev->device);
    dev->llr.reserve(cw_total + 16);
    dev->rm.reserve(rm_total + 16);
    dev->bits.reserve(bits_total + 16);
    dev->res.reserve(4 * n + 16);
    hipStream_t st = dev->stream;
    timer.mark(0, st, timing);
    for (size_t i = 0; i != n; ++i) {
      dev->h2d(static_cast<int8_t*>(dev->llr.ptr) + descs[i].rm.cw_offset, cbs[i].first.data(), cbs[i].first.size());
      if (!new_data) {
        dev->h2d(static_cast<int8_t*>(dev->rm.ptr) + descs[i].rm.rm_offset, rm_buffers[i].data(), rm_buffers[i].size());
      }
      dev->h2d(static_cast<uint8_t*>(dev->bits.ptr) + descs[i].bits_offset, cb_data[i].get_buffer().data(),
               (cb_data[i].size() + 7) / 8);
    }
    timer.mark(1, st, timing);
    decode(descs.data(), static_cast<unsigned>(n), static_cast<const int8_t*>(dev->llr.ptr),
           static_cast<int8_t*>(dev->rm.ptr), new_data, use_early_stop, nof_ldpc_iterations, st);
    timer.mark(2, st, timing);
    std::vector<int32_t> res(n);
    dev->d2h(res.data(), dev->res.ptr, 4 * n);
    for (size_t i = 0; i != n; ++i) {
      // HARQ semantics: the combined soft bits go back to the caller's buffers.
      dev->d2h(rm_buffers[i].data(), static_cast<int8_t*>(dev->rm.ptr) + descs[i].rm.rm_offset, rm_buffers[i].size());
      dev->d2h(cb_data[i].get_buffer().data(), static_cast<uint8_t*>(dev->bits.ptr) + descs[i].bits_offset,
               (cb_data[i].size() + 7) / 8);
    }
    timer.mark(3, st, timing);
    dev->sync();
    for (size_t i = 0; i != n; ++i) {
      out[i].crc_ok = res[i] > 0;
      out[i].nof_iterations = res[i] > 0 ? static_cast<unsigned>(res[i]) : nof_ldpc_iterations;
    }
    if (timing) {
      begin_stats(static_cast<unsigned>(n));
      stats.h2d_transfer_us = timer.between(0, 1);
      stats.ldpc_decode_us = timer.between(1, 2); // dematching fused into the decoder launch
      stats.d2h_transfer_us = timer.between(2, 3);
      stats.total_backend_us = timer.between(0, 3);
      stats.total_e2e_us = us_since(t0);
    }
    return out;
  }

  void rate_dematch_only(span<const log_likelihood_ratio> cb_llrs,
                         span<log_likelihood_ratio>       rm_buffer,
                         bool                             new_data,
                         const codeblock_metadata&        metadata) override
  {
    dematcher.rate_dematch(rm_buffer, cb_llrs, new_data, metadata);
  }

  // The dematched soft bits of the last resident decode stay on the device until they are
  // exported (once) into the caller's host HARQ buffers, which start with the first
  // non-punctured soft bit as the device buffers do.
  bool export_resident_softbits(span<span<log_likelihood_ratio>> rm_buffers) override
  {
    const resident_state r = resident;
    resident.valid = false;
    if (!r.valid || rm_buffers.size() != r.nof_cbs) {
      return false;
    }
    std::lock_guard<std::mutex> lock(dev->mutex);
    (void)hipSetDevice(dev->device);
    for (unsigned c = 0; c != r.nof_cbs; ++c) {
      const size_t count = std::min<size_t>(rm_buffers[c].size(), r.cb_softbits);
      dev->d2h(rm_buffers[c].data(), static_cast<int8_t*>(resident_rm.ptr) + static_cast<size_t>(c) * r.cb_softbits,
               count);
    }
    if (hipStreamSynchronize(dev->stream) != hipSuccess) {
      return false;
    }
    for (unsigned c = 0; c != r.nof_cbs; ++c) {
      // Positions the device never held start clean.
      for (size_t i = r.cb_softbits; i < rm_buffers[c].size(); ++i) {
        rm_buffers[c][i] = log_likelihood_ratio(0);
      }
    }
    return true;
  }

  // resident_llrs: the codeword's int8 LLRs in device memory, already descrambled (as
  // lh_pusch_demodulate writes them); the codeblocks' LLRs follow each other in order.
  // Dematching, decoding, TB reassembly and TB CRC run on the device; only the verdicts,
  // the codeblock results and the TB bytes come back.  Invalid or over-capacity grants are
  // rejected before any copy or launch, leaving tb_output untouched.
  tb_decode_result decode_resident_softbits(void*                                                resident_llrs,
                                            size_t                                               nof_softbits,
                                            const scrambling_config*                             scrambling_cfg,
                                            const codeblock_metadata::tb_common_metadata&        tb_common,
                                            span<const codeblock_metadata::cb_specific_metadata> cb_specific,
                                            span<uint8_t>                                        tb_output,
                                            crc_generator_poly                                   cb_crc_poly,
                                            unsigned                                             nof_ldpc_iterations,
                                            void*                                                execution_context,
                                            int /*buffer_index*/) override
  {
    const auto t0 = std::chrono::steady_clock::now();
    resident.valid = false;
    const size_t n = cb_specific.size();
    const unsigned bg = static_cast<unsigned>(tb_common.base_graph);
    const unsigned Z = static_cast<unsigned>(tb_common.lifting_size);
    const unsigned K = (bg == 1 ? 22u : 10u) * Z;
    const unsigned N = (bg == 1 ? 66u : 50u) * Z;
    const unsigned mb = (K + 7) / 8;
    const size_t tbs_bits = tb_output.size() * 8;
    const unsigned tb_crc = tbs_bits > 3824 ? 24u : 16u;
    bool ok = resident_llrs != nullptr && scrambling_cfg == nullptr && n > 0 && n <= dev->max_codeblocks &&
              tbs_bits > 0 && nof_ldpc_iterations > 0;
    size_t described = 0;
    for (size_t i = 0; ok && i != n; ++i) {
      ok = cb_specific[i].rm_length > 0 && cb_specific[i].nof_filler_bits == cb_specific[0].nof_filler_bits;
      described += cb_specific[i].rm_length;
    }
    const unsigned F = ok ? cb_specific[0].nof_filler_bits : 0;
    const unsigned data_bits = K - F - (n > 1 ? 24u : 0u);
    ok = ok && described == nof_softbits && F < K && tbs_bits + tb_crc <= n * static_cast<size_t>(data_bits);
    if (!ok) {
      return {};
    }
    std::vector<lh_pusch_cb_desc> descs(n);
    size_t cw = 0;
    for (size_t i = 0; i != n; ++i) {
      lh_pusch_cb_desc& d = descs[i];
      d.rm.cw_offset = cw;
      d.rm.rm_offset = i * static_cast<size_t>(N);
      d.rm.rm_length = cb_specific[i].rm_length;
      d.rm.base_graph = bg;
      d.rm.lifting_size = Z;
      d.rm.rv = tb_common.rv;
      d.rm.modulation_order = get_bits_per_symbol(tb_common.mod);
      d.rm.Nref = tb_common.Nref;
      d.rm.nof_filler_bits = F;
      d.bits_offset = i * static_cast<size_t>(mb);
      d.nof_crc_bits = cb_specific[i].nof_crc_bits;
      d.crc_poly = static_cast<uint32_t>(cb_crc_poly);
      cw += cb_specific[i].rm_length;
    }
    lh_tb_desc tb{};
    tb.tb_offset = 0;
    tb.cb_bits_offset = 0;
    tb.cb_bits_stride = mb;
    tb.tbs_bits = static_cast<uint32_t>(tbs_bits);
    tb.first_cb = 0;
    tb.nof_cbs = static_cast<uint32_t>(n);
    tb.cb_data_bits = data_bits;

    std::lock_guard<std::mutex> lock(dev->mutex);
    (void)hipSetDevice(dev->device);
    resident_rm.reserve(n * static_cast<size_t>(N) + 16);
    dev->bits.reserve(n * static_cast<size_t>(mb) + 16);
    dev->res.reserve(4 * n + 16);
    dev->tb.reserve(tb_output.size() + 16);
    dev->tb_status.reserve(16);
    hipStream_t st = execution_context != nullptr ? static_cast<hipStream_t>(execution_context) : dev->stream;
    timer.mark(0, st, timing);
    decode(descs.data(), static_cast<unsigned>(n), static_cast<const int8_t*>(resident_llrs),
           static_cast<int8_t*>(resident_rm.ptr), true, true, nof_ldpc_iterations, st);
    timer.mark(1, st, timing);
    check_lh(lh_tb_finalize(dev->ctx, &tb, 1, static_cast<const uint8_t*>(dev->bits.ptr),
                            static_cast<const int32_t*>(dev->res.ptr), static_cast<uint8_t*>(dev->tb.ptr),
                            static_cast<int32_t*>(dev->tb_status.ptr), st),
             "lh_tb_finalize");
    timer.mark(2, st, timing);
    int32_t status = 0;
    std::vector<int32_t> res(n);
    check_hip(hipMemcpyAsync(&status, dev->tb_status.ptr, 4, hipMemcpyDeviceToHost, st), "D2H");
    check_hip(hipMemcpyAsync(res.data(), dev->res.ptr, 4 * n, hipMemcpyDeviceToHost, st), "D2H");
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    const auto t_copy = std::chrono::steady_clock::now();
    if (status == LH_TB_CRC_OK) {
      check_hip(hipMemcpy(tb_output.data(), dev->tb.ptr, tb_output.size(), hipMemcpyDeviceToHost), "D2H");
    }
    timer.mark(3, st, timing);
    tb_decode_result out;
    out.tb_crc_ok = status == LH_TB_CRC_OK;
    unsigned passed = 0, iters = 0;
    for (int32_t r : res) {
      if (r > 0) {
        ++passed;
        iters += static_cast<unsigned>(r);
      }
    }
    out.all_cb_crcs_ok = passed == n;
    // Representative count: the mean over the codeblocks that passed, rounded.
    out.nof_iterations = passed > 0 ? (iters + passed / 2) / passed : nof_ldpc_iterations;
    resident = {true, static_cast<unsigned>(n), N};
    if (timing) {
      begin_stats(static_cast<unsigned>(n));
      stats.ldpc_decode_us = timer.between(0, 1); // dematching fused into the decoder launch
      stats.crc_check_us = timer.between(1, 2);   // TB reassembly + TB CRC
      stats.tb_output_copy_us = us_since(t_copy);
      stats.total_backend_us = timer.between(0, 2);
      stats.total_e2e_us = us_since(t0);
    }
    return out;
  }

  const pusch_resident_codeblock_decoder_timing_stats& get_last_timing_stats() const override { return stats; }

  void set_demod_timing(float grid_staging_us, float demod_sync_us) override
  {
    grid_staging = grid_staging_us;
    demod_sync = demod_sync_us;
    stats.grid_staging_us = grid_staging_us;
    stats.demod_sync_us = demod_sync_us;
  }

  void set_timing_enabled(bool enable) override { timing = enable; }

  // "learned": the learned min-sum decoder with the plain decoder as its fallback (for the
  // base graphs whose tables the backend was given); anything else ("auto", "min_sum",
  // "generic", ...): the plain decoder, bit-exact with the reference's generic decoder.
  void set_ldpc_decoder_algorithm(const std::string& algorithm) override { learned = algorithm == "learned"; }

private:
  struct resident_state {
    bool valid = false;
    unsigned nof_cbs = 0;
    unsigned cb_softbits = 0;
  };

  void begin_stats(unsigned n)
  {
    stats = pusch_resident_codeblock_decoder_timing_stats{};
    stats.nof_cbs = n;
    stats.timing_enabled = true;
    stats.grid_staging_us = grid_staging;
    stats.demod_sync_us = demod_sync;
  }

  // One decode launch of a batch into the device HARQ buffers, bits and results.
  void decode(const lh_pusch_cb_desc* descs, unsigned n, const int8_t* d_cw, int8_t* rm, bool new_data,
              bool use_early_stop, unsigned iters, hipStream_t st)
  {
    bool use_learned = learned;
    for (unsigned i = 0; use_learned && i != n; ++i) {
      use_learned = dev->learned[descs[i].rm.base_graph - 1];
    }
    uint8_t* bits = static_cast<uint8_t*>(dev->bits.ptr);
    int32_t* res = static_cast<int32_t*>(dev->res.ptr);
    if (use_learned) {
      check_lh(lh_pusch_decode_batch_learned(dev->ctx, descs, n, d_cw, rm, bits, res, new_data ? 1 : 0,
                                             use_early_stop ? 1 : 0, iters, iters, st),
               "lh_pusch_decode_batch_learned");
    } else {
      check_lh(lh_pusch_decode_batch(dev->ctx, descs, n, d_cw, rm, bits, res, new_data ? 1 : 0, use_early_stop ? 1 : 0,
                                     iters, st),
               "lh_pusch_decode_batch");
    }
  }

  std::shared_ptr<hip_device> dev;
  ldpc_rate_dematcher_hip dematcher;
  dev_buf resident_rm; // soft bits of the last resident decode (kept for export)
  call_timer timer;
  bool timing = false;
  bool learned = false;
  float grid_staging = 0.0f, demod_sync = 0.0f;
  resident_state resident;
  pusch_resident_codeblock_decoder_timing_stats stats;
};

class ldpc_decoder_factory_hip : public ldpc_decoder_factory
{
public:
  ldpc_decoder_factory_hip(const hip_backend_config& h, ldpc_decoder_factory_configuration c) :
    dev(std::make_shared<hip_device>(h)), cfg(c)
  {
  }
  std::unique_ptr<ldpc_decoder> create() override { return std::make_unique<ldpc_decoder_hip>(dev, cfg); }

private:
  std::shared_ptr<hip_device> dev;
  ldpc_decoder_factory_configuration cfg;
};

class ldpc_rate_dematcher_factory_hip : public ldpc_rate_dematcher_factory
{
public:
  explicit ldpc_rate_dematcher_factory_hip(const hip_backend_config& h) : dev(std::make_shared<hip_device>(h)) {}
  std::unique_ptr<ldpc_rate_dematcher> create() override { return std::make_unique<ldpc_rate_dematcher_hip>(dev); }

private:
  std::shared_ptr<hip_device> dev;
};

} // namespace

std::shared_ptr<ldpc_decoder_factory>
create_ldpc_decoder_factory_hip(const hip_backend_config& hcfg, ldpc_decoder_factory::ldpc_decoder_factory_configuration cfg)
{
  return std::make_shared<ldpc_decoder_factory_hip>(hcfg, cfg);
}

std::shared_ptr<ldpc_rate_dematcher_factory> create_ldpc_rate_dematcher_factory_hip(const hip_backend_config& hcfg)
{
  return std::make_shared<ldpc_rate_dematcher_factory_hip>(hcfg);
}

std::unique_ptr<pusch_resident_codeblock_decoder> create_pusch_batch_decoder_hip(const hip_backend_config& hcfg)
{
  return std::make_unique<pusch_batch_decoder_hip>(std::make_shared<hip_device>(hcfg));
}

const char* get_ldpc_decoder_type_str_hip()
{
  return "hip";
}

} // namespace ocudu
