// Adapters from the reference's class interfaces to the C ABI (include/ocudu_hip_ldpc.h).
// Per-call semantics follow the reference: spans are host memory valid for the call,
// results are host-visible on return, a decoder instance is not thread-safe (the shared
// device context serialises callers with a mutex, like the CUDA batch decoder's
// decode_mutex_).
#include "hip_ldpc_adapters.h"

#include "ocudu_hip_ldpc.h"

#include <hip/hip_runtime.h>

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
  explicit hip_device(const hip_backend_config& cfg) : device(cfg.device)
  {
    check_hip(hipSetDevice(device), "hipSetDevice");
    check_lh(lh_ctx_create(&ctx, device, cfg.max_codeblocks), "lh_ctx_create");
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  ~hip_device()
  {
    (void)hipStreamDestroy(stream);
    lh_ctx_destroy(ctx);
  }
  std::mutex mutex;
  int device;
  lh_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  dev_buf llr, rm, bits, res;

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
    const unsigned K = bg_k(cfg.base_graph) * cfg.lifting_size;
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
    d.cfg = {static_cast<uint32_t>(cfg.base_graph), cfg.lifting_size, cfg.nof_filler_bits, cfg.nof_crc_bits,
             cfg.max_iterations};
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
    const size_t n = cbs.size();
    std::vector<codeblock_result> out(n);
    if (n == 0) {
      return out;
    }
    std::vector<lh_pusch_cb_desc> descs(n);
    size_t cw_total = 0, rm_total = 0, bits_total = 0;
    for (size_t i = 0; i != n; ++i) {
      const codeblock_metadata& md = cbs[i].second;
      lh_pusch_cb_desc& d = descs[i];
      d.rm.cw_offset = cw_total;
      d.rm.rm_offset = rm_total;
      d.rm.rm_length = static_cast<uint32_t>(cbs[i].first.size());
      d.rm.base_graph = static_cast<uint32_t>(md.tb_common.base_graph);
      d.rm.lifting_size = md.tb_common.lifting_size;
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
    for (size_t i = 0; i != n; ++i) {
      dev->h2d(static_cast<int8_t*>(dev->llr.ptr) + descs[i].rm.cw_offset, cbs[i].first.data(), cbs[i].first.size());
      if (!new_data) {
        dev->h2d(static_cast<int8_t*>(dev->rm.ptr) + descs[i].rm.rm_offset, rm_buffers[i].data(), rm_buffers[i].size());
      }
      dev->h2d(static_cast<uint8_t*>(dev->bits.ptr) + descs[i].bits_offset, cb_data[i].get_buffer().data(),
               (cb_data[i].size() + 7) / 8);
    }
    check_lh(lh_pusch_decode_batch(dev->ctx, descs.data(), static_cast<unsigned>(n), static_cast<const int8_t*>(dev->llr.ptr),
                                   static_cast<int8_t*>(dev->rm.ptr), static_cast<uint8_t*>(dev->bits.ptr),
                                   static_cast<int32_t*>(dev->res.ptr), new_data ? 1 : 0, use_early_stop ? 1 : 0,
                                   nof_ldpc_iterations, dev->stream),
             "lh_pusch_decode_batch");
    std::vector<int32_t> res(n);
    dev->d2h(res.data(), dev->res.ptr, 4 * n);
    for (size_t i = 0; i != n; ++i) {
      // HARQ semantics: the combined soft bits go back to the caller's buffers.
      dev->d2h(rm_buffers[i].data(), static_cast<int8_t*>(dev->rm.ptr) + descs[i].rm.rm_offset, rm_buffers[i].size());
      dev->d2h(cb_data[i].get_buffer().data(), static_cast<uint8_t*>(dev->bits.ptr) + descs[i].bits_offset,
               (cb_data[i].size() + 7) / 8);
    }
    dev->sync();
    for (size_t i = 0; i != n; ++i) {
      out[i].crc_ok = res[i] > 0;
      out[i].nof_iterations = res[i] > 0 ? static_cast<unsigned>(res[i]) : nof_ldpc_iterations;
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

  // Soft bits are written back on every decode_batch, nothing stays resident.
  bool export_resident_softbits(span<span<log_likelihood_ratio>>) override { return false; }

private:
  std::shared_ptr<hip_device> dev;
  ldpc_rate_dematcher_hip dematcher;
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
