// C ABI of the HIP LDPC receive chain (declared in include/ocudu_hip_ldpc.h).
//
// Host side: validates the reference-style descriptors, turns them into device jobs
// (layer-count bounds, k0, filler ranges), stages them through a small ring of pinned
// buffers and launches the kernels on the caller's stream.
#include "ocudu_hip_ldpc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lh_internal.h"

namespace lh {
size_t decode_lds_bytes(int bg, int maxZ, int lmax, int nt);
hipError_t launch_decode(int bg, const dec_job* d_jobs, unsigned n, int maxZ, int lmax, const int8_t* in,
                         uint8_t* bits, int32_t* results, const uint32_t* crc_x, hipStream_t s);
hipError_t launch_rate_dematch(const rm_job* d_jobs, unsigned n, unsigned maxN, const int8_t* cw, int8_t* rm,
                               hipStream_t s);
hipError_t launch_crc(const uint64_t* d_offs, const uint32_t* d_nbits, unsigned n, const uint8_t* bits,
                      const uint32_t* xtab, int order, uint32_t* out, hipStream_t s);
} // namespace lh

using namespace lh;

namespace {

thread_local std::string g_err = "no error";

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what)
{
  return fail(LH_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int lifting_index(unsigned Z)
{
  static const unsigned a_set[8] = {2, 3, 5, 7, 9, 11, 13, 15};
  for (int i = 0; i < 8; ++i) {
    for (unsigned v = a_set[i]; v <= 384; v *= 2) {
      if (v == Z) {
        return i;
      }
    }
  }
  return -1;
}

// x^e mod G for e in [0, kCrcXLen): the contribution of one message bit to the
// remainder of an MSB-first long division with zero init (crc_calculator_generic_impl.cpp:69-91).
void crc_x_table(uint32_t poly, uint32_t* out)
{
  unsigned order = 0;
  while ((poly >> (order + 1)) != 0) {
    ++order;
  }
  const uint32_t mask = (1u << order) - 1u;
  uint32_t v = 1;
  for (int e = 0; e < kCrcXLen; ++e) {
    out[e] = v & mask;
    v <<= 1;
    if (v & (1u << order)) {
      v ^= poly;
    }
  }
}

constexpr int kRing = 4;

struct slot {
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
  hipEvent_t done = nullptr;
  bool used = false;
};

} // namespace

struct lh_ctx {
  int device = 0;
  unsigned max_cb = 0;
  size_t slot_bytes = 0;
  uint32_t* d_crc_x = nullptr;
  slot ring[kRing];
  int next = 0;
};

namespace {

// Returns a staging slot whose previous upload has completed.
slot* acquire_slot(lh_ctx* ctx)
{
  slot* s = &ctx->ring[ctx->next];
  ctx->next = (ctx->next + 1) % kRing;
  if (s->used) {
    hipEventSynchronize(s->done);
  }
  return s;
}

int upload(lh_ctx* ctx, slot* s, size_t bytes, hipStream_t stream)
{
  hipError_t e = hipMemcpyAsync(s->dev, s->host, bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) {
    return hip_fail(e, "descriptor upload");
  }
  e = hipEventRecord(s->done, stream);
  if (e != hipSuccess) {
    return hip_fail(e, "event record");
  }
  s->used = true;
  (void)ctx;
  return LH_OK;
}

struct bg_info {
  int K, M, nshort;
};

bg_info bginfo(unsigned bg)
{
  return bg == 1 ? bg_info{22, 46, 66} : bg_info{10, 42, 50};
}

// Upper bound of the layer count from the largest position that can hold a nonzero
// soft bit (ldpc_decoder_impl.cpp:85-99).
uint32_t layer_bound(const bg_info& b, unsigned Z, uint64_t max_nonzero_pos_plus1)
{
  uint64_t cbl = std::max<uint64_t>(max_nonzero_pos_plus1 + 2ull * Z, (uint64_t)(b.K + 4) * Z);
  cbl = (cbl + Z - 1) / Z * Z;
  int64_t L = (int64_t)(cbl / Z) - b.K;
  return (uint32_t)std::min<int64_t>(std::max<int64_t>(L, 4), b.M);
}

int check_decoder_cfg(const lh_decoder_config& c, unsigned i)
{
  if (c.base_graph != 1 && c.base_graph != 2) {
    return fail(LH_ERR_INVALID, "cb %u: invalid base graph %u", i, c.base_graph);
  }
  if (lifting_index(c.lifting_size) < 0) {
    return fail(LH_ERR_INVALID, "cb %u: invalid lifting size %u", i, c.lifting_size);
  }
  if (c.nof_crc_bits != 16 && c.nof_crc_bits != 24) {
    return fail(LH_ERR_INVALID, "cb %u: invalid number of CRC bits %u", i, c.nof_crc_bits);
  }
  if (c.max_iterations == 0 || c.max_iterations > 255) {
    return fail(LH_ERR_INVALID, "cb %u: invalid max iterations %u", i, c.max_iterations);
  }
  bg_info b = bginfo(c.base_graph);
  if (c.nof_filler_bits >= (unsigned)(b.K - 2) * c.lifting_size) {
    return fail(LH_ERR_INVALID, "cb %u: invalid number of filler bits %u", i, c.nof_filler_bits);
  }
  return LH_OK;
}

// Fills an rm_job from a descriptor; returns the largest possibly-nonzero position + 1.
int make_rm_job(const lh_rm_desc& d, bool new_data, rm_job& j, uint64_t& max_pos1, unsigned i)
{
  static const double k0_bg1[4] = {0, 17, 33, 56};
  static const double k0_bg2[4] = {0, 13, 25, 43};
  if (d.base_graph != 1 && d.base_graph != 2) {
    return fail(LH_ERR_INVALID, "cb %u: invalid base graph %u", i, d.base_graph);
  }
  if (lifting_index(d.lifting_size) < 0) {
    return fail(LH_ERR_INVALID, "cb %u: invalid lifting size %u", i, d.lifting_size);
  }
  if (d.rv > 3) {
    return fail(LH_ERR_INVALID, "cb %u: RV should be an integer between 0 and 3", i);
  }
  const unsigned qm = d.modulation_order;
  if (qm != 1 && qm != 2 && qm != 4 && qm != 6 && qm != 8) {
    return fail(LH_ERR_INVALID, "cb %u: invalid modulation order %u", i, qm);
  }
  if (d.rm_length == 0 || d.rm_length % qm != 0 || d.rm_length > kMaxRmLength) {
    return fail(LH_ERR_INVALID, "cb %u: invalid rate-matched length %u", i, d.rm_length);
  }
  bg_info b = bginfo(d.base_graph);
  const unsigned Z = d.lifting_size;
  const unsigned N = (unsigned)b.nshort * Z;
  const unsigned nsys = (unsigned)(b.K - 2) * Z;
  if (d.nof_filler_bits >= nsys) {
    return fail(LH_ERR_INVALID, "cb %u: invalid number of filler bits", i);
  }
  const unsigned Ncb = (d.Nref > 0) ? std::min(d.Nref, N) : N;
  if (Ncb < nsys + 1) {
    return fail(LH_ERR_INVALID, "cb %u: N_ref %u shorter than the systematic part", i, d.Nref);
  }
  const double* f = d.base_graph == 1 ? k0_bg1 : k0_bg2;
  const unsigned k0 = (unsigned)std::floor((f[d.rv] * Ncb) / N) * Z;
  j = rm_job{};
  j.cw_off = d.cw_offset;
  j.rm_off = d.rm_offset;
  j.E = d.rm_length;
  j.N = N;
  j.Ncb = Ncb;
  j.k0 = k0;
  j.info = nsys - d.nof_filler_bits;
  j.nsys = nsys;
  j.Qm = (uint16_t)qm;
  j.flags = new_data ? kFlagNewData : 0;
  // Bound of the written positions.
  if (!new_data) {
    max_pos1 = N;
  } else {
    const uint64_t F = d.nof_filler_bits, P = Ncb - F;
    const uint64_t start = k0 < j.info ? k0 : (k0 < nsys ? j.info : k0 - F);
    if (start + d.rm_length >= P) {
      max_pos1 = Ncb;
    } else {
      const uint64_t q = start + d.rm_length - 1;
      max_pos1 = (q < j.info ? q : q + F) + 1;
    }
    max_pos1 = std::max<uint64_t>(max_pos1, nsys); // fillers hold +127
  }
  return LH_OK;
}

struct group {
  std::vector<dec_job> jobs;
  int maxZ = 0;
  int lmax = 4;
};

int launch_groups(lh_ctx* ctx, slot* s, size_t off, group* g, const int8_t* in, uint8_t* bits, int32_t* results,
                  hipStream_t stream)
{
  for (int bg = 1; bg <= 2; ++bg) {
    group& gr = g[bg - 1];
    if (gr.jobs.empty()) {
      continue;
    }
    const int nt = ((gr.maxZ + 63) / 64) * 64;
    if (decode_lds_bytes(bg, gr.maxZ, gr.lmax, nt) > 160 * 1024) {
      return fail(LH_ERR_CAPACITY, "LDS budget exceeded (Z=%d, L=%d)", gr.maxZ, gr.lmax);
    }
    hipError_t e = launch_decode(bg, reinterpret_cast<const dec_job*>(s->dev + off), (unsigned)gr.jobs.size(),
                                 gr.maxZ, gr.lmax, in, bits, results, ctx->d_crc_x, stream);
    if (e != hipSuccess) {
      return hip_fail(e, "decoder launch");
    }
    off += gr.jobs.size() * sizeof(dec_job);
  }
  return LH_OK;
}

size_t stage_groups(slot* s, size_t off, group* g)
{
  for (int k = 0; k < 2; ++k) {
    if (!g[k].jobs.empty()) {
      std::memcpy(s->host + off, g[k].jobs.data(), g[k].jobs.size() * sizeof(dec_job));
      off += g[k].jobs.size() * sizeof(dec_job);
    }
  }
  return off;
}

void add_job(group* g, unsigned bg, const dec_job& j)
{
  group& gr = g[bg - 1];
  gr.jobs.push_back(j);
  gr.maxZ = std::max<int>(gr.maxZ, j.Z);
  gr.lmax = std::max<int>(gr.lmax, (int)j.lbound);
}

} // namespace

extern "C" {

const char* lh_last_error(void)
{
  return g_err.c_str();
}

int lh_ctx_create(lh_ctx** out, int device, unsigned max_codeblocks)
{
  if (out == nullptr || max_codeblocks == 0) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ctx_create");
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    return hip_fail(e, "hipSetDevice");
  }
  lh_ctx* c = new lh_ctx();
  c->device = device;
  c->max_cb = max_codeblocks;
  c->slot_bytes = (size_t)max_codeblocks * (sizeof(dec_job) + sizeof(rm_job)) * 2 + 4096;
  std::vector<uint32_t> x(3 * (size_t)kCrcXLen);
  crc_x_table(LH_CRC24A, x.data());
  crc_x_table(LH_CRC24B, x.data() + kCrcXLen);
  crc_x_table(LH_CRC16, x.data() + 2 * kCrcXLen);
  e = hipMalloc(&c->d_crc_x, x.size() * 4);
  if (e == hipSuccess) {
    e = hipMemcpy(c->d_crc_x, x.data(), x.size() * 4, hipMemcpyHostToDevice);
  }
  for (int i = 0; i < kRing && e == hipSuccess; ++i) {
    e = hipHostMalloc(&c->ring[i].host, c->slot_bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
      e = hipMalloc(&c->ring[i].dev, c->slot_bytes);
    }
    if (e == hipSuccess) {
      e = hipEventCreateWithFlags(&c->ring[i].done, hipEventDisableTiming);
    }
  }
  if (e != hipSuccess) {
    lh_ctx_destroy(c);
    return hip_fail(e, "context allocation");
  }
  *out = c;
  return LH_OK;
}

int lh_ctx_destroy(lh_ctx* c)
{
  if (c == nullptr) {
    return LH_OK;
  }
  hipSetDevice(c->device);
  for (auto& s : c->ring) {
    if (s.done) {
      hipEventSynchronize(s.done);
      hipEventDestroy(s.done);
    }
    if (s.host) {
      hipHostFree(s.host);
    }
    if (s.dev) {
      hipFree(s.dev);
    }
  }
  if (c->d_crc_x) {
    hipFree(c->d_crc_x);
  }
  delete c;
  return LH_OK;
}

int lh_ldpc_decode(lh_ctx* ctx, const lh_decode_desc* descs, unsigned n, const int8_t* d_llrs, uint8_t* d_bits,
                   int32_t* d_results, uint32_t flags, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_llrs == nullptr || d_bits == nullptr || d_results == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ldpc_decode");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  group g[2];
  for (unsigned i = 0; i < n; ++i) {
    const lh_decode_desc& d = descs[i];
    int rc = check_decoder_cfg(d.cfg, i);
    if (rc != LH_OK) {
      return rc;
    }
    const int ci = crc_index(d.crc_poly);
    if (ci < 0) {
      return fail(LH_ERR_INVALID, "cb %u: invalid CRC polynomial 0x%x", i, d.crc_poly);
    }
    bg_info b = bginfo(d.cfg.base_graph);
    const unsigned Z = d.cfg.lifting_size;
    if (d.nof_llrs > (unsigned)b.nshort * Z || d.nof_llrs < (unsigned)(b.K + 2) * Z) {
      return fail(LH_ERR_INVALID, "cb %u: input length %u outside [%u, %u]", i, d.nof_llrs, (b.K + 2) * Z,
                  b.nshort * Z);
    }
    dec_job j{};
    j.in_off = d.llr_offset;
    j.bits_off = d.bits_offset;
    j.n_in = d.nof_llrs;
    j.Z = (uint16_t)Z;
    j.ils = (uint8_t)lifting_index(Z);
    j.max_iter = (uint8_t)d.cfg.max_iterations;
    j.F = (uint16_t)d.cfg.nof_filler_bits;
    j.crc_idx = (uint8_t)ci;
    j.flags = (uint8_t)(((flags & LH_FORCE_DECODING) ? kFlagForce : 0) |
                        ((flags & LH_EARLY_STOP_SYNDROME) ? kFlagEsSyndrome : 0));
    j.lbound = layer_bound(b, Z, d.nof_llrs);
    j.out_idx = i;
    add_job(g, d.cfg.base_graph, j);
  }
  hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  size_t bytes = stage_groups(s, 0, g);
  int rc = upload(ctx, s, bytes, st);
  if (rc != LH_OK) {
    return rc;
  }
  return launch_groups(ctx, s, 0, g, d_llrs, d_bits, d_results, st);
}

int lh_rate_dematch(lh_ctx* ctx, const lh_rm_desc* descs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                    int new_data, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_cw == nullptr || d_rm == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_rate_dematch");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  std::vector<rm_job> jobs(n);
  unsigned maxN = 0;
  for (unsigned i = 0; i < n; ++i) {
    uint64_t mp = 0;
    int rc = make_rm_job(descs[i], new_data != 0, jobs[i], mp, i);
    if (rc != LH_OK) {
      return rc;
    }
    maxN = std::max(maxN, jobs[i].N);
  }
  hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  std::memcpy(s->host, jobs.data(), n * sizeof(rm_job));
  int rc = upload(ctx, s, n * sizeof(rm_job), st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_rate_dematch(reinterpret_cast<const rm_job*>(s->dev), n, maxN, d_cw, d_rm, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "dematcher launch");
}

int lh_pusch_decode_batch(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                          uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                          unsigned nof_ldpc_iterations, void* stream)
{
  if (ctx == nullptr ||
      (n > 0 && (cbs == nullptr || d_cw == nullptr || d_rm == nullptr || d_bits == nullptr || d_results == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_pusch_decode_batch");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (nof_ldpc_iterations == 0 || nof_ldpc_iterations > 255) {
    return fail(LH_ERR_INVALID, "invalid number of LDPC iterations %u", nof_ldpc_iterations);
  }
  if (n == 0) {
    return LH_OK;
  }
  std::vector<rm_job> rjobs(n);
  group g[2];
  unsigned maxN = 0;
  for (unsigned i = 0; i < n; ++i) {
    const lh_pusch_cb_desc& d = cbs[i];
    uint64_t mp = 0;
    int rc = make_rm_job(d.rm, new_data != 0, rjobs[i], mp, i);
    if (rc != LH_OK) {
      return rc;
    }
    const int ci = crc_index(d.crc_poly);
    if (ci <= 0) {
      return fail(LH_ERR_INVALID, "cb %u: invalid CRC polynomial 0x%x", i, d.crc_poly);
    }
    if (d.nof_crc_bits != 16 && d.nof_crc_bits != 24) {
      return fail(LH_ERR_INVALID, "cb %u: invalid number of CRC bits %u", i, d.nof_crc_bits);
    }
    maxN = std::max(maxN, rjobs[i].N);
    bg_info b = bginfo(d.rm.base_graph);
    const unsigned Z = d.rm.lifting_size;
    dec_job j{};
    j.in_off = d.rm.rm_offset;
    j.bits_off = d.bits_offset;
    j.n_in = rjobs[i].N;
    j.Z = (uint16_t)Z;
    j.ils = (uint8_t)lifting_index(Z);
    j.max_iter = (uint8_t)nof_ldpc_iterations;
    j.F = (uint16_t)d.rm.nof_filler_bits;
    j.crc_idx = (uint8_t)ci;
    j.flags = use_early_stop ? 0 : kFlagCrcAfter;
    j.lbound = layer_bound(b, Z, mp);
    j.out_idx = i;
    add_job(g, d.rm.base_graph, j);
  }
  hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  std::memcpy(s->host, rjobs.data(), n * sizeof(rm_job));
  size_t off = n * sizeof(rm_job);
  size_t end = stage_groups(s, off, g);
  int rc = upload(ctx, s, end, st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_rate_dematch(reinterpret_cast<const rm_job*>(s->dev), n, maxN, d_cw, d_rm, st);
  if (e != hipSuccess) {
    return hip_fail(e, "dematcher launch");
  }
  return launch_groups(ctx, s, off, g, d_rm, d_bits, d_results, st);
}

int lh_crc_calculate(lh_ctx* ctx, uint32_t poly, const uint64_t* offsets, const uint32_t* nbits, unsigned n,
                     const uint8_t* d_bits, uint32_t* d_out, void* stream)
{
  const int ci = crc_index(poly);
  if (ctx == nullptr || ci <= 0 || (n > 0 && (offsets == nullptr || nbits == nullptr || d_bits == nullptr ||
                                              d_out == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_crc_calculate");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u buffers exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  for (unsigned i = 0; i < n; ++i) {
    if (nbits[i] > (unsigned)(kCrcXLen - 24)) {
      return fail(LH_ERR_INVALID, "buffer %u: %u bits exceed the supported %d", i, nbits[i], kCrcXLen - 24);
    }
  }
  hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  std::memcpy(s->host, offsets, n * 8);
  std::memcpy(s->host + n * 8, nbits, n * 4);
  int rc = upload(ctx, s, n * 12, st);
  if (rc != LH_OK) {
    return rc;
  }
  const int order = ci == 3 ? 16 : 24;
  hipError_t e = launch_crc(reinterpret_cast<const uint64_t*>(s->dev), reinterpret_cast<const uint32_t*>(s->dev + n * 8),
                            n, d_bits, ctx->d_crc_x + (ci - 1) * kCrcXLen, order, d_out, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "CRC launch");
}

} // extern "C"
