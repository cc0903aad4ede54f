// C ABI of the HIP LDPC receive chain (declared in include/ocudu_hip_ldpc.h).
//
// Host side: validates the reference-style descriptors, turns them into device jobs
// (layer-count bounds, k0, filler ranges), stages them through a small ring of pinned
// buffers and launches the kernels on the caller's stream.
#include "ocudu_hip_ldpc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <unordered_map>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bg_tables.inc"
#include "lh_internal.h"

namespace lh {
size_t decode_item_lds_bytes(int bg, int Z, int lmax, int G, int stage, bool ext8);
hipError_t launch_decode(launch_header* d_hdr, const dec_item* d_items, const dec_job* d_jobs, unsigned grid,
                         int nt, size_t lds, int8_t* rm, const int8_t* cw, uint8_t* bits, int32_t* results,
                         const uint32_t* crc_x, const uint32_t* gtab, const uint32_t* scale_tab, int fb_iters,
                         bool pairs, bool ext8, hipStream_t s, const fused_tb_args* fused = nullptr);
// (fused: a device pointer to the arguments of the fused TB finalize, or nullptr)
int decode_occupancy(int nt, size_t lds, bool learned, bool pairs, bool ext8);
hipError_t launch_reset_counters(uint32_t* ctr, unsigned words, hipStream_t s);
hipError_t launch_tb_finalize(const lh_tb_desc_dev* d_tbs, unsigned n, const uint8_t* cb_bits, const int32_t* cb_res,
                              uint8_t* tb, int32_t* status, const uint32_t* lut, const uint32_t* pow2, hipStream_t s,
                              const tb_signal* sig = nullptr);
hipError_t launch_rate_dematch(const rm_params* d_jobs, unsigned n, unsigned maxN, const int8_t* cw, int8_t* rm,
                               hipStream_t s);
hipError_t launch_gen_codeblocks(const gen_job* d_jobs, unsigned n, int8_t* llrs, uint8_t* msgs, const uint32_t* crc_x,
                                 const uint32_t* gtab, hipStream_t s);
hipError_t launch_demod(const demod_tile* tiles, unsigned nof_tiles, const demod_job* jobs, const uint32_t* tab,
                        const float* symbols, const float* noise, int8_t* llrs, hipStream_t stream);
hipError_t launch_demux(const demux_tile* tiles, unsigned nof_tiles, const demux_sym* syms, const uint32_t* tab,
                        const int8_t* llrs, int8_t* sch, int8_t* ack, int8_t* c1, int8_t* c2, hipStream_t stream);
hipError_t launch_encode_rate_match(const gen_job* d_jobs, unsigned n, uint8_t* msgs, uint8_t* out_bits,
                                    uint8_t* cw_bits, const uint32_t* crc_x, const uint32_t* gtab, bool any_small,
                                    size_t packed_cw_bytes, hipStream_t s);
size_t nms_train_lds_bytes(int bg, int Z, int L);
hipError_t launch_nms_train(int bg, const train_job* d_jobs, unsigned n, int nt, size_t lds, int T,
                            const float* alpha, const float* beta, const float* wt, float tau, const int8_t* in,
                            const uint8_t* tgt, float* ws, float* loss, float* galpha, float* gbeta,
                            const uint32_t* gtab, int backward, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, int n_alpha, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s);
hipError_t launch_gen_tb(const gen_tb_job* d_jobs, unsigned n, uint8_t* tb, uint8_t* msgs, const uint32_t* crc_x,
                         const uint32_t* lut24a, const uint32_t* pow2, hipStream_t s);
hipError_t launch_crc(const uint64_t* d_offs, const uint32_t* d_nbits, unsigned n, const uint8_t* bits,
                      const uint32_t* xtab, int order, uint32_t* out, hipStream_t s);
hipError_t scratch_address(uint8_t** ptr);
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

const unsigned kLiftingSizes[kNofLiftingSizes] = {
    2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12,  13,  14,  15,  16,  18,  20,  22,  24,  26,  28,  30,  32,  36,  40,
    44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352,
    384};

// Position among the 51 lifting sizes and lifting set index (TS38.212 Table 5.3.2-1) of
// every Z <= 384, -1 where Z is not a lifting size: O(1) per codeblock on the host path.
struct lifting_lut {
  int8_t pos[kMaxZ + 1];
  int8_t set[kMaxZ + 1];
  lifting_lut()
  {
    for (int z = 0; z <= kMaxZ; ++z) {
      pos[z] = -1;
      set[z] = -1;
    }
    for (int i = 0; i < kNofLiftingSizes; ++i) {
      pos[kLiftingSizes[i]] = (int8_t)i;
    }
    static const unsigned a_set[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; ++i) {
      for (unsigned v = a_set[i]; v <= (unsigned)kMaxZ; v *= 2) {
        set[v] = (int8_t)i;
      }
    }
  }
};
const lifting_lut kLift;

int lifting_position(unsigned Z)
{
  return Z <= (unsigned)kMaxZ ? kLift.pos[Z] : -1;
}

// Offset of the lifted edge table of graph (bg, Z) in the device table.
uint32_t graph_offset(unsigned bg, unsigned Z)
{
  const int pos = lifting_position(Z);
  return bg == 1 ? (uint32_t)pos * 316u : (uint32_t)(kNofLiftingSizes * 316 + pos * 197);
}

int lifting_index(unsigned Z)
{
  return Z <= (unsigned)kMaxZ ? kLift.set[Z] : -1;
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

constexpr int kRing = 16;

// Tuned launch parameters of one (BG, Z): codeblocks per workgroup (0 = as many as fit)
// and LDS staging of the codeword LLRs (1 = when it fits, 0 = never).
struct tune_entry {
  uint16_t groups = 0;
  uint8_t stage = 0; // LDS staging of the codeword LLRs (off: the one-pass gathers read HBM)
  int8_t pairs = -1; // row mode (lh_tune_set_rows)
};

struct slot {
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
  hipEvent_t done = nullptr;   // the kernels that read the slot have completed
  hipEvent_t copied = nullptr; // the upload of the slot has completed
  bool used = false;
};

} // namespace

struct lh_ctx {
  int device = 0;
  int nof_cus = 256;
  unsigned max_cb = 0;
  size_t slot_bytes = 0;
  uint32_t* d_crc_x = nullptr;
  uint32_t* d_crc_xz = nullptr; // the decoder's zero-padded CRC weights (kCrcXZLen per polynomial)
  uint32_t* d_gtab = nullptr;
  uint32_t* d_demod = nullptr; // LFSR jump matrices + demapper tables (kDemodTableWords)
  uint32_t* d_tbcrc = nullptr; // CRC24A byte table (256) + x^(2^k) mod G (32)
  uint32_t* d_scale = nullptr; // learned normalisation tables (kScaleTableWords)
  tune_entry tune[2][kNofLiftingSizes] = {}; // per (BG, Z): 0 = default
  bool scale_set[2] = {false, false};
  slot ring[kRing];
  int next = 0;
  // Descriptor uploads run here, so that a call's upload overlaps the kernels of the
  // previous call on the caller's stream (default priority, non-blocking).
  hipStream_t copy = nullptr;
  // This device's copy of the static scratch (lh_scratch.hip), for the allocation-free
  // entry point, and the host workspaces that entry point packs its launch in (reserved
  // for kNoAllocMaxCbs codeblocks at creation, so packing does not reallocate).
  uint8_t* d_scratch = nullptr;
  // Resident workgroups per compute unit of the decoder instances, by launch shape.
  struct occ_entry {
    int nt = 0;
    size_t lds = 0;
    int key = -1; // learned | pairs << 1
    int occ = 0;
  } occ_cache[32];
  int occ_next = 0;
  std::vector<dec_job> ws_jobs;
  std::vector<uint8_t> ws_bytes;
  std::vector<dec_item> ws_items, ws_tmp;
};

namespace {

// Returns a staging slot whose previous upload has completed, or nullptr (error set) when
// waiting for it fails: a device error is reported, not cleared.  Slot events never refer
// to destroyed streams (rebind_ring runs before a server destroys its streams).
slot* acquire_slot(lh_ctx* ctx)
{
  slot* s = &ctx->ring[ctx->next];
  ctx->next = (ctx->next + 1) % kRing;
  if (s->used) {
    const hipError_t e = hipEventSynchronize(s->done);
    if (e != hipSuccess) {
      hip_fail(e, "wait for a descriptor staging slot");
      return nullptr;
    }
  }
  return s;
}

// Before streams that staging-slot events may have been recorded on are destroyed: wait
// for every slot and re-record its events on the context's own stream, so that no event
// refers to a destroyed stream.
void rebind_ring(lh_ctx* ctx)
{
  for (auto& s : ctx->ring) {
    if (s.used) {
      (void)hipEventSynchronize(s.done);
      (void)hipEventSynchronize(s.copied);
      (void)hipEventRecord(s.done, ctx->copy);
      (void)hipEventRecord(s.copied, ctx->copy);
    }
  }
  (void)hipGetLastError();
}

// Marks the slot busy until the work enqueued on `stream` so far (the kernels that read
// its descriptors) has completed.
void retire(slot* s, hipStream_t stream)
{
  (void)hipEventRecord(s->done, stream);
}

// Uploads a slot.  A large block goes on the copy stream, which `stream` then waits for,
// so that the copy overlaps the kernels already queued there; a small one (a single
// slot's launch, latency-bound) goes directly on `stream`, without the event hop.
constexpr size_t kCopyStreamBytes = 64 * 1024;

int upload(lh_ctx* ctx, slot* s, size_t bytes, hipStream_t stream)
{
  hipError_t e;
  if (bytes < kCopyStreamBytes) {
    e = hipMemcpyAsync(s->dev, s->host, bytes, hipMemcpyHostToDevice, stream);
  } else {
    e = hipMemcpyAsync(s->dev, s->host, bytes, hipMemcpyHostToDevice, ctx->copy);
    if (e == hipSuccess) {
      e = hipEventRecord(s->copied, ctx->copy);
    }
    if (e == hipSuccess) {
      e = hipStreamWaitEvent(stream, s->copied, 0);
    }
  }
  if (e != hipSuccess) {
    return hip_fail(e, "descriptor upload");
  }
  e = hipEventRecord(s->done, stream);
  if (e != hipSuccess) {
    return hip_fail(e, "event record");
  }
  s->used = true;
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

// Fills the dematching parameters from a descriptor; max_pos1 receives the largest
// position that can hold a nonzero soft bit, plus one.
int make_rm_params(const lh_rm_desc& d, bool new_data, rm_params& j, uint64_t& max_pos1, unsigned i)
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
  // k0 (ldpc_rate_dematcher_impl.cpp:87-88, TS38.212 Table 5.4.2.1-2)
  const double* f = d.base_graph == 1 ? k0_bg1 : k0_bg2;
  const unsigned k0 = (unsigned)std::floor((f[d.rv] * Ncb) / N) * Z;
  const unsigned F = d.nof_filler_bits;
  j = rm_params{};
  j.cw_off = d.cw_offset;
  j.rm_off = d.rm_offset;
  j.E = d.rm_length;
  j.N = N;
  j.Ncb = Ncb;
  j.info = nsys - F;
  j.nsys = nsys;
  j.P = Ncb - F;
  j.start = k0 < j.info ? k0 : (k0 < nsys ? j.info : k0 - F);
  j.Kq = d.rm_length / qm;
  // Reciprocal biased up by 2^-21: (float)i * rcp_kq then truncates to i / Kq exactly for
  // i, Kq < 2^19 (the product's rounding errors stay below the bias).
  j.rcp_kq = (float)((1.0 + 1.0 / 2097152.0) / (double)j.Kq);
  j.Qm = (uint16_t)qm;
  j.flags = new_data ? kFlagNewData : 0;
  if (!new_data) {
    max_pos1 = N;
  } else {
    if ((uint64_t)j.start + d.rm_length >= j.P) {
      max_pos1 = Ncb;
    } else {
      const uint64_t q = (uint64_t)j.start + d.rm_length - 1;
      max_pos1 = (q < j.info ? q : q + F) + 1;
    }
    max_pos1 = std::max<uint64_t>(max_pos1, nsys); // fillers hold +127
  }
  return LH_OK;
}

// Jobs of one decoder launch, turned into the (BG, Z)-sorted work queue of the
// persistent kernel.
struct launch_plan {
  std::vector<dec_job> jobs;
};

// LDS budget for staging the codeword LLRs of one item.
constexpr size_t kLdsBudget = 160 * 1024;

// Codeblocks per work item: as many as fit the workgroup (G * lanes <= threads).
int groups_per_item(int lanes, int nt)
{
  return std::max(1, std::min(kMaxGroups, nt / lanes));
}

// Whether a codeblock class in a launch of n codeblocks favours two check rows per thread
// (even Z; half the threads, each with 1.4x the instructions of one row): when a compute
// unit's resident codeblocks are limited by its waves rather than its LDS and the launch
// holds enough codeblocks to fill every compute unit at one row per thread; otherwise one
// row per thread has the lower latency per codeblock.
// A tuned row mode (lh_tune_set_rows) overrides the choice.  Resident codeblocks per
// compute unit: 16 waves of 64 threads at one row per thread (128 VGPRs), 12 waves at two
// (168 VGPRs, Z/2 threads per codeblock), i.e. at most 1.5x as many; with 1.4x the
// instructions per codeblock the mode needs the LDS to hold nearly all of them.
bool pair_rows(const lh_ctx* ctx, const dec_job& j, size_t n)
{
  const int bg = j.gt_off < kNofLiftingSizes * 316 ? 1 : 2;
  const int Z = (int)j.Z;
  const tune_entry& te = ctx->tune[bg - 1][lifting_position((unsigned)Z)];
  if ((Z & 1) || te.pairs == 0) {
    return false;
  }
  if (te.pairs == 1) {
    return true;
  }
  const int stage = ((j.flags & kFlagFused) && te.stage != 0) ? (((int)j.rm.E + 15) & ~15) : 0;
  const double by_waves = 1024.0 / Z;
  const double by_lds = (double)kLdsBudget / (double)decode_item_lds_bytes(bg, Z, j.lbound, 1, stage, false);
  return by_lds >= 1.45 * by_waves && (double)n >= by_waves * ctx->nof_cus;
}

// A packed decoder launch: four work-queue headers (pass 1 row-pair items, pass 1 one-row
// items, then the same two for the fallback pass) | items (row-pair items first) | jobs,
// and the launch shape of each row mode (index 1: row pairs).
struct plan_block {
  std::vector<uint8_t> bytes;
  std::vector<dec_item> items, tmp; // work items while packing
  size_t o_items = 0, o_jobs = 0;
  size_t ctr_words = 0;             // sub-queue counters behind the headers
  size_t size = 0;                  // bytes of the packed block
  size_t n_items[2] = {0, 0};
  unsigned grid[2] = {0, 0};        // persistent workgroups per row mode
  int nt[2] = {64, 64};
  size_t lds[2] = {0, 0};
  bool ext8 = false; // one-row launch with int8 extension-layer messages
};

// decode_occupancy, remembered per launch shape (the runtime query costs microseconds).
int occupancy(lh_ctx* ctx, int nt, size_t lds, bool learned, bool pairs, bool ext8 = false)
{
  const int key = (learned ? 1 : 0) | (pairs ? 2 : 0) | (ext8 ? 4 : 0);
  for (const auto& c : ctx->occ_cache) {
    if (c.key == key && c.nt == nt && c.lds == lds) {
      return c.occ;
    }
  }
  auto& c = ctx->occ_cache[ctx->occ_next];
  ctx->occ_next = (ctx->occ_next + 1) % 32;
  c.nt = nt;
  c.lds = lds;
  c.key = key;
  c.occ = decode_occupancy(nt, lds, learned, pairs, ext8);
  return c.occ;
}

// Sorts the jobs largest first (longest-processing-time order for the dynamic queue; equal
// (BG, Z) stay adjacent so they pack into shared items), decides the row mode of every
// (BG, Z, layer bound) class, packs items of up to G codeblocks of one class -- G from
// the tuning table, else as many as fit the workgroup -- and decides LDS staging of the
// codeword LLRs per item.
// The block is written to dst (cap bytes, e.g. a pinned staging slot), or to pb.bytes
// when dst is null.
int build_block(lh_ctx* ctx, launch_plan& plan, plan_block& pb, uint8_t* dst = nullptr, size_t cap = 0)
{
  std::vector<dec_job>& jobs = plan.jobs;
  const size_t n_cbs = jobs.size();
  // Row mode: a class with a tuned mode keeps it; the others run row pairs when every
  // class of the launch favours them (a mixed launch runs its two queues one after the
  // other), else one row per thread.
  bool pairs_launch = !jobs.empty();
  for (const dec_job& j : jobs) {
    if (!pair_rows(ctx, j, n_cbs)) {
      pairs_launch = false;
      break;
    }
  }
  auto pairs_of_job = [&](const dec_job& j) {
    const int bg = j.gt_off < kNofLiftingSizes * 316 ? 1 : 2;
    const int tuned = ctx->tune[bg - 1][lifting_position(j.Z)].pairs;
    return tuned >= 0 ? tuned == 1 : pairs_launch;
  };
  // Workgroup size per row mode: the most threads per codeblock, or a tuned packing of G
  // codeblocks of one (BG, Z), rounded up to whole waves.
  int nt[2] = {64, 64};
  for (const dec_job& j : jobs) {
    const int bg = j.gt_off < kNofLiftingSizes * 316 ? 1 : 2;
    const tune_entry& te = ctx->tune[bg - 1][lifting_position(j.Z)];
    const int m = pairs_of_job(j) ? 1 : 0;
    const int lanes = m ? (int)j.Z / 2 : (int)j.Z;
    // Untuned: codeblocks of more than half a wave are packed up to 256 threads (at most
    // 4 per workgroup); smaller ones fill one wave.
    const int g_default = lanes > 32 ? std::min(4, std::max(1, 256 / lanes)) : 1;
    nt[m] = std::max(nt[m], ((g_default * lanes + 63) / 64) * 64);
    if (te.groups > 0) {
      nt[m] = std::max(nt[m], ((te.groups * lanes + 63) / 64) * 64);
    }
  }
  nt[0] = std::min(nt[0], kMaxThreads);
  nt[1] = std::min(nt[1], kMaxThreads);
  // Total order (ties by result index), so an unstable sort, which needs no temporary
  // buffer, gives the same queue as a stable one; a batch already in that order (a
  // homogeneous slot) is not sorted again.
  const auto by_work = [](const dec_job& a, const dec_job& b) {
    const int ka = (int)a.Z * (a.lbound + 22), kb = (int)b.Z * (b.lbound + 22);
    const int bga = a.gt_off < kNofLiftingSizes * 316 ? 1 : 2, bgb = b.gt_off < kNofLiftingSizes * 316 ? 1 : 2;
    if (ka != kb) {
      return ka > kb;
    }
    if (bga != bgb) {
      return bga < bgb;
    }
    if (a.Z != b.Z) {
      return a.Z > b.Z;
    }
    return a.out_idx < b.out_idx;
  };
  if (!std::is_sorted(jobs.begin(), jobs.end(), by_work)) {
    std::sort(jobs.begin(), jobs.end(), by_work);
  }
  std::vector<dec_item>& items = pb.items;
  items.clear();
  size_t lds[2] = {0, 0};
  for (size_t i = 0; i < jobs.size();) {
    const int bg = jobs[i].gt_off < kNofLiftingSizes * 316 ? 1 : 2;
    const int Z = jobs[i].Z;
    const tune_entry& te = ctx->tune[bg - 1][lifting_position((unsigned)Z)];
    const int m = pairs_of_job(jobs[i]) ? 1 : 0;
    const int lanes = m ? Z / 2 : Z;
    int G = te.groups > 0 ? std::min<int>(te.groups, groups_per_item(lanes, nt[m])) : groups_per_item(lanes, nt[m]);
    // Fewer codeblocks per item when their soft bits would not fit the LDS.
    while (G > 1 && decode_item_lds_bytes(bg, Z, jobs[i].lbound, G, 0, false) > kLdsBudget) {
      --G;
    }
    dec_item it{};
    it.first_job = (uint32_t)i;
    it.bg = (uint8_t)bg;
    it.rows2 = (uint8_t)m;
    int lmax = 4, maxE = 0;
    bool fused = false;
    // Scatter load: new data where every position takes at most one LLR (E <= P).
    bool scatter = Z % 8 == 0;
    size_t k = i;
    while (k < jobs.size() && (int)(k - i) < G && jobs[k].Z == Z && jobs[k].gt_off == jobs[i].gt_off &&
           jobs[k].lbound == jobs[i].lbound) {
      lmax = std::max<int>(lmax, jobs[k].lbound);
      const uint8_t fl = jobs[k].flags;
      if (fl & kFlagFused) {
        fused = true;
        maxE = std::max<int>(maxE, (int)jobs[k].rm.E);
      }
      scatter = scatter && (fl & kFlagFused) && (fl & kFlagNewData) && !(fl & kFlagPassed) &&
                jobs[k].rm.E <= jobs[k].rm.P;
      ++k;
    }
    it.ngroups = (uint16_t)(k - i);
    it.lmax = (uint16_t)lmax;
    it.scatter = scatter ? 1 : 0;
    size_t need = decode_item_lds_bytes(bg, Z, lmax, it.ngroups, 0, false);
    if (need > kLdsBudget) {
      return fail(LH_ERR_CAPACITY, "LDS budget exceeded (Z=%d, L=%d)", Z, lmax);
    }
    const int stage = (fused && !scatter && te.stage != 0) ? ((maxE + 15) & ~15) : 0;
    if (stage > 0 && decode_item_lds_bytes(bg, Z, lmax, it.ngroups, stage, false) <= kLdsBudget) {
      it.stage = (uint32_t)stage;
      need = decode_item_lds_bytes(bg, Z, lmax, it.ngroups, stage, false);
    }
    // At least the scratch of a fused TB finalize (between items) and the queue word.
    lds[m] = std::max(lds[m], std::max(need, kFusedTbLdsBytes + 64));
    items.push_back(it);
    i = k;
  }
  // Row-pair items first (each mode keeps its largest-first order).
  pb.tmp.clear();
  for (const dec_item& it : items) {
    if (it.rows2) {
      pb.tmp.push_back(it);
    }
  }
  const size_t n_pair = pb.tmp.size();
  for (const dec_item& it : items) {
    if (!it.rows2) {
      pb.tmp.push_back(it);
    }
  }
  items.swap(pb.tmp);
  pb.n_items[1] = n_pair;
  pb.n_items[0] = items.size() - n_pair;
  // Persistent workgroups per row mode: as many as are resident on the device at once (the
  // plain and the learned instance alike), at most one per item.  Items beyond the first
  // round form the sub-queues, whose counters follow the headers.
  launch_header hdr[4]{};
  size_t ctr_words = 0;
  for (int m = 0; m < 2; ++m) {
    pb.grid[m] = 0;
    if (pb.n_items[m] == 0) {
      continue;
    }
    const int per_cu = std::max(1, std::min(occupancy(ctx, nt[m], lds[m], false, m == 1),
                                             occupancy(ctx, nt[m], lds[m], true, m == 1)));
    pb.grid[m] = (unsigned)std::min<size_t>(pb.n_items[m], (size_t)ctx->nof_cus * per_cu);
  }
  // The one-row launch keeps its extension-layer messages as int8 (fewer instructions per
  // layer, more LDS per codeblock) when all its items are resident at once with that LDS:
  // the latency regime, where no item waits for a workgroup.
  pb.ext8 = false;
  if (pb.n_items[0] > 0) {
    size_t lds8 = 0;
    for (size_t k = n_pair; k < items.size(); ++k) {
      const dec_item& it = items[k];
      lds8 = std::max(lds8, decode_item_lds_bytes(it.bg, (int)jobs[it.first_job].Z, it.lmax, it.ngroups,
                                                  (int)it.stage, true));
    }
    if (lds8 <= kLdsBudget) {
      const int per_cu8 = std::min(occupancy(ctx, nt[0], lds8, false, false, true),
                                   occupancy(ctx, nt[0], lds8, true, false, true));
      if (per_cu8 >= 1 && pb.n_items[0] <= (size_t)ctx->nof_cus * (size_t)per_cu8) {
        pb.ext8 = true;
        lds[0] = lds8;
        pb.grid[0] = (unsigned)pb.n_items[0];
      }
    }
  }
  for (int pass = 0; pass < 2; ++pass) {
    for (int m = 1; m >= 0; --m) {
      launch_header& h = hdr[2 * pass + (1 - m)];
      h.n_items = (uint32_t)pb.n_items[m];
      h.first_item = m == 1 ? 0u : (uint32_t)n_pair;
      if (pb.n_items[m] > pb.grid[m]) {
        h.nq = std::min<uint32_t>(kMaxSubQueues, pb.grid[m]);
        h.ctr_off = (uint32_t)((sizeof(hdr) - (size_t)(2 * pass + (1 - m)) * sizeof(launch_header)) / 4 + ctr_words);
        ctr_words += (size_t)h.nq * kCounterStride;
      }
    }
  }
  pb.ctr_words = ctr_words;
  pb.o_items = sizeof(hdr) + 4 * ctr_words;
  pb.o_jobs = pb.o_items + ((items.size() * sizeof(dec_item) + 15) & ~(size_t)15);
  pb.size = pb.o_jobs + jobs.size() * sizeof(dec_job);
  uint8_t* out = dst;
  if (out == nullptr) {
    pb.bytes.resize(pb.size);
    out = pb.bytes.data();
  } else if (pb.size > cap) {
    return fail(LH_ERR_CAPACITY, "descriptor block of %zu bytes exceeds the staging capacity %zu", pb.size, cap);
  }
  // Headers, zeroed counters, items (zero padding up to the jobs), jobs.
  std::memcpy(out, hdr, sizeof(hdr));
  std::memset(out + sizeof(hdr), 0, pb.o_items - sizeof(hdr));
  std::memcpy(out + pb.o_items, items.data(), items.size() * sizeof(dec_item));
  std::memset(out + pb.o_items + items.size() * sizeof(dec_item), 0,
              pb.o_jobs - pb.o_items - items.size() * sizeof(dec_item));
  std::memcpy(out + pb.o_jobs, jobs.data(), jobs.size() * sizeof(dec_job));
  pb.nt[0] = nt[0];
  pb.nt[1] = nt[1];
  pb.lds[0] = lds[0];
  pb.lds[1] = lds[1];
  return LH_OK;
}

// Launches a packed block resident at d_block: the decoder (learned with scale_tab) over
// the row-pair queue, then the one-row queue (each a persistent launch).  With fb_iters > 0
// the plain decoder is the fallback over the same descriptors: after a learned decoder,
// inside its launch (each workgroup re-decodes its item's failed codeblocks at once); after
// the plain one, as a second pass.
hipError_t launch_block(lh_ctx* ctx, const plan_block& pb, uint8_t* d_block, int8_t* rm, const int8_t* cw,
                        uint8_t* bits, int32_t* results, hipStream_t stream, const uint32_t* scale_tab,
                        unsigned fb_iters, const fused_tb_args* fused = nullptr)
{
  launch_header* d_hdr = reinterpret_cast<launch_header*>(d_block);
  const dec_item* d_items = reinterpret_cast<const dec_item*>(d_block + pb.o_items);
  const dec_job* d_jobs = reinterpret_cast<const dec_job*>(d_block + pb.o_jobs);
  const bool inline_fb = scale_tab != nullptr && fb_iters > 0;
  hipError_t e = hipSuccess;
  for (int pass = 0; pass < (fb_iters > 0 && !inline_fb ? 2 : 1) && e == hipSuccess; ++pass) {
    const uint32_t* tab = pass == 0 ? scale_tab : nullptr;
    const int fb = inline_fb || pass == 1 ? (int)fb_iters : 0;
    for (int m = 1; m >= 0 && e == hipSuccess; --m) {
      if (pb.n_items[m] == 0) {
        continue;
      }
      e = launch_decode(d_hdr + 2 * pass + (1 - m), d_items, d_jobs, pb.grid[m], pb.nt[m], pb.lds[m], rm, cw, bits,
                        results, ctx->d_crc_xz, ctx->d_gtab, tab, fb, m == 1, m == 0 && pb.ext8, stream, fused);
    }
  }
  return e;
}

// Packs, uploads (staging ring) and launches a plan.
int run_plan(lh_ctx* ctx, launch_plan& plan, int8_t* rm, const int8_t* cw, uint8_t* bits, int32_t* results,
             hipStream_t stream, slot* s, size_t off, const uint32_t* scale_tab = nullptr, unsigned fb_iters = 0,
             const fused_tb_args* fused = nullptr)
{
  if (plan.jobs.empty()) {
    return LH_OK;
  }
  plan_block pb;
  // Packed straight into the pinned staging slot.
  int rc = build_block(ctx, plan, pb, s->host + off, ctx->slot_bytes - off);
  if (rc != LH_OK) {
    return rc;
  }
  rc = upload(ctx, s, off + pb.size, stream);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_block(ctx, pb, s->dev + off, rm, cw, bits, results, stream, scale_tab, fb_iters, fused);
  retire(s, stream);
  return e == hipSuccess ? LH_OK : hip_fail(e, "decoder launch");
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
  // Per codeblock its descriptors; per launch the headers and sub-queue counters.
  c->slot_bytes = (size_t)max_codeblocks * (sizeof(dec_job) + sizeof(rm_params) + sizeof(dec_item) + sizeof(lh_tb_desc)) +
                  4096 + 4 * kMaxSubQueues * kCounterStride * 4;
  (void)hipDeviceGetAttribute(&c->nof_cus, hipDeviceAttributeMultiprocessorCount, device);
  std::vector<uint32_t> x(3 * (size_t)kCrcXLen);
  crc_x_table(LH_CRC24A, x.data());
  crc_x_table(LH_CRC24B, x.data() + kCrcXLen);
  crc_x_table(LH_CRC16, x.data() + 2 * kCrcXLen);
  e = hipMalloc(&c->d_crc_x, x.size() * 4);
  if (e == hipSuccess) {
    e = hipMemcpy(c->d_crc_x, x.data(), x.size() * 4, hipMemcpyHostToDevice);
  }
  {
    std::vector<uint32_t> xz(3 * (size_t)kCrcXZLen, 0u);
    for (int k = 0; k < 3; ++k) {
      const int order = k == 2 ? 16 : 24;
      for (int e2 = order; e2 < kCrcXLen; ++e2) {
        xz[(size_t)k * kCrcXZLen + kCrcPad + e2] = x[(size_t)k * kCrcXLen + e2];
      }
    }
    if (e == hipSuccess) {
      e = hipMalloc(&c->d_crc_xz, xz.size() * 4);
    }
    if (e == hipSuccess) {
      e = hipMemcpy(c->d_crc_xz, xz.data(), xz.size() * 4, hipMemcpyHostToDevice);
    }
  }
  // Lifted edge tables: (col * Z) | ((V mod Z) << 16) per edge, rows in order
  // (ldpc_graph_impl.cpp:11-39, ldpc_luts_impl.cpp:4505-4523).
  // Behind them the row-pair tables of the decoder's two-rows-per-thread mode (even Z):
  // two words per edge, col * Z | S << 16 and the 16-bit pair S | (S - T) << 16, with
  // h = Z / 2, f = (shift >= h), S = 2 (shift mod h) + f and T = Z - 1 + 2f.  Then the
  // decoder's one-row tables: two words per edge, the shift and col * Z.
  std::vector<uint32_t> gt(5 * (size_t)kGraphTableLen + kGraphTablePad);
  for (int pos = 0; pos < kNofLiftingSizes; ++pos) {
    const unsigned Z = kLiftingSizes[pos];
    const int ils = lifting_index(Z);
    auto put = [&](int bg, int k, unsigned col, unsigned v) {
      const unsigned sh = v % Z;
      const size_t o = graph_offset(bg, Z) + (size_t)k;
      gt[o] = (uint32_t)(col * Z) | ((uint32_t)sh << 16);
      gt[3 * (size_t)kGraphTableLen + 2 * o] = sh;
      gt[3 * (size_t)kGraphTableLen + 2 * o + 1] = (uint32_t)(col * Z);
      if (Z % 2 == 0) {
        const unsigned h = Z / 2, f = sh >= h ? 1u : 0u;
        const unsigned S = 2 * (sh - f * h) + f, T = Z - 1 + 2 * f;
        gt[kGraphTableLen + 2 * o] = (uint32_t)(col * Z) | (S << 16);
        gt[kGraphTableLen + 2 * o + 1] = (S & 0xffffu) | (((S - T) & 0xffffu) << 16);
      }
    };
    for (int k = 0; k < 316; ++k) {
      put(1, k, lh_bg1_col[k], lh_bg1_v[ils][k]);
    }
    for (int k = 0; k < 197; ++k) {
      put(2, k, lh_bg2_col[k], lh_bg2_v[ils][k]);
    }
  }
  // CRC24A byte table, x^(2^k) mod G, and x^(8 d 16^i) mod G (digit d < 16 of a byte count,
  // i < kTbShiftDigits) for the transport-block CRC.
  std::vector<uint32_t> tbc(256 + 32 + 16 * kTbShiftDigits);
  for (uint32_t b = 0; b < 256; ++b) {
    uint32_t v = b << 16;
    for (int k = 0; k < 8; ++k) {
      v = (v << 1) ^ ((v & 0x800000u) ? (LH_CRC24A & 0xffffffu) : 0u);
    }
    tbc[b] = v & 0xffffffu;
  }
  {
    auto mulmod = [](uint32_t a, uint32_t b) {
      uint32_t res = 0;
      for (int i = 23; i >= 0; --i) {
        res <<= 1;
        if (res & 0x1000000u) {
          res ^= LH_CRC24A;
        }
        if ((b >> i) & 1u) {
          res ^= a;
        }
      }
      return res;
    };
    uint32_t v = 2; // x^1
    for (int k = 0; k < 32; ++k) {
      tbc[256 + k] = v;
      v = mulmod(v, v);
    }
    uint32_t base = tbc[256 + 3]; // x^8: one byte
    for (int i = 0; i < kTbShiftDigits; ++i) {
      uint32_t* w = &tbc[256 + 32 + 16 * i];
      w[0] = 1;
      for (int d = 1; d < 16; ++d) {
        w[d] = mulmod(w[d - 1], base);
      }
      base = mulmod(w[15], base); // x^(8 16^(i+1))
    }
  }
  if (e == hipSuccess) {
    e = hipMalloc(&c->d_tbcrc, tbc.size() * 4);
  }
  if (e == hipSuccess) {
    e = hipMemcpy(c->d_tbcrc, tbc.data(), tbc.size() * 4, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) {
    e = hipMalloc(&c->d_scale, kScaleTableWords * 4);
  }
  if (e == hipSuccess) {
    e = hipMemset(c->d_scale, 0, kScaleTableWords * 4);
  }
  if (e == hipSuccess) {
    e = hipMalloc(&c->d_gtab, gt.size() * 4);
  }
  if (e == hipSuccess) {
    e = hipMemcpy(c->d_gtab, gt.data(), gt.size() * 4, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) {
    std::vector<uint32_t> dt(kDemodTableWords);
    build_demod_tables(dt.data());
    e = hipMalloc(&c->d_demod, dt.size() * 4);
    if (e == hipSuccess) {
      e = hipMemcpy(c->d_demod, dt.data(), dt.size() * 4, hipMemcpyHostToDevice);
    }
  }
  for (int i = 0; i < kRing && e == hipSuccess; ++i) {
    e = hipHostMalloc(&c->ring[i].host, c->slot_bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
      e = hipMalloc(&c->ring[i].dev, c->slot_bytes);
    }
    if (e == hipSuccess) {
      e = hipEventCreateWithFlags(&c->ring[i].done, hipEventDisableTiming);
    }
    if (e == hipSuccess) {
      e = hipEventCreateWithFlags(&c->ring[i].copied, hipEventDisableTiming);
    }
  }
  if (e == hipSuccess) {
    e = hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking);
  }
  if (e != hipSuccess) {
    lh_ctx_destroy(c);
    return hip_fail(e, "context allocation");
  }
  // With the device current: its address of the static scratch.
  e = scratch_address(&c->d_scratch);
  if (e != hipSuccess) {
    lh_ctx_destroy(c);
    return hip_fail(e, "hipGetSymbolAddress(scratch)");
  }
  const size_t ws_cbs = std::min<size_t>(max_codeblocks, kNoAllocMaxCbs);
  c->ws_jobs.reserve(ws_cbs);
  c->ws_items.reserve(ws_cbs);
  c->ws_tmp.reserve(ws_cbs);
  c->ws_bytes.reserve(kScratchBytes);
  *out = c;
  return LH_OK;
}

int lh_ctx_destroy(lh_ctx* c)
{
  if (c == nullptr) {
    return LH_OK;
  }
  (void)hipSetDevice(c->device);
  for (auto& s : c->ring) {
    if (s.done) {
      (void)hipEventSynchronize(s.done);
      (void)hipEventDestroy(s.done);
    }
    if (s.copied) {
      (void)hipEventSynchronize(s.copied);
      (void)hipEventDestroy(s.copied);
    }
    if (s.host) {
      (void)hipHostFree(s.host);
    }
    if (s.dev) {
      (void)hipFree(s.dev);
    }
  }
  if (c->copy) {
    (void)hipStreamSynchronize(c->copy);
    (void)hipStreamDestroy(c->copy);
  }
  if (c->d_crc_xz) {
    (void)hipFree(c->d_crc_xz);
  }
  if (c->d_crc_x) {
    (void)hipFree(c->d_crc_x);
  }
  if (c->d_demod) {
    (void)hipFree(c->d_demod);
  }
  if (c->d_gtab) {
    (void)hipFree(c->d_gtab);
  }
  if (c->d_tbcrc) {
    (void)hipFree(c->d_tbcrc);
  }
  if (c->d_scale) {
    (void)hipFree(c->d_scale);
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
  launch_plan plan;
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
    j.lbound = (uint16_t)layer_bound(b, Z, d.nof_llrs);
    j.gt_off = graph_offset(d.cfg.base_graph, Z);
    j.out_idx = i;
    plan.jobs.push_back(j);
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  return run_plan(ctx, plan, const_cast<int8_t*>(d_llrs), nullptr, d_bits, d_results, st, s, 0);
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
  std::vector<rm_params> jobs(n);
  unsigned maxN = 0;
  for (unsigned i = 0; i < n; ++i) {
    uint64_t mp = 0;
    int rc = make_rm_params(descs[i], new_data != 0, jobs[i], mp, i);
    if (rc != LH_OK) {
      return rc;
    }
    maxN = std::max(maxN, jobs[i].N);
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  std::memcpy(s->host, jobs.data(), n * sizeof(rm_params));
  int rc = upload(ctx, s, n * sizeof(rm_params), st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_rate_dematch(reinterpret_cast<const rm_params*>(s->dev), n, maxN, d_cw, d_rm, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "dematcher launch");
}

} // extern "C"

namespace {

// Decoder jobs of a PUSCH batch (fused dematching + decoding, HARQ write-back).
// new_data_per_cb (optional) overrides new_data per codeblock.
// passed_iters (optional): per codeblock, 0 or the iteration count of an earlier
// transmission whose CRC passed (combine only, report that count).
int pusch_jobs(const lh_pusch_cb_desc* cbs, unsigned n, int new_data_all, int use_early_stop,
               unsigned nof_ldpc_iterations, launch_plan& plan, const uint8_t* new_data_per_cb = nullptr,
               const uint8_t* passed_iters = nullptr)
{
  if (nof_ldpc_iterations == 0 || nof_ldpc_iterations > 255) {
    return fail(LH_ERR_INVALID, "invalid number of LDPC iterations %u", nof_ldpc_iterations);
  }
  plan.jobs.reserve(plan.jobs.size() + n);
  for (unsigned i = 0; i < n; ++i) {
    const lh_pusch_cb_desc& d = cbs[i];
    const int new_data = new_data_per_cb ? new_data_per_cb[i] : new_data_all;
    uint64_t mp = 0;
    dec_job j{};
    int rc = make_rm_params(d.rm, new_data != 0, j.rm, mp, i);
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
    bg_info b = bginfo(d.rm.base_graph);
    const unsigned Z = d.rm.lifting_size;
    j.in_off = d.rm.rm_offset;
    j.bits_off = d.bits_offset;
    j.n_in = j.rm.N;
    j.Z = (uint16_t)Z;
    j.ils = (uint8_t)lifting_index(Z);
    j.max_iter = (uint8_t)nof_ldpc_iterations;
    j.F = (uint16_t)d.rm.nof_filler_bits;
    j.crc_idx = (uint8_t)ci;
    j.flags = (uint8_t)((use_early_stop ? 0 : kFlagCrcAfter) | kFlagFused | kFlagHarqWrite |
                        (new_data ? kFlagNewData : 0));
    j.lbound = (uint16_t)layer_bound(b, Z, mp);
    j.gt_off = graph_offset(d.rm.base_graph, Z);
    j.out_idx = i;
    if (passed_iters != nullptr && passed_iters[i] > 0 && !new_data) {
      j.flags |= kFlagPassed;
      j.preset = passed_iters[i];
    }
    plan.jobs.push_back(j);
  }
  return LH_OK;
}

int check_pusch_args(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                     uint8_t* d_bits, int32_t* d_results, const char* what)
{
  if (ctx == nullptr ||
      (n > 0 && (cbs == nullptr || d_cw == nullptr || d_rm == nullptr || d_bits == nullptr || d_results == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to %s", what);
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  return LH_OK;
}

} // namespace

extern "C" {

int lh_pusch_decode_batch(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                          uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                          unsigned nof_ldpc_iterations, void* stream)
{
  int rc = check_pusch_args(ctx, cbs, n, d_cw, d_rm, d_bits, d_results, "lh_pusch_decode_batch");
  if (rc != LH_OK) {
    return rc;
  }
  launch_plan plan;
  rc = pusch_jobs(cbs, n, new_data, use_early_stop, nof_ldpc_iterations, plan);
  if (rc != LH_OK || n == 0) {
    return rc;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  return run_plan(ctx, plan, d_rm, d_cw, d_bits, d_results, st, s, 0);
}

int lh_pusch_decode_batch_noalloc(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw,
                                  int8_t* d_rm, uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                                  unsigned nof_ldpc_iterations, void* stream)
{
  int rc = check_pusch_args(ctx, cbs, n, d_cw, d_rm, d_bits, d_results, "lh_pusch_decode_batch_noalloc");
  if (rc != LH_OK) {
    return rc;
  }
  if (n > kNoAllocMaxCbs) {
    return fail(LH_ERR_CAPACITY, "lh_pusch_decode_batch_noalloc: %u codeblocks exceed the static scratch maximum of %u",
                n, kNoAllocMaxCbs);
  }
  if (ctx->d_scratch == nullptr) {
    return fail(LH_ERR_INVALID, "lh_pusch_decode_batch_noalloc: no scratch address for device %d", ctx->device);
  }
  // The launch is packed in the context's reserved host vectors (swapped in and back out,
  // never reallocated) and its descriptor block goes to the static scratch.
  launch_plan plan;
  plan_block pb;
  struct swap_guard {
    lh_ctx* c;
    launch_plan& pl;
    plan_block& b;
    swap_guard(lh_ctx* c_, launch_plan& pl_, plan_block& b_) : c(c_), pl(pl_), b(b_) { swap_all(); }
    ~swap_guard() { swap_all(); }
    void swap_all()
    {
      pl.jobs.swap(c->ws_jobs);
      b.bytes.swap(c->ws_bytes);
      b.items.swap(c->ws_items);
      b.tmp.swap(c->ws_tmp);
    }
  } guard(ctx, plan, pb);
  plan.jobs.clear();
  rc = pusch_jobs(cbs, n, new_data, use_early_stop, nof_ldpc_iterations, plan);
  if (rc != LH_OK || n == 0) {
    return rc;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  // Packed in the slot's pinned host half, then copied to the scratch on the caller's
  // stream, behind the previous call's kernels, which read the same scratch.  The
  // sub-queue counters arrive as zeros with the block.
  rc = build_block(ctx, plan, pb, s->host, std::min(kScratchBytes, ctx->slot_bytes));
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = hipMemcpyAsync(ctx->d_scratch, s->host, pb.size, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    return hip_fail(e, "descriptor upload to the scratch");
  }
  s->used = true;
  e = launch_block(ctx, pb, ctx->d_scratch, d_rm, d_cw, d_bits, d_results, st, nullptr, 0);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "decoder launch");
}

int lh_set_scale_table(lh_ctx* ctx, unsigned base_graph, const uint16_t* alpha_q, const int16_t* beta_q,
                       unsigned iters, unsigned layers)
{
  if (ctx == nullptr || alpha_q == nullptr || beta_q == nullptr || (base_graph != 1 && base_graph != 2)) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_set_scale_table");
  }
  const unsigned M = base_graph == 1 ? 46 : 42;
  if (iters == 0 || iters > (unsigned)kLearnedIters || layers != M) {
    return fail(LH_ERR_INVALID, "scale table of BG%u must be 1..%d iterations x %u layers", base_graph, kLearnedIters,
                M);
  }
  std::vector<uint32_t> tab((size_t)kLearnedIters * M);
  for (unsigned t = 0; t < (unsigned)kLearnedIters; ++t) {
    const unsigned src = std::min(t, iters - 1);
    for (unsigned l = 0; l < M; ++l) {
      const unsigned a = alpha_q[src * M + l];
      if (a > 511) {
        return fail(LH_ERR_INVALID, "scale table entry (%u, %u): alpha above 511/256", src, l);
      }
      tab[t * M + l] = a | ((uint32_t)(uint16_t)beta_q[src * M + l] << 16);
    }
  }
  (void)hipSetDevice(ctx->device);
  uint32_t* dst = ctx->d_scale + (base_graph == 1 ? 0 : kLearnedIters * 46);
  hipError_t e = hipMemcpy(dst, tab.data(), tab.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    return hip_fail(e, "scale table upload");
  }
  ctx->scale_set[base_graph - 1] = true;
  return LH_OK;
}

int lh_pusch_decode_batch_learned(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw,
                                  int8_t* d_rm, uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                                  unsigned nof_ldpc_iterations, unsigned nof_fallback_iterations, void* stream)
{
  int rc = check_pusch_args(ctx, cbs, n, d_cw, d_rm, d_bits, d_results, "lh_pusch_decode_batch_learned");
  if (rc != LH_OK) {
    return rc;
  }
  launch_plan plan;
  rc = pusch_jobs(cbs, n, new_data, use_early_stop, nof_ldpc_iterations, plan);
  if (rc != LH_OK || n == 0) {
    return rc;
  }
  if (nof_fallback_iterations > 255) {
    return fail(LH_ERR_INVALID, "invalid number of fallback iterations %u", nof_fallback_iterations);
  }
  for (unsigned i = 0; i < n; ++i) {
    if (!ctx->scale_set[cbs[i].rm.base_graph - 1]) {
      return fail(LH_ERR_INVALID, "no learned scale table for BG%u", cbs[i].rm.base_graph);
    }
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  return run_plan(ctx, plan, d_rm, d_cw, d_bits, d_results, st, s, 0, ctx->d_scale, nof_fallback_iterations);
}

} // extern "C"

struct lh_slot_graph {
  lh_ctx* ctx = nullptr;
  uint8_t* d_block = nullptr; // decoder descriptors (resident)
  uint8_t* d_tbs = nullptr;   // TB descriptors (resident)
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int validate_tbs(const lh_tb_desc* tbs, unsigned n)
{
  for (unsigned i = 0; i < n; ++i) {
    const lh_tb_desc& d = tbs[i];
    if (d.tbs_bits == 0 || d.tbs_bits % 8 != 0 || d.nof_cbs == 0 || d.tbs_bits / 8 >= (1u << (4 * kTbShiftDigits))) {
      return fail(LH_ERR_INVALID, "tb %u: invalid size", i);
    }
    if (d.nof_cbs > 1 && (d.cb_data_bits == 0 || (uint64_t)d.cb_data_bits * d.nof_cbs < (uint64_t)d.tbs_bits + 24 ||
                          (uint64_t)d.cb_data_bits * (d.nof_cbs - 1) > d.tbs_bits)) {
      return fail(LH_ERR_INVALID, "tb %u: codeblock data bits do not cover the TB and its CRC", i);
    }
  }
  return LH_OK;
}

} // namespace

extern "C" {

int lh_slot_graph_build(lh_ctx* ctx, const lh_pusch_cb_desc* cbs, unsigned n, const int8_t* d_cw, int8_t* d_rm,
                        uint8_t* d_bits, int32_t* d_results, int new_data, int use_early_stop,
                        unsigned nof_ldpc_iterations, int learned, unsigned nof_fallback_iterations,
                        const lh_tb_desc* tbs, unsigned n_tbs, uint8_t* d_tb, int32_t* d_tb_status,
                        lh_slot_graph** out)
{
  if (out == nullptr) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_slot_graph_build");
  }
  *out = nullptr;
  int rc = check_pusch_args(ctx, cbs, n, d_cw, d_rm, d_bits, d_results, "lh_slot_graph_build");
  if (rc != LH_OK) {
    return rc;
  }
  if (n == 0 || (n_tbs > 0 && (tbs == nullptr || d_tb == nullptr || d_tb_status == nullptr)) || n_tbs > ctx->max_cb ||
      nof_fallback_iterations > 255) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_slot_graph_build");
  }
  rc = validate_tbs(tbs, n_tbs);
  if (rc != LH_OK) {
    return rc;
  }
  if (learned) {
    for (unsigned i = 0; i < n; ++i) {
      if (!ctx->scale_set[cbs[i].rm.base_graph - 1]) {
        return fail(LH_ERR_INVALID, "no learned scale table for BG%u", cbs[i].rm.base_graph);
      }
    }
  }
  launch_plan plan;
  rc = pusch_jobs(cbs, n, new_data, use_early_stop, nof_ldpc_iterations, plan);
  if (rc != LH_OK) {
    return rc;
  }
  plan_block pb;
  rc = build_block(ctx, plan, pb);
  if (rc != LH_OK) {
    return rc;
  }
  (void)hipSetDevice(ctx->device);
  lh_slot_graph* g = new lh_slot_graph();
  g->ctx = ctx;
  hipError_t e = hipMalloc(&g->d_block, pb.size);
  if (e == hipSuccess) {
    e = hipMemcpy(g->d_block, pb.bytes.data(), pb.size, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && n_tbs > 0) {
    e = hipMalloc(&g->d_tbs, n_tbs * sizeof(lh_tb_desc));
    if (e == hipSuccess) {
      e = hipMemcpy(g->d_tbs, tbs, n_tbs * sizeof(lh_tb_desc), hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess) {
    e = hipStreamCreate(&g->cap);
  }
  // Graph: rewind the sub-queue counters, decode (+ fallback), reassemble the TBs.
  if (e == hipSuccess) {
    e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeRelaxed);
  }
  if (e == hipSuccess) {
    // Every replay starts with zero sub-queue counters (a small kernel, no memset node).
    hipError_t le = pb.ctr_words == 0
                        ? hipSuccess
                        : launch_reset_counters(reinterpret_cast<uint32_t*>(g->d_block + 4 * sizeof(launch_header)),
                                                (unsigned)pb.ctr_words, g->cap);
    if (le == hipSuccess) {
      le = launch_block(ctx, pb, g->d_block, d_rm, d_cw, d_bits, d_results, g->cap, learned ? ctx->d_scale : nullptr,
                        learned ? nof_fallback_iterations : 0);
    }
    if (le == hipSuccess && n_tbs > 0) {
      le = launch_tb_finalize(reinterpret_cast<const lh_tb_desc_dev*>(g->d_tbs), n_tbs, d_bits, d_results, d_tb,
                              d_tb_status, ctx->d_tbcrc, ctx->d_tbcrc + 256, g->cap);
    }
    e = hipStreamEndCapture(g->cap, &g->graph);
    if (le != hipSuccess) {
      e = le;
    }
  }
  if (e == hipSuccess) {
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  }
  if (e != hipSuccess) {
    lh_slot_graph_destroy(g);
    return hip_fail(e, "slot graph build");
  }
  *out = g;
  return LH_OK;
}

int lh_slot_graph_launch(lh_slot_graph* g, void* stream)
{
  if (g == nullptr || g->exec == nullptr) {
    return fail(LH_ERR_INVALID, "invalid slot graph");
  }
  (void)hipSetDevice(g->ctx->device);
  hipError_t e = hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream));
  return e == hipSuccess ? LH_OK : hip_fail(e, "slot graph launch");
}

int lh_slot_graph_destroy(lh_slot_graph* g)
{
  if (g == nullptr) {
    return LH_OK;
  }
  (void)hipSetDevice(g->ctx->device);
  if (g->cap) {
    (void)hipStreamSynchronize(g->cap);
  }
  (void)hipDeviceSynchronize(); // a replay may still read the descriptors
  if (g->exec) {
    (void)hipGraphExecDestroy(g->exec);
  }
  if (g->graph) {
    (void)hipGraphDestroy(g->graph);
  }
  if (g->cap) {
    (void)hipStreamDestroy(g->cap);
  }
  if (g->d_block) {
    (void)hipFree(g->d_block);
  }
  if (g->d_tbs) {
    (void)hipFree(g->d_tbs);
  }
  delete g;
  return LH_OK;
}

int lh_tune_set(lh_ctx* ctx, unsigned base_graph, unsigned lifting_size, unsigned groups, int stage_llrs)
{
  if (ctx == nullptr || (base_graph != 1 && base_graph != 2) || lifting_position(lifting_size) < 0 ||
      groups > (unsigned)kMaxGroups) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_tune_set");
  }
  tune_entry& t = ctx->tune[base_graph - 1][lifting_position(lifting_size)];
  t.groups = (uint16_t)groups;
  t.stage = stage_llrs ? 1 : 0;
  return LH_OK;
}

int lh_tune_set_rows(lh_ctx* ctx, unsigned base_graph, unsigned lifting_size, int row_pairs)
{
  if (ctx == nullptr || (base_graph != 1 && base_graph != 2) || lifting_position(lifting_size) < 0 ||
      row_pairs < -1 || row_pairs > 1 || (row_pairs == 1 && (lifting_size & 1))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_tune_set_rows");
  }
  ctx->tune[base_graph - 1][lifting_position(lifting_size)].pairs = (int8_t)row_pairs;
  return LH_OK;
}

int lh_tb_finalize(lh_ctx* ctx, const lh_tb_desc* tbs, unsigned n, const uint8_t* d_cb_bits, const int32_t* d_cb_results,
                   uint8_t* d_tb, int32_t* d_tb_status, void* stream)
{
  static_assert(sizeof(lh_tb_desc) == sizeof(lh_tb_desc_dev), "TB descriptor layout");
  if (ctx == nullptr ||
      (n > 0 && (tbs == nullptr || d_cb_bits == nullptr || d_cb_results == nullptr || d_tb == nullptr ||
                 d_tb_status == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_tb_finalize");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u transport blocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  int rc0 = validate_tbs(tbs, n);
  if (rc0 != LH_OK) {
    return rc0;
  }
  if (n == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  std::memcpy(s->host, tbs, n * sizeof(lh_tb_desc));
  int rc = upload(ctx, s, n * sizeof(lh_tb_desc), st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_tb_finalize(reinterpret_cast<const lh_tb_desc_dev*>(s->dev), n, d_cb_bits, d_cb_results, d_tb,
                                    d_tb_status, ctx->d_tbcrc, ctx->d_tbcrc + 256, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "TB finalize launch");
}

int lh_gen_codeblocks(lh_ctx* ctx, const lh_gen_desc* descs, unsigned n, int8_t* d_llrs, uint8_t* d_msgs, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_llrs == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_gen_codeblocks");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  gen_job* jobs = reinterpret_cast<gen_job*>(s->host);
  for (unsigned i = 0; i < n; ++i) {
    const lh_gen_desc& d = descs[i];
    const int ci = crc_index(d.crc_poly);
    if (ci < 0) {
      return fail(LH_ERR_INVALID, "cb %u: unknown CRC polynomial 0x%x", i, d.crc_poly);
    }
    if (d.modulation_order != 2 && d.modulation_order != 4 && d.modulation_order != 6 && d.modulation_order != 8) {
      return fail(LH_ERR_INVALID, "cb %u: the generator maps QPSK to 256QAM, not order %u", i, d.modulation_order);
    }
    if (!(d.n0 >= 0.0)) {
      return fail(LH_ERR_INVALID, "cb %u: invalid noise variance", i);
    }
    lh_rm_desc rd{};
    rd.cw_offset = d.llr_offset;
    rd.rm_length = d.rm_length;
    rd.base_graph = d.base_graph;
    rd.lifting_size = d.lifting_size;
    rd.rv = d.rv;
    rd.modulation_order = d.modulation_order;
    rd.Nref = d.Nref;
    rd.nof_filler_bits = d.nof_filler_bits;
    gen_job& j = jobs[i];
    j = gen_job{};
    uint64_t mp = 0;
    int rc = make_rm_params(rd, true, j.rm, mp, i);
    if (rc != LH_OK) {
      return rc;
    }
    const bg_info b = bginfo(d.base_graph);
    const unsigned order = ci == 0 ? 0 : (ci == 3 ? 16 : 24);
    if ((uint64_t)d.nof_filler_bits + order >= (uint64_t)b.K * d.lifting_size) {
      return fail(LH_ERR_INVALID, "cb %u: no data bits left", i);
    }
    j.msg_off = d.msg_offset;
    j.seed_msg = d.seed_msg;
    j.seed_noise = d.seed_noise;
    j.n0 = d.n0;
    j.gt_off = graph_offset(d.base_graph, d.lifting_size);
    j.Z = (uint16_t)d.lifting_size;
    j.bg = (uint8_t)d.base_graph;
    j.crc_idx = (uint8_t)ci;
    j.F = (uint16_t)d.nof_filler_bits;
    j.write_msg = d_msgs != nullptr ? 1 : 0;
    j.msg_input = (d.flags & LH_GEN_MSG_INPUT) ? 1 : 0;
    if (j.msg_input && d_msgs == nullptr) {
      return fail(LH_ERR_INVALID, "cb %u: message input without a message buffer", i);
    }
  }
  int rc = upload(ctx, s, n * sizeof(gen_job), st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_gen_codeblocks(reinterpret_cast<const gen_job*>(s->dev), n, d_llrs, d_msgs, ctx->d_crc_x,
                                       ctx->d_gtab, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "generator launch");
}

int lh_ldpc_encode_rate_match(lh_ctx* ctx, const lh_encode_desc* descs, unsigned n, const uint8_t* d_msgs,
                              uint8_t* d_out, uint8_t* d_codewords, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_msgs == nullptr || d_out == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ldpc_encode_rate_match");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  gen_job* jobs = reinterpret_cast<gen_job*>(s->host);
  for (unsigned i = 0; i < n; ++i) {
    const lh_encode_desc& d = descs[i];
    lh_rm_desc rd{};
    rd.cw_offset = d.out_offset;
    rd.rm_length = d.rm_length;
    rd.base_graph = d.base_graph;
    rd.lifting_size = d.lifting_size;
    rd.rv = d.rv;
    rd.modulation_order = d.modulation_order;
    rd.Nref = d.Nref;
    rd.nof_filler_bits = d.nof_filler_bits;
    gen_job& j = jobs[i];
    j = gen_job{};
    uint64_t mp = 0;
    int rc = make_rm_params(rd, true, j.rm, mp, i);
    if (rc != LH_OK) {
      return rc;
    }
    j.msg_off = d.msg_offset;
    j.gt_off = graph_offset(d.base_graph, d.lifting_size);
    j.Z = (uint16_t)d.lifting_size;
    j.bg = (uint8_t)d.base_graph;
    j.F = (uint16_t)d.nof_filler_bits;
    j.msg_input = 1;
    j.want_cw = d_codewords != nullptr ? 1 : 0;
    j.cwb_off = d.cw_offset;
  }
  int rc = upload(ctx, s, n * sizeof(gen_job), st);
  if (rc != LH_OK) {
    return rc;
  }
  // Lifting sizes below 32 take the bit-per-byte encoder; the others the word-parallel one,
  // whose LDS holds the codeword columns its codeblocks need.
  bool any_small = false;
  size_t cw_bytes = 0;
  for (unsigned i = 0; i < n; ++i) {
    const gen_job& j = jobs[i];
    if (j.Z < 32) {
      any_small = true;
      continue;
    }
    const int K = j.bg == 1 ? 22 : 10, M = j.bg == 1 ? 46 : 42;
    cw_bytes = std::max(cw_bytes, (size_t)(K + encode_parity_rows(j.rm, K, M, j.Z, j.want_cw != 0)) * j.Z);
  }
  hipError_t e = launch_encode_rate_match(reinterpret_cast<const gen_job*>(s->dev), n, const_cast<uint8_t*>(d_msgs),
                                          d_out, d_codewords, ctx->d_crc_x, ctx->d_gtab, any_small, cw_bytes, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "encoder launch");
}

int lh_pusch_demodulate(lh_ctx* ctx, const lh_demod_desc* descs, unsigned n, const float* d_symbols,
                        const float* d_noise_vars, int8_t* d_llrs, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_symbols == nullptr || d_noise_vars == nullptr ||
                                   d_llrs == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_pusch_demodulate");
  }
  size_t nof_tiles = 0;
  for (unsigned i = 0; i < n; ++i) {
    const lh_demod_desc& d = descs[i];
    const uint32_t m = d.modulation;
    if (m != 0 && m != 1 && m != 2 && m != 4 && m != 6 && m != 8) {
      return fail(LH_ERR_INVALID, "block %u: modulation %u is not 0, 1, 2, 4, 6 or 8", i, m);
    }
    const uint64_t bits = (uint64_t)d.nof_symbols * (m == 0 ? 1 : m);
    if ((uint64_t)d.seq_offset + bits > (1ull << 25)) {
      return fail(LH_ERR_INVALID, "block %u: sequence positions beyond 2^25", i);
    }
    nof_tiles += (d.nof_symbols + 255) / 256;
  }
  const size_t bytes = (size_t)n * sizeof(demod_job) + nof_tiles * sizeof(demod_tile);
  if (bytes > ctx->slot_bytes) {
    return fail(LH_ERR_CAPACITY, "%u blocks of %zu tiles exceed the context's staging (%zu bytes)", n, nof_tiles,
                ctx->slot_bytes);
  }
  if (nof_tiles == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  demod_job* jobs = reinterpret_cast<demod_job*>(s->host);
  demod_tile* tiles = reinterpret_cast<demod_tile*>(s->host + (size_t)n * sizeof(demod_job));
  size_t t = 0;
  for (unsigned i = 0; i < n; ++i) {
    const lh_demod_desc& d = descs[i];
    demod_job& j = jobs[i];
    j.sym_off = d.symbol_offset;
    j.llr_off = d.llr_offset;
    j.nof_symbols = d.nof_symbols;
    j.modulation = d.modulation;
    const uint32_t nt = (d.nof_symbols + 255) / 256;
    std::vector<uint32_t> s1(nt), s2(nt);
    demod_tile_states(d.c_init, d.seq_offset, d.modulation == 0 ? 1 : (int)d.modulation, nt, s1.data(), s2.data());
    for (uint32_t k = 0; k < nt; ++k) {
      tiles[t++] = demod_tile{i, k, s1[k], s2[k]};
    }
  }
  int rc = upload(ctx, s, bytes, st);
  if (rc != LH_OK) {
    return rc;
  }
  const uint8_t* dv = reinterpret_cast<const uint8_t*>(s->dev);
  hipError_t e = launch_demod(reinterpret_cast<const demod_tile*>(dv + (size_t)n * sizeof(demod_job)),
                              (unsigned)nof_tiles, reinterpret_cast<const demod_job*>(dv), ctx->d_demod, d_symbols,
                              d_noise_vars, d_llrs, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "demodulator launch");
}

int lh_ulsch_information(const lh_ulsch_config* cfg, lh_ulsch_info* info)
{
  if (cfg == nullptr || info == nullptr) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ulsch_information");
  }
  const char* why = nullptr;
  if (ulsch_information(*cfg, *info, &why) != 0) {
    return fail(LH_ERR_INVALID, "UL-SCH information: %s", why);
  }
  return LH_OK;
}

int lh_ulsch_demux_get_sizes(const lh_ulsch_demux_desc* desc, lh_ulsch_demux_sizes* sizes)
{
  if (desc == nullptr || sizes == nullptr) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ulsch_demux_get_sizes");
  }
  const char* why = nullptr;
  if (demux_plan(*desc, nullptr, nullptr, sizes, &why) != 0) {
    return fail(LH_ERR_INVALID, "UL-SCH demultiplexer: %s", why);
  }
  return LH_OK;
}

int lh_ulsch_demultiplex(lh_ctx* ctx, const lh_ulsch_demux_desc* descs, unsigned n, const int8_t* d_llrs,
                         int8_t* d_sch, int8_t* d_harq_ack, int8_t* d_csi_part1, int8_t* d_csi_part2, void* stream)
{
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_llrs == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_ulsch_demultiplex");
  }
  std::vector<demux_sym> syms((size_t)n * 14);
  size_t nof_syms = 0, nof_tiles = 0;
  for (unsigned i = 0; i < n; ++i) {
    const char* why = nullptr;
    unsigned ns = 0;
    lh_ulsch_demux_sizes sz;
    if (demux_plan(descs[i], syms.data() + nof_syms, &ns, &sz, &why) != 0) {
      return fail(LH_ERR_INVALID, "codeword %u: %s", i, why);
    }
    if ((sz.nof_sch && d_sch == nullptr) || (sz.nof_harq_ack && d_harq_ack == nullptr) ||
        (sz.nof_csi_part1 && d_csi_part1 == nullptr) || (sz.nof_csi_part2 && d_csi_part2 == nullptr)) {
      return fail(LH_ERR_INVALID, "codeword %u: an output it writes is NULL", i);
    }
    for (unsigned k = 0; k < ns; ++k) {
      const demux_sym& s = syms[nof_syms + k];
      const uint32_t units = s.M * s.nbpr / s.unit;
      nof_tiles += (units + kDemuxTileUnits - 1) / kDemuxTileUnits;
    }
    nof_syms += ns;
  }
  const size_t sym_bytes = (nof_syms * sizeof(demux_sym) + 15) & ~(size_t)15;
  const size_t bytes = sym_bytes + nof_tiles * sizeof(demux_tile);
  if (bytes > ctx->slot_bytes) {
    return fail(LH_ERR_CAPACITY, "%u codewords of %zu tiles exceed the context's staging (%zu bytes)", n, nof_tiles,
                ctx->slot_bytes);
  }
  if (nof_tiles == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  std::memcpy(s->host, syms.data(), nof_syms * sizeof(demux_sym));
  demux_tile* tiles = reinterpret_cast<demux_tile*>(s->host + sym_bytes);
  size_t t = 0;
  for (size_t k = 0; k < nof_syms; ++k) {
    const uint32_t units = syms[k].M * syms[k].nbpr / syms[k].unit;
    for (uint32_t j = 0; j * kDemuxTileUnits < units; ++j) {
      tiles[t++] = demux_tile{(uint32_t)k, j};
    }
  }
  int rc = upload(ctx, s, bytes, st);
  if (rc != LH_OK) {
    return rc;
  }
  const uint8_t* dv = reinterpret_cast<const uint8_t*>(s->dev);
  hipError_t e = launch_demux(reinterpret_cast<const demux_tile*>(dv + sym_bytes), (unsigned)nof_tiles,
                              reinterpret_cast<const demux_sym*>(dv), ctx->d_demod, d_llrs, d_sch, d_harq_ack,
                              d_csi_part1, d_csi_part2, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "demultiplexer launch");
}

} // extern "C"

namespace {

// Workspace floats of one training codeblock: hist[T][E_L][Z] + gl[T][K Z] + gnext[E_L][Z],
// rounded up to 64 floats.
size_t train_ws_floats(unsigned bg, unsigned Z, unsigned L, unsigned T)
{
  const uint16_t* rs = bg == 1 ? lh_bg1_row_start : lh_bg2_row_start;
  const size_t EL = rs[L], K = bg == 1 ? 22 : 10;
  const size_t f = (size_t)Z * (T * EL + T * K + EL);
  return (f + 63) & ~(size_t)63;
}

int check_train_desc(unsigned bg, const lh_train_desc& d, unsigned i)
{
  if (lifting_index(d.lifting_size) < 0) {
    return fail(LH_ERR_INVALID, "cb %u: invalid lifting size %u", i, d.lifting_size);
  }
  const bg_info b = bginfo(bg);
  if (d.nof_layers < 4 || d.nof_layers > (unsigned)b.M) {
    return fail(LH_ERR_INVALID, "cb %u: invalid number of layers %u", i, d.nof_layers);
  }
  if (d.nof_llrs > (unsigned)b.nshort * d.lifting_size) {
    return fail(LH_ERR_INVALID, "cb %u: %u soft bits exceed the codeword", i, d.nof_llrs);
  }
  if (d.nof_filler_bits >= (unsigned)(b.K - 2) * d.lifting_size) {
    return fail(LH_ERR_INVALID, "cb %u: invalid number of filler bits", i);
  }
  return LH_OK;
}

} // namespace

extern "C" {

size_t lh_nms_workspace_bytes(unsigned base_graph, const lh_train_desc* descs, unsigned n, unsigned iters)
{
  if ((base_graph != 1 && base_graph != 2) || descs == nullptr) {
    return 0;
  }
  size_t f = 0;
  for (unsigned i = 0; i < n; ++i) {
    f += train_ws_floats(base_graph, descs[i].lifting_size, descs[i].nof_layers, iters);
  }
  return 4 * f;
}

int lh_nms_train_step(lh_ctx* ctx, unsigned base_graph, const lh_train_desc* descs, unsigned n, unsigned iters,
                      const float* d_alpha, const float* d_beta, const float* w, float tau, const int8_t* d_llrs,
                      const uint8_t* d_msgs, void* d_workspace, size_t workspace_bytes, float* d_loss,
                      float* d_grad_alpha, float* d_grad_beta, int backward, void* stream)
{
  if (ctx == nullptr || (base_graph != 1 && base_graph != 2) || w == nullptr ||
      (n > 0 && (descs == nullptr || d_alpha == nullptr || d_beta == nullptr || d_llrs == nullptr ||
                 d_msgs == nullptr || d_workspace == nullptr || d_loss == nullptr)) ||
      (backward && (d_grad_alpha == nullptr || d_grad_beta == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_nms_train_step");
  }
  if (iters == 0 || iters > (unsigned)kLearnedIters) {
    return fail(LH_ERR_INVALID, "invalid number of iterations %u (1..%d)", iters, kLearnedIters);
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u codeblocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  if (n == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  train_job* jobs = reinterpret_cast<train_job*>(s->host);
  size_t ws = 0;
  int maxZ = 0;
  size_t lds = 0;
  for (unsigned i = 0; i < n; ++i) {
    const lh_train_desc& d = descs[i];
    int rc = check_train_desc(base_graph, d, i);
    if (rc != LH_OK) {
      return rc;
    }
    train_job& j = jobs[i];
    j = train_job{};
    j.in_off = d.llr_offset;
    j.tgt_off = d.msg_offset;
    j.ws_off = ws;
    j.n_in = d.nof_llrs;
    j.gt_off = graph_offset(base_graph, d.lifting_size);
    j.Z = (uint16_t)d.lifting_size;
    j.L = (uint16_t)d.nof_layers;
    j.F = (uint16_t)d.nof_filler_bits;
    ws += train_ws_floats(base_graph, d.lifting_size, d.nof_layers, iters);
    maxZ = std::max<int>(maxZ, (int)d.lifting_size);
    lds = std::max(lds, nms_train_lds_bytes((int)base_graph, (int)d.lifting_size, (int)d.nof_layers));
  }
  if (4 * ws > workspace_bytes) {
    return fail(LH_ERR_CAPACITY, "workspace of %zu bytes, %zu needed", workspace_bytes, 4 * ws);
  }
  const size_t o_w = ((size_t)n * sizeof(train_job) + 15) & ~(size_t)15;
  std::memcpy(s->host + o_w, w, iters * sizeof(float));
  int rc = upload(ctx, s, o_w + iters * sizeof(float), st);
  if (rc != LH_OK) {
    return rc;
  }
  const int nt = std::max(64, ((maxZ + 63) / 64) * 64);
  hipError_t e = launch_nms_train((int)base_graph, reinterpret_cast<const train_job*>(s->dev), n, nt, lds, (int)iters,
                                  d_alpha, d_beta, reinterpret_cast<const float*>(s->dev + o_w), tau, d_llrs, d_msgs,
                                  static_cast<float*>(d_workspace), d_loss, d_grad_alpha, d_grad_beta, ctx->d_gtab,
                                  backward, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "training launch");
}

int lh_nms_adam(lh_ctx* ctx, float* d_params, const float* d_grads, float* d_m, float* d_v, unsigned n,
                unsigned n_alpha, float lr, float beta1, float beta2, float eps, unsigned step, void* stream)
{
  if (ctx == nullptr || d_params == nullptr || d_grads == nullptr || d_m == nullptr || d_v == nullptr || step == 0 ||
      n_alpha > n) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_nms_adam");
  }
  (void)hipSetDevice(ctx->device);
  const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
  hipError_t e = launch_adam(d_params, d_grads, d_m, d_v, (int)n, (int)n_alpha, lr, beta1, beta2, eps, c1, c2,
                             reinterpret_cast<hipStream_t>(stream));
  return e == hipSuccess ? LH_OK : hip_fail(e, "adam launch");
}

int lh_gen_transport_blocks(lh_ctx* ctx, const lh_gen_tb_desc* descs, unsigned n, uint8_t* d_tb, uint8_t* d_msgs,
                            void* stream)
{
  static_assert(sizeof(lh_gen_tb_desc) == sizeof(gen_tb_job), "TB generator descriptor layout");
  if (ctx == nullptr || (n > 0 && (descs == nullptr || d_tb == nullptr || d_msgs == nullptr))) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_gen_transport_blocks");
  }
  if (n > ctx->max_cb) {
    return fail(LH_ERR_CAPACITY, "batch of %u transport blocks exceeds the context capacity %u", n, ctx->max_cb);
  }
  for (unsigned i = 0; i < n; ++i) {
    const lh_gen_tb_desc& d = descs[i];
    const unsigned C = d.nof_cbs, K = d.segment_length, F = d.nof_filler_bits;
    const unsigned L = d.tbs_bits <= 3824 ? 16 : 24;
    const bool kok = (K % 22 == 0 && lifting_position(K / 22) >= 0) || (K % 10 == 0 && lifting_position(K / 10) >= 0);
    if (d.tbs_bits == 0 || d.tbs_bits % 8 != 0 || C == 0 || !kok || F >= K || d.msg_stride < (K + 7) / 8) {
      return fail(LH_ERR_INVALID, "tb %u: invalid generator descriptor", i);
    }
    const uint64_t D = C > 1 ? (uint64_t)K - F - 24 : (uint64_t)K - F;
    if (K - F <= (C > 1 ? 24u : 0u) || D * C < (uint64_t)d.tbs_bits + L || (C == 1 && D != (uint64_t)d.tbs_bits + L)) {
      return fail(LH_ERR_INVALID, "tb %u: segmentation does not hold the TB and its CRC", i);
    }
  }
  if (n == 0) {
    return LH_OK;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  std::memcpy(s->host, descs, n * sizeof(lh_gen_tb_desc));
  int rc = upload(ctx, s, n * sizeof(lh_gen_tb_desc), st);
  if (rc != LH_OK) {
    return rc;
  }
  hipError_t e = launch_gen_tb(reinterpret_cast<const gen_tb_job*>(s->dev), n, d_tb, d_msgs, ctx->d_crc_x,
                               ctx->d_tbcrc, ctx->d_tbcrc + 256, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "TB generator launch");
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
  (void)hipSetDevice(ctx->device);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    return LH_ERR_HIP;
  }
  std::memcpy(s->host, offsets, n * 8);
  std::memcpy(s->host + n * 8, nbits, n * 4);
  int rc = upload(ctx, s, n * 12, st);
  if (rc != LH_OK) {
    return rc;
  }
  const int order = ci == 3 ? 16 : 24;
  hipError_t e = launch_crc(reinterpret_cast<const uint64_t*>(s->dev), reinterpret_cast<const uint32_t*>(s->dev + n * 8),
                            n, d_bits, ctx->d_crc_x + (ci - 1) * kCrcXLen, order, d_out, st);
  retire(s, st);
  return e == hipSuccess ? LH_OK : hip_fail(e, "CRC launch");
}

} // extern "C"

// ===================================================================================
// Slot-deadline serving runtime
// ===================================================================================

namespace {

double now_us()
{
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A HARQ process on the device: the codeblocks' soft bits (N_short*Z each), then their
// decoded message bits (one message per codeblock), and per codeblock the iteration count
// of the transmission whose CRC passed, or 0 (rx_buffer's codeblock CRC flags and data
// bits, rx_buffer.h:29-63).
struct harq_entry {
  uint64_t off = 0, bytes = 0;
  uint64_t msg_off = 0;
  uint32_t nof_cbs = 0, Z = 0, bg = 0, tbs = 0;
  std::vector<uint8_t> cb_iters;
  bool busy = false;
  uint32_t expiry = 0; // slot at which it expires when not in use (expire_timeout_slots > 0)
};

uint64_t harq_bytes(uint32_t bg, uint32_t Z, uint32_t nof_cbs, uint64_t* msg_at = nullptr)
{
  const uint64_t soft = (uint64_t)nof_cbs * (bg == 1 ? 66u : 50u) * Z;
  const uint64_t mb = ((bg == 1 ? 22u : 10u) * Z + 7) / 8;
  if (msg_at != nullptr) {
    *msg_at = soft;
  }
  return soft + (uint64_t)nof_cbs * mb;
}

struct srv_request {
  lh_tb_request req{};
  lh_segmentation seg{};
  std::vector<uint32_t> rm, off;
  uint64_t harq_off = 0;
  double t_submit = 0;
};

// A launch's results (TB verdicts | codeblock results | TB bytes) are written by the TB
// finalize kernel straight into one pinned, device-mapped host block (offsets per launch);
// the kernel's last workgroup then stores the launch's sequence number into `h_flag`, which
// the host polls.  The decoder's codeblock results stay in device memory (d_res).
struct srv_slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int32_t* d_res = nullptr;
  uint32_t* d_counter = nullptr;
  uint8_t* h_out = nullptr;
  uint8_t* h_out_dev = nullptr; // device address of h_out
  uint32_t* h_flag = nullptr;
  uint32_t* h_flag_dev = nullptr;
  uint32_t seq = 0;
  int32_t* h_res = nullptr;
  int32_t* h_st = nullptr;
  uint8_t* h_tb = nullptr;
  std::vector<srv_request> reqs;
  std::vector<uint32_t> first_cb, tb_off;
  bool busy = false;
};

// First-fit allocator over the HARQ arena (offsets), coalescing on free.
struct arena_alloc {
  std::map<uint64_t, uint64_t> free_; // offset -> bytes
  explicit arena_alloc(uint64_t bytes = 0) { if (bytes) free_[0] = bytes; }
  bool alloc(uint64_t bytes, uint64_t& off)
  {
    bytes = (bytes + 255) & ~(uint64_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        const uint64_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) {
          free_[off + bytes] = rest;
        }
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes)
  {
    bytes = (bytes + 255) & ~(uint64_t)255;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t free_bytes() const
  {
    uint64_t f = 0;
    for (auto& kv : free_) {
      f += kv.second;
    }
    return f;
  }
};

constexpr unsigned kMaxMsgBytes = (22 * 384) / 8;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Result block of a launch of at most ncb codeblocks (and as many TBs).
size_t srv_out_bytes(size_t ncb)
{
  return 2 * align256(ncb * 4) + ncb * kMaxMsgBytes;
}

} // namespace

struct lh_server {
  lh_ctx* ctx = nullptr;
  lh_server_config cfg{};
  int8_t* d_arena = nullptr;
  arena_alloc arena;
  std::unordered_map<uint32_t, harq_entry> harq;
  std::deque<srv_request> pending;
  std::deque<srv_request> ready;
  std::vector<srv_slot> slots;
  std::deque<lh_tb_result> results;
  std::unordered_map<uint64_t, uint8_t*> tb_dest; // host TB destinations by request id
  unsigned next_slot = 0;
  uint32_t cur_slot = 0; // last lh_server_run_slot slot
  bool have_slot = false;
};

namespace {

bool harq_matches(const harq_entry& h, const srv_request& r)
{
  return h.nof_cbs == r.seg.nof_cbs && h.Z == r.seg.lifting_size && h.bg == r.req.base_graph &&
         h.tbs == r.req.tbs_bits;
}

void push_result(lh_server* sv, const srv_request& r, int32_t status, unsigned ok, unsigned iters)
{
  lh_tb_result res{};
  res.id = r.req.id;
  res.harq_process = r.req.harq_process;
  res.tb_crc_ok = status;
  res.nof_cbs = r.seg.nof_cbs;
  res.nof_cbs_ok = ok;
  res.max_iterations = iters;
  res.latency_us = now_us() - r.t_submit;
  sv->results.push_back(res);
}

// Admits pending requests whose HARQ buffers can be placed (FIFO; a request whose process
// is in flight waits, later requests may pass it).
void admit(lh_server* sv)
{
  for (auto it = sv->pending.begin(); it != sv->pending.end();) {
    srv_request& r = *it;
    auto h = sv->harq.find(r.req.harq_process);
    uint64_t msg_at = 0;
    const uint64_t bytes = harq_bytes(r.req.base_graph, r.seg.lifting_size, r.seg.nof_cbs, &msg_at);
    if (h != sv->harq.end() && h->second.busy) {
      ++it;
      continue;
    }
    if (!r.req.new_data) {
      if (h == sv->harq.end() || !harq_matches(h->second, r)) {
        push_result(sv, r, LH_TB_REJECTED, 0, 0);
        it = sv->pending.erase(it);
        continue;
      }
    } else if (h != sv->harq.end() && !harq_matches(h->second, r)) {
      sv->arena.release(h->second.off, h->second.bytes);
      sv->harq.erase(h);
      h = sv->harq.end();
    }
    if (h == sv->harq.end()) {
      uint64_t off = 0;
      if (!sv->arena.alloc(bytes, off)) {
        ++it; // waits for memory
        continue;
      }
      harq_entry e;
      e.off = off;
      e.bytes = bytes;
      e.msg_off = off + msg_at;
      e.cb_iters.assign(r.seg.nof_cbs, 0);
      e.nof_cbs = r.seg.nof_cbs;
      e.Z = r.seg.lifting_size;
      e.bg = r.req.base_graph;
      e.tbs = r.req.tbs_bits;
      h = sv->harq.emplace(r.req.harq_process, e).first;
    }
    h->second.busy = true;
    // Reserved now (rx_buffer_pool_impl.cpp:79-80): at the request's slot, or at the current
    // slot when the request waited for memory past it.
    const uint32_t at = (sv->have_slot && (int32_t)(sv->cur_slot - r.req.slot) > 0) ? sv->cur_slot : r.req.slot;
    h->second.expiry = at + sv->cfg.expire_timeout_slots;
    r.harq_off = h->second.off;
    sv->ready.push_back(std::move(r));
    it = sv->pending.erase(it);
  }
}

// Retires a completed slot: results, HARQ release on TB CRC pass, TB bytes to the caller.
void retire_slot(lh_server* sv, srv_slot& sl)
{
  for (size_t k = 0; k < sl.reqs.size(); ++k) {
    const srv_request& r = sl.reqs[k];
    const int32_t st = sl.h_st[k];
    unsigned ok = 0, its = 0;
    for (uint32_t c = 0; c < r.seg.nof_cbs; ++c) {
      const int32_t x = sl.h_res[sl.first_cb[k] + c];
      ok += x > 0 ? 1u : 0u;
      its = std::max<unsigned>(its, x > 0 ? (unsigned)x : 0u);
    }
    auto dst = sv->tb_dest.find(r.req.id);
    if (dst != sv->tb_dest.end()) {
      if (st == LH_TB_CRC_OK && dst->second != nullptr) {
        std::memcpy(dst->second, sl.h_tb + sl.tb_off[k], r.req.tbs_bits / 8);
      }
      sv->tb_dest.erase(dst);
    }
    auto h = sv->harq.find(r.req.harq_process);
    if (h != sv->harq.end()) {
      h->second.busy = false;
      if (st == LH_TB_CRC_OK) {
        sv->arena.release(h->second.off, h->second.bytes);
        sv->harq.erase(h);
      } else if (st == LH_TB_CRC_FAILED) {
        // Every codeblock passed but the TB CRC did not: all are decoded again next time
        // (pusch_decoder_impl.cpp:657-669, reset_codeblocks_crc).
        std::fill(h->second.cb_iters.begin(), h->second.cb_iters.end(), (uint8_t)0);
      } else {
        for (uint32_t c = 0; c < r.seg.nof_cbs && c < h->second.cb_iters.size(); ++c) {
          const int32_t x = sl.h_res[sl.first_cb[k] + c];
          if (x > 0) {
            h->second.cb_iters[c] = (uint8_t)std::min(x, 255);
          }
        }
      }
    }
    push_result(sv, r, st, ok, its);
  }
  sl.reqs.clear();
  sl.busy = false;
}

// A slot whose launch failed or was never made: nothing of it will complete.  Its requests
// are reported as rejected and leave no state behind (destinations, busy processes); a new
// transmission's HARQ process is dropped, since its soft bits were never written.
void reject_slot(lh_server* sv, srv_slot& sl)
{
  for (auto& r : sl.reqs) {
    auto h = sv->harq.find(r.req.harq_process);
    if (h != sv->harq.end()) {
      h->second.busy = false;
      if (r.req.new_data) {
        sv->arena.release(h->second.off, h->second.bytes);
        sv->harq.erase(h);
      }
    }
    sv->tb_dest.erase(r.req.id);
    push_result(sv, r, LH_TB_REJECTED, 0, 0);
  }
  sl.reqs.clear();
  sl.busy = false;
}

} // namespace

extern "C" {

int lh_server_create(lh_ctx* ctx, const lh_server_config* cfg, lh_server** out)
{
  if (ctx == nullptr || cfg == nullptr || out == nullptr || cfg->max_cbs_per_slot == 0 ||
      cfg->max_cbs_per_slot > ctx->max_cb || cfg->max_slots_in_flight == 0 || cfg->max_slots_in_flight > 8 ||
      cfg->harq_budget_bytes < align256(harq_bytes(1, 384, 1)) || cfg->nof_ldpc_iterations == 0 || cfg->nof_ldpc_iterations > 255 ||
      cfg->nof_fallback_iterations > 255) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_create");
  }
  (void)hipSetDevice(ctx->device);
  lh_server* sv = new lh_server();
  sv->ctx = ctx;
  sv->cfg = *cfg;
  hipError_t e = hipMalloc(&sv->d_arena, cfg->harq_budget_bytes);
  sv->arena = arena_alloc(cfg->harq_budget_bytes);
  const size_t ncb = cfg->max_cbs_per_slot;
  sv->slots.resize(cfg->max_slots_in_flight);
  for (auto& sl : sv->slots) {
    if (e == hipSuccess) e = hipStreamCreate(&sl.stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&sl.d_res, ncb * 4);
    if (e == hipSuccess) e = hipMalloc(&sl.d_counter, 4);
    if (e == hipSuccess) e = hipMemset(sl.d_counter, 0, 4);
    if (e == hipSuccess) {
      e = hipHostMalloc(&sl.h_out, srv_out_bytes(ncb), hipHostMallocMapped | hipHostMallocCoherent);
    }
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.h_out_dev), sl.h_out, 0);
    if (e == hipSuccess) {
      e = hipHostMalloc(&sl.h_flag, 64, hipHostMallocMapped | hipHostMallocCoherent);
    }
    if (e == hipSuccess) {
      *sl.h_flag = 0;
      e = hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.h_flag_dev), sl.h_flag, 0);
    }
  }
  if (e != hipSuccess) {
    lh_server_destroy(sv);
    return hip_fail(e, "server allocation");
  }
  *out = sv;
  return LH_OK;
}

int lh_server_destroy(lh_server* sv)
{
  if (sv == nullptr) {
    return LH_OK;
  }
  (void)hipSetDevice(sv->ctx->device);
  for (auto& sl : sv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  }
  rebind_ring(sv->ctx); // the staging slots' events may name the server's streams
  for (auto& sl : sv->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.d_res) (void)hipFree(sl.d_res);
    if (sl.d_counter) (void)hipFree(sl.d_counter);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    if (sl.h_flag) (void)hipHostFree(sl.h_flag);
  }
  if (sv->d_arena) (void)hipFree(sv->d_arena);
  delete sv;
  return LH_OK;
}

} // extern "C"

namespace {

int submit_one(lh_server* sv, const lh_tb_request* req, uint8_t* tb_out, double t_submit)
{
  if (req == nullptr || req->d_llrs == nullptr || (req->base_graph != 1 && req->base_graph != 2)) {
    return fail(LH_ERR_INVALID, "invalid transport-block request");
  }
  srv_request r;
  r.req = *req;
  r.rm.resize(512);
  r.off.resize(512);
  int rc = lh_segment_rx(req->base_graph, req->tbs_bits, req->modulation_order, req->nof_layers, req->nof_ch_symbols,
                         &r.seg, r.rm.data(), r.off.data(), 512);
  if (rc != LH_OK) {
    return fail(rc, "request %llu: invalid transport block", (unsigned long long)req->id);
  }
  r.rm.resize(r.seg.nof_cbs);
  r.off.resize(r.seg.nof_cbs);
  // Every codeblock's dematching parameters are checked here, so that a launch never
  // fails on them (rate-matched length, N_ref, RV, modulation order).
  for (uint32_t c = 0; c < r.seg.nof_cbs; ++c) {
    lh_rm_desc d{};
    d.rm_length = r.rm[c];
    d.base_graph = req->base_graph;
    d.lifting_size = r.seg.lifting_size;
    d.rv = req->rv;
    d.modulation_order = req->modulation_order;
    d.Nref = req->Nref;
    d.nof_filler_bits = r.seg.nof_filler_bits;
    rm_params j;
    uint64_t mp = 0;
    rc = make_rm_params(d, req->new_data != 0, j, mp, c);
    if (rc != LH_OK) {
      return fail(LH_ERR_INVALID, "request %llu: %s", (unsigned long long)req->id, lh_last_error());
    }
  }
  // The HARQ process's whole footprint (soft bits and message bits, as the arena places it).
  const uint64_t bytes = align256(harq_bytes(req->base_graph, r.seg.lifting_size, r.seg.nof_cbs));
  if (r.seg.nof_cbs > sv->cfg.max_cbs_per_slot || bytes > sv->cfg.harq_budget_bytes) {
    return fail(LH_ERR_CAPACITY, "request %llu: %u codeblocks (%llu HARQ bytes) cannot be served",
                (unsigned long long)req->id, r.seg.nof_cbs, (unsigned long long)bytes);
  }
  r.t_submit = t_submit;
  sv->tb_dest[req->id] = tb_out;
  sv->pending.push_back(std::move(r));
  return LH_OK;
}

} // namespace

extern "C" {

int lh_server_submit(lh_server* sv, const lh_tb_request* req, uint8_t* tb_out)
{
  if (sv == nullptr) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_submit");
  }
  int rc = submit_one(sv, req, tb_out, now_us());
  admit(sv);
  return rc;
}

int lh_server_submit_batch(lh_server* sv, const lh_tb_request* reqs, unsigned n, uint8_t* const* tb_outs)
{
  if (sv == nullptr || (n > 0 && reqs == nullptr)) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_submit_batch");
  }
  const double t = now_us();
  int rc = LH_OK;
  for (unsigned i = 0; i < n && rc == LH_OK; ++i) {
    rc = submit_one(sv, &reqs[i], tb_outs ? tb_outs[i] : nullptr, t);
  }
  admit(sv);
  return rc;
}

int lh_server_step(lh_server* sv)
{
  if (sv == nullptr) {
    return fail(LH_ERR_INVALID, "invalid server");
  }
  admit(sv);
  if (sv->ready.empty()) {
    return LH_OK;
  }
  // A free slot (launch resources); none free: the work waits for the next step.
  srv_slot* sl = nullptr;
  for (unsigned k = 0; k < sv->slots.size(); ++k) {
    srv_slot& c = sv->slots[(sv->next_slot + k) % sv->slots.size()];
    if (!c.busy) {
      sl = &c;
      sv->next_slot = (unsigned)((&c - sv->slots.data() + 1) % sv->slots.size());
      break;
    }
  }
  if (sl == nullptr) {
    return LH_OK;
  }
  // Continuous batching: everything admitted since the last launch, up to the slot size.
  unsigned ncb = 0;
  while (!sv->ready.empty() && ncb + sv->ready.front().seg.nof_cbs <= sv->cfg.max_cbs_per_slot) {
    ncb += sv->ready.front().seg.nof_cbs;
    sl->reqs.push_back(std::move(sv->ready.front()));
    sv->ready.pop_front();
  }
  const int8_t* cw_base = sl->reqs[0].req.d_llrs;
  for (auto& r : sl->reqs) {
    cw_base = std::min(cw_base, r.req.d_llrs);
  }
  std::vector<lh_pusch_cb_desc> cbs;
  std::vector<uint8_t> nd, passed;
  std::vector<lh_tb_desc> tbs;
  sl->first_cb.clear();
  sl->tb_off.clear();
  uint64_t tbo = 0;
  for (auto& r : sl->reqs) {
    const lh_tb_request& q = r.req;
    const unsigned C = r.seg.nof_cbs, Z = r.seg.lifting_size, N = (q.base_graph == 1 ? 66u : 50u) * Z;
    const unsigned mb = ((q.base_graph == 1 ? 22u : 10u) * Z + 7) / 8;
    const uint32_t poly = C > 1 ? LH_CRC24B : (q.tbs_bits <= 3824 ? LH_CRC16 : LH_CRC24A);
    // The codeblocks' message bits live in the HARQ process (kept across transmissions).
    harq_entry& he = sv->harq.at(q.harq_process);
    if (q.new_data) {
      std::fill(he.cb_iters.begin(), he.cb_iters.end(), (uint8_t)0);
    }
    const uint64_t msg_base = he.msg_off;
    sl->first_cb.push_back((uint32_t)cbs.size());
    sl->tb_off.push_back((uint32_t)tbo);
    lh_tb_desc t{};
    t.tb_offset = tbo;
    t.cb_bits_offset = msg_base;
    t.cb_bits_stride = mb;
    t.tbs_bits = q.tbs_bits;
    t.first_cb = (uint32_t)cbs.size();
    t.nof_cbs = C;
    t.cb_data_bits = r.seg.cb_data_bits;
    tbs.push_back(t);
    for (unsigned c = 0; c < C; ++c) {
      lh_pusch_cb_desc d{};
      d.rm.cw_offset = (uint64_t)(q.d_llrs - cw_base) + r.off[c];
      d.rm.rm_offset = r.harq_off + (uint64_t)c * N;
      d.rm.rm_length = r.rm[c];
      d.rm.base_graph = q.base_graph;
      d.rm.lifting_size = Z;
      d.rm.rv = q.rv;
      d.rm.modulation_order = q.modulation_order;
      d.rm.Nref = q.Nref;
      d.rm.nof_filler_bits = r.seg.nof_filler_bits;
      d.bits_offset = msg_base + (uint64_t)c * mb;
      d.nof_crc_bits = r.seg.nof_cb_crc_bits;
      d.crc_poly = poly;
      cbs.push_back(d);
      nd.push_back(q.new_data ? 1 : 0);
      passed.push_back(c < he.cb_iters.size() ? he.cb_iters[c] : (uint8_t)0);
    }
    tbo += ((q.tbs_bits / 8) + 3) & ~3u; // word-aligned TB offsets (the finalize writes words)
  }
  launch_plan plan;
  int rc = pusch_jobs(cbs.data(), (unsigned)cbs.size(), 1, 1, sv->cfg.nof_ldpc_iterations, plan, nd.data(),
                      passed.data());
  if (rc == LH_OK) {
    rc = validate_tbs(tbs.data(), (unsigned)tbs.size());
  }
  if (rc != LH_OK) {
    reject_slot(sv, *sl);
    return rc;
  }
  lh_ctx* ctx = sv->ctx;
  (void)hipSetDevice(ctx->device);
  // Results block of this launch in host memory: TB verdicts, codeblock results, TB bytes.
  const size_t res_at = align256(tbs.size() * 4), tb_at = res_at + align256(cbs.size() * 4);
  sl->h_st = reinterpret_cast<int32_t*>(sl->h_out);
  sl->h_res = reinterpret_cast<int32_t*>(sl->h_out + res_at);
  sl->h_tb = sl->h_out + tb_at;
  tb_signal sig{};
  sig.res_out = reinterpret_cast<int32_t*>(sl->h_out_dev + res_at);
  sig.counter = sl->d_counter;
  sig.flag = sl->h_flag_dev;
  sig.seq = ++sl->seq;
  // One upload carries, at the start of the staging slot, the TB descriptors, each
  // codeblock's transport block and each transport block's codeblock count, then the
  // decoder's descriptor block: the decoder launch reassembles and checks every transport
  // block as soon as its last codeblock is decoded (fused TB finalize).
  slot* s = acquire_slot(ctx);
  if (s == nullptr) {
    reject_slot(sv, *sl);
    return LH_ERR_HIP;
  }
  const size_t tb_desc_bytes = align256(tbs.size() * sizeof(lh_tb_desc));
  const size_t cb_tb_at = tb_desc_bytes, tb_left_at = cb_tb_at + align256(cbs.size() * 4);
  const size_t fa_at = tb_left_at + align256(tbs.size() * 4);
  const size_t prefix = fa_at + align256(sizeof(fused_tb_args));
  std::memcpy(s->host, tbs.data(), tbs.size() * sizeof(lh_tb_desc));
  uint32_t* h_cb_tb = reinterpret_cast<uint32_t*>(s->host + cb_tb_at);
  uint32_t* h_tb_left = reinterpret_cast<uint32_t*>(s->host + tb_left_at);
  for (size_t k = 0; k < tbs.size(); ++k) {
    h_tb_left[k] = tbs[k].nof_cbs;
    for (uint32_t c = 0; c < tbs[k].nof_cbs; ++c) {
      h_cb_tb[tbs[k].first_cb + c] = (uint32_t)k;
    }
  }
  fused_tb_args fa{};
  fa.tbs = reinterpret_cast<const lh_tb_desc_dev*>(s->dev);
  fa.cb_tb = reinterpret_cast<const uint32_t*>(s->dev + cb_tb_at);
  fa.tb_left = reinterpret_cast<uint32_t*>(s->dev + tb_left_at);
  fa.tb_out = sl->h_out_dev + tb_at;
  fa.status = reinterpret_cast<int32_t*>(sl->h_out_dev);
  fa.lut24a = ctx->d_tbcrc;
  fa.sig = sig;
  fa.n_tbs = (uint32_t)tbs.size();
  std::memcpy(s->host + fa_at, &fa, sizeof(fa)); // uploaded with the descriptors
  uint8_t* arena_bytes = reinterpret_cast<uint8_t*>(sv->d_arena);
  rc = run_plan(ctx, plan, sv->d_arena, cw_base, arena_bytes, sl->d_res, sl->stream, s, prefix,
                sv->cfg.learned ? ctx->d_scale : nullptr, sv->cfg.learned ? sv->cfg.nof_fallback_iterations : 0,
                reinterpret_cast<const fused_tb_args*>(s->dev + fa_at));
  hipError_t e = hipSuccess;
  if (rc == LH_OK) {
    e = hipEventRecord(sl->done, sl->stream);
  }
  if (rc != LH_OK || e != hipSuccess) {
    if (rc == LH_OK) {
      rc = hip_fail(e, "server slot launch");
    }
    reject_slot(sv, *sl);
    return rc;
  }
  sl->busy = true;
  return LH_OK;
}

int lh_server_poll(lh_server* sv, lh_tb_result* out, unsigned max, unsigned* n_out, int wait)
{
  if (sv == nullptr || n_out == nullptr || (max > 0 && out == nullptr)) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_poll");
  }
  *n_out = 0;
  bool any = false;
  for (auto& sl : sv->slots) {
    any = any || sl.busy;
  }
  if (wait && sv->results.empty() && any) {
    // Block until the oldest in-flight slot completes.
    for (unsigned k = 0; k < sv->slots.size(); ++k) {
      srv_slot& sl = sv->slots[(sv->next_slot + k) % sv->slots.size()];
      if (sl.busy) {
        const hipError_t e = hipEventSynchronize(sl.done);
        if (e != hipSuccess) {
          return hip_fail(e, "server slot wait");
        }
        // The slot's work has completed, so its completion word must be there.
        if (__atomic_load_n(sl.h_flag, __ATOMIC_ACQUIRE) != sl.seq) {
          return fail(LH_ERR_HIP, "server slot completed without its completion word");
        }
        break;
      }
    }
  }
  // A launch is complete once its finalize kernel has stored the launch's sequence number
  // (the kernel's host writes precede it).
  for (auto& sl : sv->slots) {
    if (sl.busy && __atomic_load_n(sl.h_flag, __ATOMIC_ACQUIRE) == sl.seq) {
      retire_slot(sv, sl);
    }
  }
  admit(sv);
  while (*n_out < max && !sv->results.empty()) {
    out[(*n_out)++] = sv->results.front();
    sv->results.pop_front();
  }
  return LH_OK;
}

int lh_server_harq_release(lh_server* sv, uint32_t harq_process)
{
  if (sv == nullptr) {
    return fail(LH_ERR_INVALID, "invalid server");
  }
  auto h = sv->harq.find(harq_process);
  if (h == sv->harq.end()) {
    return LH_OK;
  }
  if (h->second.busy) {
    return fail(LH_ERR_INVALID, "HARQ process %u is in flight", harq_process);
  }
  sv->arena.release(h->second.off, h->second.bytes);
  sv->harq.erase(h);
  return LH_OK;
}

int lh_server_run_slot(lh_server* sv, uint32_t slot)
{
  if (sv == nullptr) {
    return fail(LH_ERR_INVALID, "invalid server");
  }
  sv->cur_slot = slot;
  sv->have_slot = true;
  if (sv->cfg.expire_timeout_slots == 0) {
    return LH_OK;
  }
  for (auto it = sv->harq.begin(); it != sv->harq.end();) {
    harq_entry& h = it->second;
    if (h.busy) {
      h.expiry = slot + sv->cfg.expire_timeout_slots; // in use: renewed (rx_buffer_pool_impl.cpp:100-105)
      ++it;
    } else if ((int32_t)(slot - h.expiry) >= 0) {
      sv->arena.release(h.off, h.bytes);
      it = sv->harq.erase(it);
    } else {
      ++it;
    }
  }
  admit(sv); // memory freed by expiry admits waiting requests
  return LH_OK;
}

int lh_server_serve(lh_server* sv, const lh_serve_tb* tbs, unsigned n_tbs, const lh_serve_config* cfg,
                    lh_tb_result* results, double* slot_latency_us, double* wall_s)
{
  if (sv == nullptr || cfg == nullptr || results == nullptr || slot_latency_us == nullptr ||
      (n_tbs > 0 && tbs == nullptr) || cfg->slot_us <= 0.0 || cfg->nof_slots == 0) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_serve");
  }
  const unsigned rtt = cfg->harq_rtt_slots;
  const unsigned last_slot = cfg->nof_slots + rtt + 2;
  std::vector<std::vector<uint32_t>> fresh(cfg->nof_slots), retx(last_slot);
  for (unsigned i = 0; i < n_tbs; ++i) {
    results[2 * i] = results[2 * i + 1] = lh_tb_result{};
    results[2 * i].id = 2 * (uint64_t)i;
    results[2 * i + 1].id = 2 * (uint64_t)i + 1;
    results[2 * i].tb_crc_ok = results[2 * i + 1].tb_crc_ok = LH_TB_NOT_SENT;
    if (tbs[i].slot < cfg->nof_slots) {
      fresh[tbs[i].slot].push_back(i);
    }
  }
  std::vector<double> start(last_slot, -1.0);
  std::vector<int> outstanding(last_slot, 0);
  std::vector<uint32_t> sent_in(2 * (size_t)n_tbs, 0);
  for (unsigned k = 0; k < last_slot; ++k) {
    slot_latency_us[k] = -1.0;
  }
  unsigned k = 0, first_open = 0;
  size_t open = 0; // slots with outstanding transmissions
  std::vector<lh_tb_result> got(256);
  std::vector<lh_tb_request> reqs;
  std::vector<uint8_t*> outs;
  const double t0 = now_us();
  int rc = LH_OK;
  auto complete = [&](const lh_tb_result& r) {
    if (r.id >= 2 * (uint64_t)n_tbs) {
      return;
    }
    results[r.id] = r;
    const uint32_t s = sent_in[r.id];
    if (--outstanding[s] == 0) {
      slot_latency_us[s] = now_us() - start[s];
      --open;
    }
    const unsigned i = (unsigned)(r.id / 2), tx = (unsigned)(r.id % 2);
    if (r.tb_crc_ok != LH_TB_CRC_OK) {
      const unsigned due = std::max(s + rtt, k);
      if (tx == 0 && due < last_slot) {
        retx[due].push_back(i);
      } else {
        (void)lh_server_harq_release(sv, tbs[i].tx[0].harq_process);
      }
    }
  };
  auto drain = [&](int wait) {
    unsigned n = 0;
    int r = lh_server_poll(sv, got.data(), (unsigned)got.size(), &n, wait);
    for (unsigned j = 0; j < n; ++j) {
      complete(got[j]);
    }
    return r;
  };
  while (rc == LH_OK && (k < last_slot || open > 0)) {
    if (cfg->realtime) {
      rc = drain(0);
      if (rc != LH_OK) {
        break;
      }
    }
    const double now = now_us();
    while (first_open < k && outstanding[first_open] == 0) {
      ++first_open;
    }
    // Deterministic release: slot k once the slots whose retransmissions fall in it are
    // done; or when nothing is in flight or ready (requests wait for HARQ memory that only
    // the run's progress frees).
    bool due = cfg->realtime != 0 || open == 0 || first_open + rtt >= k + 1;
    if (!cfg->realtime && !due) {
      lh_server_stats_t st;
      (void)lh_server_stats(sv, &st);
      due = st.slots_in_flight == 0 && st.ready == 0;
    }
    if (now - t0 > cfg->timeout_s * 1e6) {
      rc = fail(LH_ERR_CAPACITY, "serving loop exceeded %.1f s at slot %u", cfg->timeout_s, k);
      break;
    }
    if (k < last_slot && ((cfg->realtime && now >= t0 + k * cfg->slot_us) || (!cfg->realtime && due))) {
      (void)lh_server_run_slot(sv, k);
      reqs.clear();
      outs.clear();
      const std::vector<uint32_t>* lists[2] = {k < cfg->nof_slots ? &fresh[k] : nullptr, &retx[k]};
      for (int tx = 0; tx < 2; ++tx) {
        if (lists[tx] == nullptr) {
          continue;
        }
        for (uint32_t i : *lists[tx]) {
          lh_tb_request q = tbs[i].tx[tx];
          q.id = 2 * (uint64_t)i + tx;
          q.slot = k;
          sent_in[q.id] = k;
          reqs.push_back(q);
          outs.push_back(tbs[i].tb_out[tx]);
        }
      }
      if (!reqs.empty()) {
        start[k] = now_us();
        outstanding[k] += (int)reqs.size();
        ++open;
        rc = lh_server_submit_batch(sv, reqs.data(), (unsigned)reqs.size(), outs.data());
        if (rc != LH_OK) {
          break;
        }
      }
      ++k;
    }
    rc = lh_server_step(sv);
    if (rc == LH_OK) {
      rc = drain(!cfg->realtime && (k >= last_slot || !due) ? 1 : 0);
    }
  }
  if (wall_s != nullptr) {
    *wall_s = (now_us() - t0) * 1e-6;
  }
  return rc;
}

int lh_server_stats(lh_server* sv, lh_server_stats_t* st)
{
  if (sv == nullptr || st == nullptr) {
    return fail(LH_ERR_INVALID, "invalid arguments to lh_server_stats");
  }
  st->harq_free_bytes = sv->arena.free_bytes();
  st->harq_processes = (uint32_t)sv->harq.size();
  st->pending = (uint32_t)sv->pending.size();
  st->ready = (uint32_t)sv->ready.size();
  uint32_t busy = 0;
  for (auto& sl : sv->slots) {
    busy += sl.busy ? 1u : 0u;
  }
  st->slots_in_flight = busy;
  return LH_OK;
}

} // extern "C"
