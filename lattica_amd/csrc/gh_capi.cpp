// extern "C" entry points of libgridhip.so (declared in include/gridhip.h).
//
// Host-memory calls stage buffers through a grow-only device staging area,
// enqueue the kernels on the context stream and wait; device-memory calls only
// enqueue. Validation mirrors the reference: dimension checks of
// commit/src/linear.rs:256-327 and the canonical-value check of
// PrimeField's Valid impl (algebra/src/arith/prime.rs:723-728).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>

#include "gh_internal.h"

using namespace gh;

namespace gh {

// Grow-only device buffer; the stream is drained before a buffer is freed.
static int grow(gh_ctx* ctx, void** buf, size_t* have, size_t need) {
  if (*have >= need) return GH_OK;
  if (*buf) {
    GH_HIP(hipStreamSynchronize(ctx->stream));
    GH_HIP(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
  }
  size_t sz = need < (1u << 20) ? (1u << 20) : need;
  sz = (sz + 4095) & ~(size_t)4095;
  GH_HIP(hipMalloc(buf, sz));
  *have = sz;
  return GH_OK;
}

int ctx_workspace(gh_ctx* ctx, size_t bytes, void** out) {
  // a pending finalize reads its partial sums from the workspace: issue it
  // before the caller reuses or regrows the buffer
  int st = ctx_flush(ctx);
  if (st) return st;
  st = grow(ctx, &ctx->work, &ctx->work_bytes, bytes);
  if (st) return st;
  *out = ctx->work;
  return GH_OK;
}

int ctx_staging(gh_ctx* ctx, size_t bytes, void** out) {
  int st = grow(ctx, &ctx->stage, &ctx->stage_bytes, bytes);
  if (st) return st;
  *out = ctx->stage;
  return GH_OK;
}

void prof_pair(gh_ctx* ctx, int phase, hipEvent_t* start, hipEvent_t* stop) {
  *start = *stop = nullptr;
  if (!ctx->prof_on) return;
  while (ctx->prof_used + 2 > ctx->prof_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ctx->prof_pool.push_back(e);
  }
  *start = ctx->prof_pool[ctx->prof_used++];
  *stop = ctx->prof_pool[ctx->prof_used++];
  ctx->prof_recs.push_back({phase, *start, *stop});
}

// Host inputs are checked on the host before upload.
int check_canonical_host(const gh_ring* ring, const void* p, size_t words) {
  if (ring->word_bytes == 4) {
    const uint32_t* w = (const uint32_t*)p;
    const uint32_t q = (uint32_t)ring->q;
    for (size_t i = 0; i < words; i++)
      if (w[i] >= q) return GH_E_NOT_CANONICAL;
  } else {
    const uint64_t* w = (const uint64_t*)p;
    for (size_t i = 0; i < words; i++)
      if (w[i] >= ring->q) return GH_E_NOT_CANONICAL;
  }
  return GH_OK;
}

hipEvent_t prof_mark(gh_ctx* ctx) {
  if (!ctx->prof_on) return nullptr;
  if (ctx->prof_used == ctx->prof_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->prof_pool.push_back(e);
  }
  hipEvent_t e = ctx->prof_pool[ctx->prof_used++];
  if (hipEventRecord(e, ctx->stream) != hipSuccess) return nullptr;
  return e;
}

void prof_span(gh_ctx* ctx, int phase, hipEvent_t a, hipEvent_t b) {
  if (ctx->prof_on && a && b) ctx->prof_recs.push_back({phase, a, b});
}

}  // namespace gh

namespace {

constexpr int kVersion = 1;

// Every call but a device-memory gh_commit starts here: a finalize held back
// by the previous commitment is issued first, so whatever the call reads or
// frees sees the finished commitment.
int set_device(gh_ctx* ctx) {
  GH_HIP(hipSetDevice(ctx->device));
  return ctx_flush(ctx);
}

int check_mem(int mem) { return (mem == GH_MEM_HOST || mem == GH_MEM_DEVICE) ? GH_OK : GH_E_INVALID_ARGUMENT; }

// Rejects device data holding a word >= q (gh_check_canonical).
int check_canonical(gh_ring* ring, const void* dev, size_t words) {
  if (words == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  int* flag = (int*)ctx->misc;
  GH_HIP(hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
  int st = launch_check_canonical(ring, dev, words, flag);
  if (st) return st;
  int host = 0;
  GH_HIP(hipMemcpyAsync(&host, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return host ? GH_E_NOT_CANONICAL : GH_OK;
}

int h2d(gh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) GH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return GH_OK;
}
int d2h(gh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) GH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return GH_OK;
}
// Waits for the context stream: polls first, so a short wait ends within
// about a microsecond of the work (a blocking wait wakes tens of microseconds
// late), then blocks in the runtime once the wait has lasted 2 ms.
int sync(gh_ctx* ctx) {
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipSuccess) return GH_OK;
    if (e != hipErrorNotReady) return e == hipErrorOutOfMemory ? GH_E_OOM : GH_E_DEVICE;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
  }
  GH_HIP(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

#define TRY(x)              \
  do {                      \
    int _s = (x);           \
    if (_s != GH_OK) return _s; \
  } while (0)

}  // namespace

extern "C" {

GH_API const char* gh_status_string(int s) {
  switch (s) {
    case GH_OK: return "ok";
    case GH_E_LENGTH_MISMATCH: return "length mismatch";
    case GH_E_LENGTH_NOT_POW2: return "length is not a power of two";
    case GH_E_UNSUPPORTED_SIZE: return "unsupported NTT size";
    case GH_E_UNSUPPORTED_TWIST: return "unsupported twist (no 2d-th root of unity)";
    case GH_E_INVALID_PARAMETERS: return "invalid commitment parameters";
    case GH_E_DIMENSION_MISMATCH: return "dimension mismatch";
    case GH_E_OPENING_NORM_EXCEEDED: return "opening norm exceeded configured bound";
    case GH_E_INVALID_MESSAGE_ENCODING: return "invalid message encoding";
    case GH_E_INVALID_OPENING: return "invalid opening";
    case GH_E_INVALID_ARGUMENT: return "invalid argument";
    case GH_E_NOT_CANONICAL: return "input value not a canonical residue";
    case GH_E_UNSUPPORTED_MODULUS: return "modulus has no GPU field implementation";
    case GH_E_DEVICE: return "HIP runtime error";
    case GH_E_OOM: return "device out of memory";
    case GH_E_BUSY: return "busy";
    case GH_E_NO_DEVICE: return "no HIP device";
  }
  return "unknown status";
}

GH_API int gh_version(void) { return kVersion; }

GH_API int gh_device_count(int* count) {
  if (!count) return GH_E_INVALID_ARGUMENT;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return GH_OK;
}

GH_API int gh_ctx_create(int device, gh_ctx** out) {
  if (!out) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return GH_E_NO_DEVICE;
  if (device < 0 || device >= n) return GH_E_INVALID_ARGUMENT;
  GH_HIP(hipSetDevice(device));
  gh_ctx* ctx = new (std::nothrow) gh_ctx();
  if (!ctx) return GH_E_OOM;
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return GH_E_DEVICE;
  }
  ctx->num_cus = prop.multiProcessorCount;
  if (hipMalloc(&ctx->misc, 4096) != hipSuccess) {
    gh_ctx_destroy(ctx);
    return GH_E_OOM;
  }
  // the device's own copy of the no-allocation scratch array (device current)
  if (int st = scratch_address(&ctx->scratch)) {
    gh_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return GH_OK;
}

GH_API int gh_ctx_destroy(gh_ctx* ctx) {
  if (!ctx) return GH_E_INVALID_ARGUMENT;
  hipSetDevice(ctx->device);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->stream) ctx_flush(ctx);
  }
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->work) hipFree(ctx->work);
  if (ctx->stage) hipFree(ctx->stage);
  if (ctx->misc) hipFree(ctx->misc);
  if (ctx->batch_done) hipFree(ctx->batch_done);
  for (hipEvent_t e : ctx->prof_pool) hipEventDestroy(e);
  if (ctx->upload_done) hipEventDestroy(ctx->upload_done);
  if (ctx->upload_host) hipHostFree(ctx->upload_host);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return GH_OK;
}

GH_API int gh_ctx_synchronize(gh_ctx* ctx) {
  if (!ctx) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  return sync(ctx);
}

GH_API int gh_ctx_profile_enable(gh_ctx* ctx, int on) {
  if (!ctx) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->prof_on = on != 0;
  return GH_OK;
}

GH_API int gh_ctx_profile_read(gh_ctx* ctx, int phase, double* total_ms, uint64_t* count) {
  if (!ctx || !total_ms || !count) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  TRY(sync(ctx));
  double sum = 0;
  uint64_t n = 0;
  for (const auto& r : ctx->prof_recs) {
    if (r.phase != phase) continue;
    float ms = 0;
    GH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    sum += ms;
    n++;
  }
  *total_ms = sum;
  *count = n;
  // every phase has been read out once the caller asks for the last one
  if (phase == 2) {
    ctx->prof_recs.clear();
    ctx->prof_used = 0;
  }
  return GH_OK;
}

GH_API int gh_ctx_set_tuning(gh_ctx* ctx, const char* key, int64_t value) {
  if (!ctx || !key) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  gh_tuning& t = ctx->tune;
  if (!std::strcmp(key, "mac_variant")) {
    if (value != 0 && value != 1) return GH_E_INVALID_ARGUMENT;
    t.mac_variant = (int)value;
  } else if (!std::strcmp(key, "mac_waves")) {
    if (value != 4 && value != 8 && value != 16) return GH_E_INVALID_ARGUMENT;
    t.mac_waves = (int)value;
  } else if (!std::strcmp(key, "mac_blocks_per_cu")) {
    if (value < 1 || value > 8) return GH_E_INVALID_ARGUMENT;
    t.mac_blocks_per_cu = (int)value;
  } else if (!std::strcmp(key, "quad_max")) {
    if (value < 0) return GH_E_INVALID_ARGUMENT;
    t.quad_max = value;
  } else if (!std::strcmp(key, "fused")) {
    if (value < 0 || value > 2) return GH_E_INVALID_ARGUMENT;
    t.fused = (int)value;
  } else if (!std::strcmp(key, "fused_depth")) {
    if (value < 1 || value > 4) return GH_E_INVALID_ARGUMENT;
    t.fused_depth = (int)value;
  } else if (!std::strcmp(key, "batch_chunk_cols")) {
    if (value < 16 || value > 4096) return GH_E_INVALID_ARGUMENT;
    t.batch_chunk_cols = (int)value;
  } else if (!std::strcmp(key, "batch_chunk_cols_gold")) {
    if (value < 16 || value > 4096) return GH_E_INVALID_ARGUMENT;
    t.batch_chunk_cols_gold = (int)value;
  } else if (!std::strcmp(key, "batch_group")) {
    if (value != 4 && value != 8 && value != 16) return GH_E_INVALID_ARGUMENT;
    t.batch_group = (int)value;
  } else if (!std::strcmp(key, "defer_finalize")) {
    if (value != 0 && value != 1) return GH_E_INVALID_ARGUMENT;
    // turning deferral off completes a finalize held back until now
    if (!value) TRY(set_device(ctx));
    t.defer_finalize = (int)value;
  } else if (!std::strcmp(key, "batch_fused")) {
    if (value != 0 && value != 1) return GH_E_INVALID_ARGUMENT;
    t.batch_fused = (int)value;
  } else if (!std::strcmp(key, "batch_fused_cols")) {
    if (value < 32 || value > 4096 || value % 32) return GH_E_INVALID_ARGUMENT;
    t.batch_fused_cols = (int)value;
  } else if (!std::strcmp(key, "stage1_tasks_per_block")) {
    if (value < -1 || value > (int64_t)kStage1MaxTpb) return GH_E_INVALID_ARGUMENT;
    t.stage1_tasks_per_block = (int)value;
  } else if (!std::strcmp(key, "q32_ring_group")) {
    if (value != 4 && value != 8) return GH_E_INVALID_ARGUMENT;
    t.q32_ring_group = (int)value;
  } else if (!std::strcmp(key, "xform_quad")) {
    if (value < -1 || value > 1) return GH_E_INVALID_ARGUMENT;
    t.xform_quad = (int)value;
  } else if (!std::strcmp(key, "quad_lds")) {
    if (value < 0 || value > 1) return GH_E_INVALID_ARGUMENT;
    t.quad_lds = (int)value;
  } else if (!std::strcmp(key, "negacyclic_quad")) {
    if (value < -1 || value > 1) return GH_E_INVALID_ARGUMENT;
    t.negacyclic_quad = (int)value;
  } else if (!std::strcmp(key, "batch_mac_ring")) {
    if (value != 0 && (value < 2 || value > 4)) return GH_E_INVALID_ARGUMENT;
    t.batch_mac_ring = (int)value;
  } else {
    return GH_E_INVALID_ARGUMENT;
  }
  return GH_OK;
}

GH_API int gh_ctx_get_tuning(gh_ctx* ctx, const char* key, int64_t* value) {
  if (!ctx || !key || !value) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const gh_tuning& t = ctx->tune;
  if (!std::strcmp(key, "mac_variant")) *value = t.mac_variant;
  else if (!std::strcmp(key, "mac_waves")) *value = t.mac_waves;
  else if (!std::strcmp(key, "mac_blocks_per_cu")) *value = t.mac_blocks_per_cu;
  else if (!std::strcmp(key, "quad_max")) *value = t.quad_max;
  else if (!std::strcmp(key, "fused")) *value = t.fused;
  else if (!std::strcmp(key, "fused_depth")) *value = t.fused_depth;
  else if (!std::strcmp(key, "batch_chunk_cols")) *value = t.batch_chunk_cols;
  else if (!std::strcmp(key, "batch_chunk_cols_gold")) *value = t.batch_chunk_cols_gold;
  else if (!std::strcmp(key, "batch_group")) *value = t.batch_group;
  else if (!std::strcmp(key, "defer_finalize")) *value = t.defer_finalize;
  else if (!std::strcmp(key, "stage1_tasks_per_block")) *value = t.stage1_tasks_per_block;
  else if (!std::strcmp(key, "batch_fused")) *value = t.batch_fused;
  else if (!std::strcmp(key, "batch_fused_cols")) *value = t.batch_fused_cols;
  else if (!std::strcmp(key, "batch_mac_ring")) *value = t.batch_mac_ring;
  else if (!std::strcmp(key, "negacyclic_quad")) *value = t.negacyclic_quad;
  else if (!std::strcmp(key, "q32_ring_group")) *value = t.q32_ring_group;
  else if (!std::strcmp(key, "quad_lds")) *value = t.quad_lds;
  else if (!std::strcmp(key, "xform_quad")) *value = t.xform_quad;
  else return GH_E_INVALID_ARGUMENT;
  return GH_OK;
}

GH_API int gh_measure_bandwidth(gh_ctx* ctx, size_t bytes, int reps, double* read_gbps,
                                double* copy_gbps) {
  if (!ctx || !read_gbps || !copy_gbps || reps < 1 || bytes < 16) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  bytes &= ~(size_t)15;
  // below 512 MiB three source buffers are read in turn, so that the 256 MB
  // Infinity Cache holds none of them when it is read again
  const int nsrc = bytes <= ((size_t)512 << 20) ? 3 : 1;
  void *src[3] = {nullptr, nullptr, nullptr}, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = GH_OK;
  for (int i = 0; i < nsrc && st == GH_OK; i++)
    if (hipMalloc(&src[i], bytes) != hipSuccess) st = GH_E_OOM;
  if (st == GH_OK && hipMalloc(&dst, bytes) != hipSuccess) st = GH_E_OOM;
  if (st == GH_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = GH_E_DEVICE;
  for (int i = 0; i < nsrc && st == GH_OK; i++)
    if (hipMemsetAsync(src[i], 0x5a, bytes, ctx->stream) != hipSuccess) st = GH_E_DEVICE;
  double best[2] = {0, 0};
  for (int copy = 0; copy < 2 && st == GH_OK; copy++) {
    for (int r = 0; r < reps + nsrc && st == GH_OK; r++) {  // the first pass over the sources warms up
      st = launch_stream_probe(ctx, src[r % nsrc], dst, bytes, copy, e0, e1);
      float ms = 0;
      if (st == GH_OK && (hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = GH_E_DEVICE;
      if (st == GH_OK && r >= nsrc && ms > 0) {
        const double gbps = (copy ? 2.0 : 1.0) * (double)bytes / (ms * 1e-3) / 1e9;
        if (gbps > best[copy]) best[copy] = gbps;
      }
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(ctx->stream);
  for (int i = 0; i < nsrc; i++)
    if (src[i]) hipFree(src[i]);
  if (dst) hipFree(dst);
  *read_gbps = best[0];
  *copy_gbps = best[1];
  return st;
}

GH_API int gh_measure_issue(gh_ctx* ctx, int waves_per_simd, double out_us[4]) {
  if (!ctx || !out_us || waves_per_simd < 1 || waves_per_simd > 8) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  // 4-wave blocks: waves_per_simd blocks per CU
  const uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)waves_per_simd;
  int* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = GH_OK;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(int)) != hipSuccess) st = GH_E_OOM;
  if (st == GH_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = GH_E_DEVICE;
  for (int kind = 0; kind < 4 && st == GH_OK; kind++) {
    double best = 0;
    for (int r = 0; r < 4 && st == GH_OK; r++) {  // the first run warms up
      st = launch_issue_probe(ctx, kind, blocks, out, e0, e1);
      float ms = 0;
      if (st == GH_OK && (hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = GH_E_DEVICE;
      if (st == GH_OK && r > 0 && (best == 0 || ms * 1e3 < best)) best = ms * 1e3;
    }
    out_us[kind] = best;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(ctx->stream);
  if (out) hipFree(out);
  return st;
}

GH_API int gh_measure_valu(gh_ctx* ctx, int waves_per_simd, double out_gips[10]) {
  if (!ctx || !out_gips || waves_per_simd < 1 || waves_per_simd > 8) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)waves_per_simd;
  uint32_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = GH_OK;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess) st = GH_E_OOM;
  if (st == GH_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = GH_E_DEVICE;
  // wave instructions of the measured kind per launch: 4 waves per block,
  // 2048 iterations of eight
  const double winst = (double)blocks * 4 * 2048 * 8;
  for (int kind = 0; kind < kValuKinds && st == GH_OK; kind++) {
    double best = 0;
    for (int r = 0; r < 4 && st == GH_OK; r++) {  // the first run warms up
      st = launch_valu_probe(ctx, kind, blocks, out, e0, e1);
      float ms = 0;
      if (st == GH_OK && (hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = GH_E_DEVICE;
      if (st == GH_OK && r > 0 && ms > 0 && (best == 0 || ms < best)) best = ms;
    }
    // kind 6 is a three-instruction step
    const double per_step = kind == 6 ? 3.0 : 1.0;
    out_gips[kind] = best > 0 ? per_step * winst / (best * 1e-3) / 1e9 : 0;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(ctx->stream);
  if (out) hipFree(out);
  return st;
}

GH_API int gh_ctx_num_cus(gh_ctx* ctx, int* out) {
  if (!ctx || !out) return GH_E_INVALID_ARGUMENT;
  *out = ctx->num_cus;
  return GH_OK;
}

GH_API int gh_alloc(gh_ctx* ctx, size_t bytes, void** dev) {
  if (!ctx || !dev) return GH_E_INVALID_ARGUMENT;
  TRY(set_device(ctx));
  *dev = nullptr;
  GH_HIP(hipMalloc(dev, bytes ? bytes : 1));
  return GH_OK;
}

GH_API int gh_free(gh_ctx* ctx, void* dev) {
  if (!ctx) return GH_E_INVALID_ARGUMENT;
  if (!dev) return GH_OK;
  TRY(set_device(ctx));
  GH_HIP(hipStreamSynchronize(ctx->stream));
  GH_HIP(hipFree(dev));
  return GH_OK;
}

GH_API int gh_upload(gh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return GH_E_INVALID_ARGUMENT;
  TRY(set_device(ctx));
  TRY(h2d(ctx, dst, src, bytes));
  return sync(ctx);
}

GH_API int gh_download(gh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return GH_E_INVALID_ARGUMENT;
  TRY(set_device(ctx));
  TRY(d2h(ctx, dst, src, bytes));
  return sync(ctx);
}

GH_API int gh_ring_create(gh_ctx* ctx, uint64_t q, uint32_t d, gh_ring** out) {
  if (!ctx || !out) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  gh_ring* ring = new (std::nothrow) gh_ring();
  if (!ring) return GH_E_OOM;
  int st = plan_ring(q, d, ring);
  if (st != GH_OK) {
    delete ring;
    return st;
  }
  ring->ctx = ctx;
  if ((st = set_device(ctx)) != GH_OK) { delete ring; return st; }
  const size_t tab_bytes = ring->host_tab.size() * sizeof(uint64_t);
  if (hipMalloc(&ring->dev_tab, tab_bytes) != hipSuccess) {
    delete ring;
    return GH_E_OOM;
  }
  if (hipMemcpy(ring->dev_tab, ring->host_tab.data(), tab_bytes, hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(ring->dev_tab);
    delete ring;
    return GH_E_DEVICE;
  }
  *out = ring;
  return GH_OK;
}

GH_API int gh_ring_destroy(gh_ring* ring) {
  if (!ring) return GH_E_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(ring->ctx->mu);
    TRY(set_device(ring->ctx));  // a pending finalize may read this ring's table
  }
  hipStreamSynchronize(ring->ctx->stream);
  if (ring->dev_tab) hipFree(ring->dev_tab);
  delete ring;
  return GH_OK;
}

GH_API int gh_ring_query(const gh_ring* ring, gh_ring_info* out) {
  if (!ring || !out) return GH_E_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof *out);
  out->q = ring->q;
  out->d = ring->d;
  out->word_bytes = ring->word_bytes;
  out->kind = (uint32_t)ring->kind;
  out->wire_bytes = ring->wire_bytes;
  out->generator = ring->generator;
  out->psi = ring->psi;
  out->omega = ring->omega;
  out->d_inv = ring->d_inv;
  return GH_OK;
}

}  // extern "C"

namespace {

int xform_entry(gh_ring* ring, void* values, size_t count, int mem, int mode) {
  if (!ring || (count && !values)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (count == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t bytes = count * ring->d * ring->word_bytes;
  if (mem == GH_MEM_DEVICE) return launch_xform(ring, mode, values, values, count);
  TRY(check_canonical_host(ring, values, count * ring->d));
  void* dev;
  TRY(ctx_staging(ctx, bytes, &dev));
  TRY(h2d(ctx, dev, values, bytes));
  TRY(launch_xform(ring, mode, dev, dev, count));
  TRY(d2h(ctx, values, dev, bytes));
  return sync(ctx);
}

int binary_entry(gh_ring* ring, const void* a, const void* b, void* out, size_t count, int mem,
                 bool negacyclic) {
  if (!ring || (count && (!a || !b || !out))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (count == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t words = count * ring->d;
  const size_t bytes = words * ring->word_bytes;
  if (mem == GH_MEM_DEVICE)
    return negacyclic ? launch_negacyclic_mul(ring, a, b, out, count)
                      : launch_pointwise_mul(ring, a, b, out, count);
  TRY(check_canonical_host(ring, a, words));
  TRY(check_canonical_host(ring, b, words));
  void* dev;
  TRY(ctx_staging(ctx, 2 * bytes, &dev));
  void* da = dev;
  void* db = (char*)dev + bytes;
  TRY(h2d(ctx, da, a, bytes));
  TRY(h2d(ctx, db, b, bytes));
  TRY(negacyclic ? launch_negacyclic_mul(ring, da, db, da, count)
                 : launch_pointwise_mul(ring, da, db, da, count));
  TRY(d2h(ctx, out, da, bytes));
  return sync(ctx);
}

}  // namespace

extern "C" {

GH_API int gh_ntt_forward(gh_ring* r, void* v, size_t c, int mem) { return xform_entry(r, v, c, mem, kCyclicFwd); }
GH_API int gh_ntt_inverse(gh_ring* r, void* v, size_t c, int mem) { return xform_entry(r, v, c, mem, kCyclicInv); }
GH_API int gh_twisted_forward(gh_ring* r, void* v, size_t c, int mem) { return xform_entry(r, v, c, mem, kTwistedFwd); }
GH_API int gh_twisted_inverse(gh_ring* r, void* v, size_t c, int mem) { return xform_entry(r, v, c, mem, kTwistedInv); }

GH_API int gh_negacyclic_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count, int mem) {
  return binary_entry(ring, a, b, out, count, mem, true);
}
GH_API int gh_pointwise_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count, int mem) {
  return binary_entry(ring, a, b, out, count, mem, false);
}

}  // extern "C"

namespace {

// Ring slice hooks: dst op= src (src unused by kSliceScale) on `len` words.
int slice_entry(gh_ring* ring, int op, void* dst, const void* src, uint64_t scalar, size_t len,
                int mem) {
  const bool has_src = op != kSliceScale;
  if (!ring || (len && (!dst || (has_src && !src)))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if ((op == kSliceScale || op == kSliceAxpy) && scalar >= ring->q) return GH_E_NOT_CANONICAL;
  if (len == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (mem == GH_MEM_DEVICE) return launch_slice(ring, op, dst, src, scalar, len);
  TRY(check_canonical_host(ring, dst, len));
  if (has_src) TRY(check_canonical_host(ring, src, len));
  const size_t bytes = len * ring->word_bytes;
  const size_t bpad = (bytes + 255) & ~(size_t)255;  // both operands 16-byte aligned
  void* dev;
  TRY(ctx_staging(ctx, (has_src ? 2 : 1) * bpad, &dev));
  void* dsrc = has_src ? (char*)dev + bpad : nullptr;
  TRY(h2d(ctx, dev, dst, bytes));
  if (has_src) TRY(h2d(ctx, dsrc, src, bytes));
  TRY(launch_slice(ring, op, dev, dsrc, scalar, len));
  TRY(d2h(ctx, dst, dev, bytes));
  return sync(ctx);
}

}  // namespace

extern "C" {

GH_API int gh_add_assign(gh_ring* ring, void* dst, const void* src, size_t len, int mem) {
  return slice_entry(ring, kSliceAdd, dst, src, 0, len, mem);
}
GH_API int gh_sub_assign(gh_ring* ring, void* dst, const void* src, size_t len, int mem) {
  return slice_entry(ring, kSliceSub, dst, src, 0, len, mem);
}
GH_API int gh_scalar_mul(gh_ring* ring, void* dst, uint64_t scalar, size_t len, int mem) {
  return slice_entry(ring, kSliceScale, dst, nullptr, scalar, len, mem);
}
GH_API int gh_add_assign_scaled(gh_ring* ring, void* dst, const void* src, uint64_t scalar,
                                size_t len, int mem) {
  return slice_entry(ring, kSliceAxpy, dst, src, scalar, len, mem);
}

GH_API int gh_dot_product(gh_ring* ring, const void* a, const void* b, size_t len, void* out,
                          int mem) {
  if (!ring || !out || (len && (!a || !b))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes;
  void* partial;
  TRY(ctx_workspace(ctx, dot_partial_words(ring, len) * wb, &partial));
  if (mem == GH_MEM_DEVICE) return launch_dot(ring, a, b, len, partial, out);
  TRY(check_canonical_host(ring, a, len));
  TRY(check_canonical_host(ring, b, len));
  const size_t bytes = len * wb;
  const size_t bpad = (bytes + 255) & ~(size_t)255;
  void* dev;
  TRY(ctx_staging(ctx, 2 * bpad + 256, &dev));
  void* da = dev;
  void* db = (char*)dev + bpad;
  void* dout = (char*)dev + 2 * bpad;
  TRY(h2d(ctx, da, a, bytes));
  TRY(h2d(ctx, db, b, bytes));
  TRY(launch_dot(ring, da, db, len, partial, dout));
  TRY(d2h(ctx, out, dout, wb));
  return sync(ctx);
}

}  // extern "C"

namespace {

// Builds a resident key from device sources (context lock held): the
// [col][row][slot] layout of both blocks, then, for a coefficient-domain key,
// prepare_poly_matrix (linear.rs:345-359): one forward transform; the layout
// only permutes whole polynomials.
int key_create_dev(gh_ring* ring, uint32_t rows, const void* src_msg, uint32_t msg_len,
                   const void* src_open, uint32_t open_len, int domain, gh_key** out) {
  gh_ctx* ctx = ring->ctx;
  const uint32_t cols = msg_len + open_len;
  const size_t key_bytes = (size_t)rows * cols * ring->d * ring->word_bytes;
  gh_key* key = new (std::nothrow) gh_key();
  if (!key) return GH_E_OOM;
  key->ring = ring;
  key->rows = rows;
  key->msg_len = msg_len;
  key->open_len = open_len;
  if (hipMalloc(&key->dev_a, key_bytes) != hipSuccess) {
    delete key;
    return GH_E_OOM;
  }
  int st = launch_key_layout(ring, src_msg, rows, msg_len, 0, cols, key->dev_a, domain);
  if (st == GH_OK && open_len)
    st = launch_key_layout(ring, src_open, rows, open_len, msg_len, cols, key->dev_a, domain);
  if (st == GH_OK && domain == GH_DOMAIN_COEFF)
    st = launch_xform(ring, kTwistedFwd, key->dev_a, key->dev_a, (size_t)rows * cols);
  if (st == GH_OK) st = sync(ctx);
  if (st != GH_OK) {
    hipFree(key->dev_a);
    delete key;
    return st;
  }
  *out = key;
  return GH_OK;
}

}  // namespace

extern "C" {

GH_API int gh_key_create(gh_ring* ring, uint32_t rows, const void* a_msg, uint32_t msg_len,
                         const void* a_open, uint32_t open_len, int domain, int mem, gh_key** out) {
  if (!ring || !out) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  // CommitmentDimensions::validate (linear.rs:45-59): message and commitment
  // lengths must be non-zero; the opening block may be empty here.
  if (rows == 0 || msg_len == 0) return GH_E_INVALID_PARAMETERS;
  if (!a_msg || (open_len && !a_open)) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes, d = ring->d;
  const uint32_t cols = msg_len + open_len;
  if (cols < msg_len) return GH_E_INVALID_PARAMETERS;
  const size_t msg_bytes = (size_t)rows * msg_len * d * wb;
  const size_t open_bytes = (size_t)rows * open_len * d * wb;
  const void* src_msg = a_msg;
  const void* src_open = a_open;
  if (mem == GH_MEM_HOST) {
    TRY(check_canonical_host(ring, a_msg, (size_t)rows * msg_len * d));
    if (open_len) TRY(check_canonical_host(ring, a_open, (size_t)rows * open_len * d));
    void* dev;
    TRY(ctx_staging(ctx, msg_bytes + open_bytes, &dev));
    TRY(h2d(ctx, dev, a_msg, msg_bytes));
    if (open_len) TRY(h2d(ctx, (char*)dev + msg_bytes, a_open, open_bytes));
    src_msg = dev;
    src_open = (char*)dev + msg_bytes;
  }
  return key_create_dev(ring, rows, src_msg, msg_len, src_open, open_len, domain, out);
}

GH_API int gh_key_destroy(gh_key* key) {
  if (!key) return GH_E_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(key->ring->ctx->mu);
    TRY(set_device(key->ring->ctx));
  }
  hipStreamSynchronize(key->ring->ctx->stream);
  if (key->dev_a) hipFree(key->dev_a);
  delete key;
  return GH_OK;
}

GH_API int gh_key_dims(const gh_key* key, uint32_t* rows, uint32_t* msg_len, uint32_t* open_len) {
  if (!key) return GH_E_INVALID_ARGUMENT;
  if (rows) *rows = key->rows;
  if (msg_len) *msg_len = key->msg_len;
  if (open_len) *open_len = key->open_len;
  return GH_OK;
}

GH_API int gh_commit(gh_key* key, const void* message, uint32_t msg_len, const void* randomness,
                     uint32_t open_len, int domain, void* t_out, int mem) {
  if (!key || !message || !t_out || (open_len && !randomness)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  // validate_message / validate_opening_randomness (linear.rs:270-293)
  if (msg_len != key->msg_len || open_len != key->open_len) return GH_E_DIMENSION_MISMATCH;
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  GH_HIP(hipSetDevice(ctx->device));
  const size_t wb = ring->word_bytes, d = ring->d;
  const uint32_t cols = msg_len + open_len;
  const size_t msg_bytes = (size_t)msg_len * d * wb, open_bytes = (size_t)open_len * d * wb;
  const size_t t_bytes = (size_t)key->rows * d * wb;
  // device memory, no opening: asynchronous, and the finalize may be held
  // back until the next call on the context (the next such commitment runs it
  // inside its own launch)
  if (mem == GH_MEM_DEVICE && open_len == 0)
    return launch_commit(key, message, cols, domain, t_out, ctx->tune.defer_finalize != 0);
  TRY(ctx_flush(ctx));
  void* dev;
  TRY(ctx_staging(ctx, msg_bytes + open_bytes + t_bytes, &dev));
  void* ds = dev;
  void* dt = (char*)dev + msg_bytes + open_bytes;
  if (mem == GH_MEM_HOST) {
    TRY(check_canonical_host(ring, message, (size_t)msg_len * d));
    if (open_len) TRY(check_canonical_host(ring, randomness, (size_t)open_len * d));
    TRY(h2d(ctx, ds, message, msg_bytes));
    if (open_len) TRY(h2d(ctx, (char*)ds + msg_bytes, randomness, open_bytes));
  } else {
    GH_HIP(hipMemcpyAsync(ds, message, msg_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GH_HIP(hipMemcpyAsync((char*)ds + msg_bytes, randomness, open_bytes, hipMemcpyDeviceToDevice,
                          ctx->stream));
  }
  TRY(launch_commit(key, ds, cols, domain, mem == GH_MEM_HOST ? dt : t_out));
  if (mem == GH_MEM_HOST) {
    TRY(d2h(ctx, t_out, dt, t_bytes));
    return sync(ctx);
  }
  return GH_OK;
}

GH_API int gh_commit_noalloc(gh_key* key, const void* message, uint32_t msg_len, void* t_out) {
  if (!key || !message || !t_out) return GH_E_INVALID_ARGUMENT;
  if (msg_len != key->msg_len) return GH_E_DIMENSION_MISMATCH;
  if (key->open_len != 0) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = key->ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  GH_HIP(hipSetDevice(ctx->device));
  return launch_commit_noalloc(key, message, msg_len, t_out, ctx->scratch);
}

GH_API int gh_decompose_balanced(gh_ring* ring, const void* values, size_t count, uint64_t base,
                                 uint32_t limbs, void* out, int mem) {
  if (!ring || (count && (!values || !out))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (base < 2 || base > (1ull << 32) || limbs == 0) return GH_E_INVALID_PARAMETERS;
  if (count == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes, d = ring->d;
  if (mem == GH_MEM_DEVICE) return launch_decompose(ring, values, count, base, limbs, out);
  TRY(check_canonical_host(ring, values, count * d));
  const size_t in_bytes = count * d * wb, out_bytes = in_bytes * limbs;
  void* dev;
  TRY(ctx_staging(ctx, in_bytes + out_bytes, &dev));
  TRY(h2d(ctx, dev, values, in_bytes));
  TRY(launch_decompose(ring, dev, count, base, limbs, (char*)dev + in_bytes));
  TRY(d2h(ctx, out, (char*)dev + in_bytes, out_bytes));
  return sync(ctx);
}

GH_API int gh_garbage_h(gh_ring* ring, const void* phi, const void* s, uint32_t r, uint32_t n,
                        int domain, void* h_out, int mem) {
  if (!ring || (r && (!phi || !s || !h_out))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  if (r == 0) return GH_OK;
  if (r > 0xffff || (uint64_t)r * n > 0xffffffffull) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t h_bytes = (size_t)r * (r + 1) / 2 * d * wb;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (n == 0) {  // empty witness: every h_ij is zero
    if (mem == GH_MEM_HOST) { std::memset(h_out, 0, h_bytes); return GH_OK; }
    GH_HIP(hipMemsetAsync(h_out, 0, h_bytes, ctx->stream));
    return GH_OK;
  }
  if (mem == GH_MEM_DEVICE) return launch_garbage_h(ring, phi, s, r, n, domain, h_out);
  const size_t in_bytes = (size_t)r * n * d * wb;
  TRY(check_canonical_host(ring, phi, (size_t)r * n * d));
  TRY(check_canonical_host(ring, s, (size_t)r * n * d));
  void* dev;
  TRY(ctx_staging(ctx, 2 * in_bytes + h_bytes, &dev));
  char* b = (char*)dev;
  TRY(h2d(ctx, b, phi, in_bytes));
  TRY(h2d(ctx, b + in_bytes, s, in_bytes));
  TRY(launch_garbage_h(ring, b, b + in_bytes, r, n, domain, b + 2 * in_bytes));
  TRY(d2h(ctx, h_out, b + 2 * in_bytes, h_bytes));
  return sync(ctx);
}

GH_API int gh_amortize(gh_ring* ring, const void* c, const void* s, uint32_t r, uint32_t n,
                       int domain, void* z_out, int mem) {
  if (!ring || (n && (!z_out || (r && (!c || !s))))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  if (n == 0) return GH_OK;
  if ((uint64_t)r * n > 0xffffffffull) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t z_bytes = (size_t)n * d * wb;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (r == 0) {  // empty sum
    if (mem == GH_MEM_HOST) { std::memset(z_out, 0, z_bytes); return GH_OK; }
    GH_HIP(hipMemsetAsync(z_out, 0, z_bytes, ctx->stream));
    return GH_OK;
  }
  const size_t c_bytes = (size_t)r * d * wb, s_bytes = (size_t)r * n * d * wb;
  const size_t scratch = domain == GH_DOMAIN_COEFF ? s_bytes : 0;
  void* dev;
  if (mem == GH_MEM_DEVICE) {
    void* scr = nullptr;
    if (scratch) TRY(ctx_staging(ctx, scratch, &scr));
    return launch_amortize(ring, c, s, r, n, domain, z_out, scr);
  }
  TRY(check_canonical_host(ring, c, (size_t)r * d));
  TRY(check_canonical_host(ring, s, (size_t)r * n * d));
  TRY(ctx_staging(ctx, c_bytes + s_bytes + z_bytes + scratch, &dev));
  char* b = (char*)dev;
  TRY(h2d(ctx, b, c, c_bytes));
  TRY(h2d(ctx, b + c_bytes, s, s_bytes));
  char* zd = b + c_bytes + s_bytes;
  TRY(launch_amortize(ring, b, b + c_bytes, r, n, domain, zd, scratch ? zd + z_bytes : nullptr));
  TRY(d2h(ctx, z_out, zd, z_bytes));
  return sync(ctx);
}

GH_API int gh_bdlop_commit(gh_key* mask, gh_key* key, const void* message, uint32_t msg_len,
                           const void* randomness, uint32_t open_len, int domain, void* u_out,
                           void* v_out, int mem) {
  if (!mask || !key || !message || !randomness || !u_out || !v_out) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  if (mask->ring != key->ring) return GH_E_INVALID_ARGUMENT;
  // validate_bdlop_key / validate_message / validate_opening_randomness
  // (commit/src/bdlop/mod.rs:56-64, commit/src/linear.rs:270-293)
  if (open_len == 0 || msg_len != key->msg_len || open_len != key->open_len ||
      mask->msg_len != open_len || mask->open_len != 0)
    return GH_E_DIMENSION_MISMATCH;
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t msg_bytes = (size_t)msg_len * d * wb, open_bytes = (size_t)open_len * d * wb;
  const size_t u_bytes = (size_t)mask->rows * d * wb, v_bytes = (size_t)key->rows * d * wb;
  // [m | r] contiguous: v streams [A_msg | A_open] over it, u streams A_mask over r
  void* dev;
  TRY(ctx_staging(ctx, msg_bytes + open_bytes + u_bytes + v_bytes, &dev));
  char* ds = (char*)dev;
  char* du = ds + msg_bytes + open_bytes;
  char* dv = du + u_bytes;
  if (mem == GH_MEM_HOST) {
    TRY(check_canonical_host(ring, message, (size_t)msg_len * d));
    TRY(check_canonical_host(ring, randomness, (size_t)open_len * d));
    TRY(h2d(ctx, ds, message, msg_bytes));
    TRY(h2d(ctx, ds + msg_bytes, randomness, open_bytes));
  } else {
    GH_HIP(hipMemcpyAsync(ds, message, msg_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GH_HIP(hipMemcpyAsync(ds + msg_bytes, randomness, open_bytes, hipMemcpyDeviceToDevice,
                          ctx->stream));
  }
  const bool host = mem == GH_MEM_HOST;
  TRY(launch_commit(key, ds, msg_len + open_len, domain, host ? dv : v_out));
  TRY(launch_commit(mask, ds + msg_bytes, open_len, domain, host ? du : u_out));
  if (host) {
    TRY(d2h(ctx, u_out, du, u_bytes));
    TRY(d2h(ctx, v_out, dv, v_bytes));
    return sync(ctx);
  }
  return GH_OK;
}

}  // extern "C"

extern "C" {

GH_API int gh_check_canonical(gh_ring* ring, const void* values, size_t words, int mem) {
  if (!ring || (words && !values)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (mem == GH_MEM_HOST) return check_canonical_host(ring, values, words);
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  return check_canonical(ring, values, words);
}

GH_API int gh_inner_commit(gh_key* key, const void* s, uint32_t r, uint32_t n, int domain,
                           void* t_out, int mem) {
  if (!key || (r && (!s || !t_out))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  if (n != key->msg_len || key->open_len != 0) return GH_E_DIMENSION_MISMATCH;
  if (r == 0) return GH_OK;
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t s_bytes = (size_t)r * n * d * wb, t_bytes = (size_t)r * key->rows * d * wb;
  const char* sd = (const char*)s;
  char* td = (char*)t_out;
  if (mem == GH_MEM_HOST) {
    TRY(check_canonical_host(ring, s, (size_t)r * n * d));
    void* dev;
    TRY(ctx_staging(ctx, s_bytes + t_bytes, &dev));
    TRY(h2d(ctx, dev, s, s_bytes));
    sd = (const char*)dev;
    td = (char*)dev + s_bytes;
  }
  if (domain == GH_DOMAIN_COEFF && r > 1) {
    // the r parts share the key: one batched launch reads A once per group of parts
    std::vector<gh_commit_item> items(r);
    for (uint32_t i = 0; i < r; i++)
      items[i] = gh_commit_item{sd + (size_t)i * n * d * wb, n, 0,
                                td + (size_t)i * key->rows * d * wb};
    TRY(launch_commit_batch(key, items.data(), r));
  } else {
    for (uint32_t i = 0; i < r; i++)
      TRY(launch_commit(key, sd + (size_t)i * n * d * wb, n, domain,
                        td + (size_t)i * key->rows * d * wb));
  }
  if (mem == GH_MEM_HOST) {
    TRY(d2h(ctx, t_out, td, t_bytes));
    return sync(ctx);
  }
  return GH_OK;
}

GH_API int gh_commit_batch(gh_key* key, const gh_commit_item* items, size_t count, int mem) {
  if (!key || (count && !items)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  const uint32_t cols = key->msg_len + key->open_len;
  for (size_t i = 0; i < count; i++) {
    if (!items[i].witness || !items[i].t_out) return GH_E_INVALID_ARGUMENT;
    if (items[i].len == 0 || items[i].len > cols) return GH_E_DIMENSION_MISMATCH;
  }
  if (count == 0) return GH_OK;
  gh_ring* ring = key->ring;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (mem == GH_MEM_DEVICE && count == 1) {
    // one item: gh_commit's path, its finalize held back the same way (so no
    // flush here: the launch issues a pending finalize itself)
    GH_HIP(hipSetDevice(ctx->device));
    return launch_commit(key, items[0].witness, items[0].len, GH_DOMAIN_COEFF, items[0].t_out,
                         ctx->tune.defer_finalize != 0);
  }
  TRY(set_device(ctx));
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t t_bytes = (size_t)key->rows * d * wb;
  if (mem == GH_MEM_DEVICE) return launch_commit_batch(key, items, count);
  size_t total = 0;
  for (size_t i = 0; i < count; i++) {
    TRY(check_canonical_host(ring, items[i].witness, (size_t)items[i].len * d));
    total += (size_t)items[i].len * d * wb;
  }
  void* dev;
  TRY(ctx_staging(ctx, total + count * t_bytes, &dev));
  char* sp = (char*)dev;
  char* tp = (char*)dev + total;
  std::vector<gh_commit_item> staged(count);
  for (size_t i = 0; i < count; i++) {
    const size_t b = (size_t)items[i].len * d * wb;
    TRY(h2d(ctx, sp, items[i].witness, b));
    staged[i] = gh_commit_item{sp, items[i].len, 0, tp + i * t_bytes};
    sp += b;
  }
  TRY(launch_commit_batch(key, staged.data(), count));
  for (size_t i = 0; i < count; i++) TRY(d2h(ctx, items[i].t_out, tp + i * t_bytes, t_bytes));
  return sync(ctx);
}

GH_API int gh_gram(gh_ring* ring, const void* s, uint32_t r, uint32_t n, int domain, void* g_out,
                   int mem) {
  if (!ring || (r && (!s || !g_out))) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  if (r == 0) return GH_OK;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes, d = ring->d;
  const size_t npairs = (size_t)r * (r + 1) / 2;
  const size_t g_bytes = npairs * d * wb;
  if (n == 0) {  // empty witness: every g_ij is zero
    if (mem == GH_MEM_HOST) { std::memset(g_out, 0, g_bytes); return GH_OK; }
    std::lock_guard<std::mutex> lk(ctx->mu);
    TRY(set_device(ctx));
    GH_HIP(hipMemsetAsync(g_out, 0, g_bytes, ctx->stream));
    return GH_OK;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (mem == GH_MEM_DEVICE) return launch_gram(ring, s, r, n, domain, g_out);
  const size_t s_bytes = (size_t)r * n * d * wb;
  TRY(check_canonical_host(ring, s, (size_t)r * n * d));
  void* dev;
  TRY(ctx_staging(ctx, s_bytes + g_bytes, &dev));
  TRY(h2d(ctx, dev, s, s_bytes));
  TRY(launch_gram(ring, dev, r, n, domain, (char*)dev + s_bytes));
  TRY(d2h(ctx, g_out, (char*)dev + s_bytes, g_bytes));
  return sync(ctx);
}

GH_API int gh_jl_project(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
                         uint32_t num_parts, uint64_t cols, void* p_modq, int64_t* p_exact,
                         int8_t* flat, int mem) {
  if (!ring || !seed || !p_modq || (num_parts && cols && !parts)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  // rows > 256 would need more than one keystream block per column per row
  // byte in LDS; the reference profile uses 256 (labrador/src/params/jl_profile.rs:9-41).
  if (rows == 0 || rows > 256) return GH_E_INVALID_ARGUMENT;
  if (cols % ring->d != 0) return GH_E_INVALID_ARGUMENT;
  if (p_exact && ring->kind != GH_RING_Q32) return GH_E_INVALID_ARGUMENT;
  // the int64 centred sum is exact below 2^26 polynomials
  if (p_exact && (uint64_t)num_parts * (cols / ring->d) >= kJlExactMaxPolys)
    return GH_E_INVALID_ARGUMENT;
  if (cols > 0xffffffffull / ((rows + 3) / 4)) return GH_E_INVALID_ARGUMENT;  // 32-bit block counter
  if (num_parts > 65535) return GH_E_INVALID_ARGUMENT;  // one grid row per part
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const size_t in_bytes = (size_t)num_parts * cols * wb;
  const size_t p_bytes = (size_t)rows * wb, e_bytes = (size_t)rows * sizeof(int64_t);
  const size_t f_bytes = flat ? (size_t)rows * num_parts * cols : 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (num_parts == 0 || cols == 0) {
    if (mem == GH_MEM_HOST) {
      std::memset(p_modq, 0, p_bytes);
      if (p_exact) std::memset(p_exact, 0, e_bytes);
      return GH_OK;
    }
    GH_HIP(hipMemsetAsync(p_modq, 0, p_bytes, ctx->stream));
    if (p_exact) GH_HIP(hipMemsetAsync(p_exact, 0, e_bytes, ctx->stream));
    return GH_OK;
  }
  if (mem == GH_MEM_DEVICE)
    return launch_jl(ring, seed, rows, parts, num_parts, cols, p_modq, p_exact, flat);
  TRY(check_canonical_host(ring, parts, (size_t)num_parts * cols));
  void* dev;
  TRY(ctx_staging(ctx, in_bytes + p_bytes + e_bytes + f_bytes + 64, &dev));
  char* base = (char*)dev;
  char* dp = base + in_bytes;
  int64_t* de = (int64_t*)(((uintptr_t)(dp + p_bytes) + 7) & ~(uintptr_t)7);
  int8_t* df = flat ? (int8_t*)(de + rows) : nullptr;
  TRY(h2d(ctx, base, parts, in_bytes));
  TRY(launch_jl(ring, seed, rows, base, num_parts, cols, dp, p_exact ? de : nullptr, df));
  TRY(d2h(ctx, p_modq, dp, p_bytes));
  if (p_exact) TRY(d2h(ctx, p_exact, de, e_bytes));
  if (flat) TRY(d2h(ctx, flat, df, f_bytes));
  return sync(ctx);
}

GH_API int gh_jl_launch_shape(int kernel, uint64_t total_polys, int num_cus,
                              gh_jl_shape* out) {
  if (!out || num_cus < 1) return GH_E_INVALID_ARGUMENT;
  if (total_polys > kJlMaxPolys) return GH_E_INVALID_ARGUMENT;
  JlShape s;
  if (kernel == GH_JL_KERNEL_PROJECT) jl_proj_shape(total_polys, num_cus, &s);
  else if (kernel == GH_JL_KERNEL_EVALS) jl_evals_shape(total_polys, num_cus, &s);
  else return GH_E_INVALID_ARGUMENT;
  out->chunk_polys = s.chunk_polys;
  out->per_unit = s.per_unit;
  out->blocks = s.blocks;
  out->chunks = s.chunks;
  out->terms = s.terms;
  out->bound = s.bound;
  return GH_OK;
}

}  // extern "C"

// ---- canonical wire formats -------------------------------------------------
// Headers are read and written here; element runs move through the device
// pack/unpack kernels (gh_wire.hip).

namespace {

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

struct WireKind {
  uint32_t hdr;   // header bytes
  bool per_poly;  // items are polynomials (else bare field elements)
};

int wire_kind(int kind, WireKind* k) {
  switch (kind) {
    case GH_WIRE_POLYS: *k = {0, true}; return GH_OK;
    case GH_WIRE_RINGVEC: *k = {8, true}; return GH_OK;
    case GH_WIRE_RINGMAT: *k = {16, true}; return GH_OK;
    case GH_WIRE_VEC: *k = {4, true}; return GH_OK;
    case GH_WIRE_ELEMS: *k = {4, false}; return GH_OK;
    default: return GH_E_INVALID_ARGUMENT;
  }
}

uint64_t get_le(const uint8_t* p, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}
void put_le(uint8_t* p, uint64_t v, int n) {
  for (int i = 0; i < n; i++) p[i] = (uint8_t)(v >> (8 * i));
}

// elements and payload bytes of rows x cols items; GH_E_INVALID_DATA when the
// shape overflows (RingMat's checked_mul, lattice/types.rs:464-466), as a
// 63-bit byte count bounds what the reference could reserve
int wire_shape(const gh_ring* ring, const WireKind& k, uint64_t rows, uint64_t cols,
               uint64_t* elems, uint64_t* bytes) {
  const unsigned __int128 items = (unsigned __int128)rows * cols;
  const unsigned __int128 e = items * (k.per_poly ? ring->d : 1);
  const unsigned __int128 b = e * ring->wire_bytes + k.hdr;
  if (b >> 63) return GH_E_INVALID_DATA;
  *elems = (uint64_t)e;
  *bytes = (uint64_t)b - k.hdr;
  return GH_OK;
}

void write_header(int kind, uint8_t* out, uint64_t rows, uint64_t cols) {
  switch (kind) {
    case GH_WIRE_RINGVEC: put_le(out, rows, 8); break;
    case GH_WIRE_RINGMAT: put_le(out, rows, 8); put_le(out + 8, cols, 8); break;
    case GH_WIRE_VEC:
    case GH_WIRE_ELEMS: put_le(out, rows, 4); break;
    default: break;
  }
}

// Reads the header at p; rows and cols come back (GH_WIRE_POLYS keeps *rows).
int read_header(int kind, const uint8_t* p, uint64_t len, uint64_t* rows, uint64_t* cols) {
  WireKind k;
  TRY(wire_kind(kind, &k));
  if (len < k.hdr) return GH_E_UNEXPECTED_END;
  switch (kind) {
    case GH_WIRE_RINGVEC: *rows = get_le(p, 8); *cols = 1; break;
    case GH_WIRE_RINGMAT: *rows = get_le(p, 8); *cols = get_le(p + 8, 8); break;
    case GH_WIRE_VEC:
    case GH_WIRE_ELEMS: *rows = get_le(p, 4); *cols = 1; break;
    default: *cols = 1; break;
  }
  return GH_OK;
}

// A run of `count` elements at byte offset `off` of an encoded image, to or
// from `words` (device; null: check only).
struct Run {
  uint64_t off, count;
  void* words;
};

// Elements a truncated payload of `avail` bytes still checks before the
// reference reports UnexpectedEnd: every whole element of a coefficient-domain
// polynomial (CyclotomicPolyRing decodes element by element, poly/ring.rs:488-504)
// or field-element vector, but only whole polynomials in the NTT domain
// (TwistedNttPoly checks its length first, poly/twisted_ntt.rs:663-684).
uint64_t truncated_prefix(const gh_ring* ring, const WireKind& k, int domain, uint64_t avail) {
  const uint64_t n = avail / ring->wire_bytes;
  return (k.per_poly && domain == GH_DOMAIN_NTT) ? n / ring->d * ring->d : n;
}

// Unpacks runs of `bytes` (host; the first `nbytes` are copied) on the device;
// *bad reports an element >= q. Context lock held.
int unpack_runs(gh_ring* ring, const uint8_t* bytes, uint64_t nbytes, const std::vector<Run>& runs,
                bool* bad) {
  gh_ctx* ctx = ring->ctx;
  *bad = false;
  std::vector<WireSeg> segs;
  for (const Run& r : runs)
    if (r.count) segs.push_back(WireSeg{r.off, (uint64_t)(uintptr_t)r.words, r.count, 0, 0});
  if (segs.empty()) return GH_OK;
  std::vector<WireChunk> chunks;
  wire_chunks(segs, &chunks);
  if (chunks.size() > 0xffffffffull) return GH_E_INVALID_ARGUMENT;
  const size_t o_seg = 0, o_chk = al(segs.size() * sizeof(WireSeg));
  const size_t o_in = o_chk + al(chunks.size() * sizeof(WireChunk));
  void* ws;
  TRY(ctx_workspace(ctx, o_in + nbytes, &ws));
  char* w = (char*)ws;
  for (WireSeg& s : segs) s.bytes += (uint64_t)(uintptr_t)(w + o_in);
  uint32_t* flag = (uint32_t*)ctx->misc;
  TRY(h2d(ctx, w + o_in, bytes, nbytes));
  TRY(h2d(ctx, w + o_seg, segs.data(), segs.size() * sizeof(WireSeg)));
  TRY(h2d(ctx, w + o_chk, chunks.data(), chunks.size() * sizeof(WireChunk)));
  GH_HIP(hipMemsetAsync(flag, 0, sizeof(uint32_t), ctx->stream));
  TRY(enqueue_wire_unpack(ring, ctx->stream, (const WireSeg*)(w + o_seg),
                          (const WireChunk*)(w + o_chk), (uint32_t)chunks.size(), flag));
  uint32_t host = 0;
  TRY(d2h(ctx, &host, flag, sizeof host));
  TRY(sync(ctx));
  *bad = host != 0;
  return GH_OK;
}

// Packs runs (device words) into an image of `nbytes` on the device and copies
// it to `out`; headers are written by the caller afterwards. Context lock held.
int pack_runs(gh_ring* ring, const std::vector<Run>& runs, uint64_t nbytes, uint8_t* out) {
  gh_ctx* ctx = ring->ctx;
  std::vector<WireSeg> segs;
  for (const Run& r : runs)
    if (r.count) segs.push_back(WireSeg{r.off, (uint64_t)(uintptr_t)r.words, r.count, 0, 0});
  if (segs.empty()) return GH_OK;
  std::vector<WireChunk> chunks;
  wire_chunks(segs, &chunks);
  if (chunks.size() > 0xffffffffull) return GH_E_INVALID_ARGUMENT;
  const size_t o_seg = 0, o_chk = al(segs.size() * sizeof(WireSeg));
  const size_t o_out = o_chk + al(chunks.size() * sizeof(WireChunk));
  void* ws;
  TRY(ctx_workspace(ctx, o_out + nbytes, &ws));
  char* w = (char*)ws;
  for (WireSeg& s : segs) s.bytes += (uint64_t)(uintptr_t)(w + o_out);
  TRY(h2d(ctx, w + o_seg, segs.data(), segs.size() * sizeof(WireSeg)));
  TRY(h2d(ctx, w + o_chk, chunks.data(), chunks.size() * sizeof(WireChunk)));
  TRY(enqueue_wire_pack(ring, ctx->stream, (const WireSeg*)(w + o_seg),
                        (const WireChunk*)(w + o_chk), (uint32_t)chunks.size()));
  TRY(d2h(ctx, out, w + o_out, nbytes));
  return sync(ctx);
}

}  // namespace

extern "C" {

GH_API int gh_wire_size(const gh_ring* ring, int kind, uint64_t rows, uint64_t cols,
                        uint64_t* bytes) {
  if (!ring || !bytes) return GH_E_INVALID_ARGUMENT;
  WireKind k;
  TRY(wire_kind(kind, &k));
  if (kind != GH_WIRE_RINGMAT && cols != 1) return GH_E_INVALID_ARGUMENT;
  // Vec lengths are u32 (serialize/src/lib.rs:217-218)
  if ((kind == GH_WIRE_VEC || kind == GH_WIRE_ELEMS) && rows > 0xffffffffull)
    return GH_E_INVALID_DATA;
  uint64_t elems, payload;
  TRY(wire_shape(ring, k, rows, cols, &elems, &payload));
  *bytes = k.hdr + payload;
  return GH_OK;
}

GH_API int gh_wire_encode(gh_ring* ring, int kind, uint64_t rows, uint64_t cols, const void* src,
                          int mem, void* out, uint64_t out_cap, uint64_t* written) {
  if (!ring || !written) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  uint64_t size;
  TRY(gh_wire_size(ring, kind, rows, cols, &size));
  *written = size;
  if (!out) return GH_OK;
  if (out_cap < size) return GH_E_INVALID_ARGUMENT;
  WireKind k;
  TRY(wire_kind(kind, &k));
  uint64_t elems, payload;
  TRY(wire_shape(ring, k, rows, cols, &elems, &payload));
  if (elems && !src) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  uint8_t* o = (uint8_t*)out;
  if (elems) {
    const void* words = src;
    if (mem == GH_MEM_HOST) {
      TRY(check_canonical_host(ring, src, (size_t)elems));
      void* dev;
      const size_t wbytes = (size_t)elems * ring->word_bytes;
      TRY(ctx_staging(ctx, wbytes, &dev));
      TRY(h2d(ctx, dev, src, wbytes));
      words = dev;
    }
    TRY(pack_runs(ring, {Run{0, elems, const_cast<void*>(words)}}, payload, o + k.hdr));
  }
  write_header(kind, o, rows, cols);
  return GH_OK;
}

GH_API int gh_wire_peek(const gh_ring* ring, int kind, const void* bytes, uint64_t len,
                        uint64_t* rows, uint64_t* cols, uint64_t* total) {
  if (!ring || !rows || !cols || (len && !bytes)) return GH_E_INVALID_ARGUMENT;
  WireKind k;
  TRY(wire_kind(kind, &k));
  uint64_t r = *rows, c = 1;
  TRY(read_header(kind, (const uint8_t*)bytes, len, &r, &c));
  uint64_t elems, payload;
  TRY(wire_shape(ring, k, r, c, &elems, &payload));
  *rows = r;
  *cols = c;
  if (total) *total = k.hdr + payload;
  return GH_OK;
}

GH_API int gh_wire_decode(gh_ring* ring, int kind, int domain, const void* bytes, uint64_t len,
                          uint64_t rows, uint64_t cols, void* dst, int mem, uint64_t* consumed) {
  if (!ring || (len && !bytes)) return GH_E_INVALID_ARGUMENT;
  TRY(check_mem(mem));
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  WireKind k;
  TRY(wire_kind(kind, &k));
  const uint8_t* in = (const uint8_t*)bytes;
  uint64_t r = rows, c = 1;
  TRY(read_header(kind, in, len, &r, &c));
  uint64_t elems, payload;
  TRY(wire_shape(ring, k, r, c, &elems, &payload));
  if (r != rows || c != cols) return GH_E_DIMENSION_MISMATCH;
  if (elems && !dst) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  const uint64_t avail = len - k.hdr;
  bool bad = false;
  if (avail < payload) {
    const uint64_t n = truncated_prefix(ring, k, domain, avail);
    TRY(unpack_runs(ring, in, k.hdr + n * ring->wire_bytes, {Run{k.hdr, n, nullptr}}, &bad));
    return bad ? GH_E_INVALID_DATA : GH_E_UNEXPECTED_END;
  }
  const size_t wbytes = (size_t)elems * ring->word_bytes;
  void* words = dst;
  if (mem == GH_MEM_HOST && elems) TRY(ctx_staging(ctx, wbytes, &words));
  TRY(unpack_runs(ring, in, k.hdr + payload, {Run{k.hdr, elems, words}}, &bad));
  if (bad) return GH_E_INVALID_DATA;
  // deserialize_exact (serialize/src/lib.rs:79-83)
  if (!consumed && len != k.hdr + payload) return GH_E_INVALID_DATA;
  if (mem == GH_MEM_HOST && elems) {
    TRY(d2h(ctx, dst, words, wbytes));
    TRY(sync(ctx));
  }
  if (consumed) *consumed = k.hdr + payload;
  return GH_OK;
}

GH_API int gh_key_save(gh_key* key, int domain, void* out, uint64_t out_cap, uint64_t* written) {
  if (!key || !written) return GH_E_INVALID_ARGUMENT;
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  gh_ring* ring = key->ring;
  const uint64_t d = ring->d, eb = ring->wire_bytes, wb = ring->word_bytes;
  const uint64_t m1 = (uint64_t)key->rows * key->msg_len, m2 = (uint64_t)key->rows * key->open_len;
  const uint64_t b1 = m1 * d * eb, b2 = m2 * d * eb;
  const uint64_t size = 16 + b1 + 16 + b2;
  *written = size;
  if (!out) return GH_OK;
  if (out_cap < size) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  void* dev;
  TRY(ctx_staging(ctx, (size_t)(m1 + m2) * d * wb, &dev));
  char* msg = (char*)dev;
  char* open = msg + m1 * d * wb;
  TRY(launch_key_rowmajor(ring, key->dev_a, key->rows, key->msg_len, 0, msg));
  TRY(launch_key_rowmajor(ring, key->dev_a, key->rows, key->open_len, key->msg_len, open));
  // finish_twisted_ring_mat (poly/twisted_ntt.rs:318-330): back to coefficients
  if (domain == GH_DOMAIN_COEFF) TRY(launch_xform(ring, kTwistedInv, dev, dev, (size_t)(m1 + m2)));
  uint8_t* o = (uint8_t*)out;
  TRY(pack_runs(ring, {Run{16, m1 * d, msg}, Run{16 + b1 + 16, m2 * d, open}}, size, o));
  write_header(GH_WIRE_RINGMAT, o, key->rows, key->msg_len);
  write_header(GH_WIRE_RINGMAT, o + 16 + b1, key->rows, key->open_len);
  return GH_OK;
}

GH_API int gh_key_load(gh_ring* ring, const void* bytes, uint64_t len, int domain,
                       uint64_t* consumed, gh_key** out) {
  if (!ring || !out || (len && !bytes)) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  if (domain != GH_DOMAIN_COEFF && domain != GH_DOMAIN_NTT) return GH_E_INVALID_ARGUMENT;
  const WireKind k{16, true};
  const uint8_t* in = (const uint8_t*)bytes;
  const uint64_t eb = ring->wire_bytes, wb = ring->word_bytes;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  // LinearCommitmentKey::deserialize (commit/src/linear.rs:181-186): a_msg then
  // a_open, each RingMat::deserialize; a truncated or invalid stream reports
  // what the first failing element or header would.
  uint64_t rows1 = 0, cols1 = 1;
  TRY(read_header(GH_WIRE_RINGMAT, in, len, &rows1, &cols1));
  uint64_t e1, b1;
  TRY(wire_shape(ring, k, rows1, cols1, &e1, &b1));
  bool bad = false;
  if (len - 16 < b1) {
    const uint64_t n = truncated_prefix(ring, k, domain, len - 16);
    TRY(unpack_runs(ring, in, 16 + n * eb, {Run{16, n, nullptr}}, &bad));
    return bad ? GH_E_INVALID_DATA : GH_E_UNEXPECTED_END;
  }
  const uint64_t off2 = 16 + b1;
  uint64_t rows2 = 0, cols2 = 1, e2 = 0, b2 = 0;
  int st2 = read_header(GH_WIRE_RINGMAT, in + off2, len - off2, &rows2, &cols2);
  if (st2 == GH_OK) st2 = wire_shape(ring, k, rows2, cols2, &e2, &b2);
  if (st2 != GH_OK || len - off2 - 16 < b2) {
    uint64_t n2 = 0;
    if (st2 == GH_OK) n2 = truncated_prefix(ring, k, domain, len - off2 - 16);
    TRY(unpack_runs(ring, in, off2 + (n2 ? 16 + n2 * eb : 0),
                    {Run{16, e1, nullptr}, Run{off2 + 16, n2, nullptr}}, &bad));
    if (bad) return GH_E_INVALID_DATA;
    return st2 != GH_OK ? st2 : GH_E_UNEXPECTED_END;
  }
  const uint64_t total = off2 + 16 + b2;
  void* dev = nullptr;
  const size_t w1 = (size_t)e1 * wb, w2 = (size_t)e2 * wb;
  if (e1 + e2) TRY(ctx_staging(ctx, w1 + w2, &dev));
  char* msg = (char*)dev;
  char* open = msg ? msg + w1 : nullptr;
  TRY(unpack_runs(ring, in, total, {Run{16, e1, msg}, Run{off2 + 16, e2, open}}, &bad));
  if (bad) return GH_E_INVALID_DATA;
  if (!consumed && len != total) return GH_E_INVALID_DATA;
  // Valid for LinearCommitmentKey (linear.rs:188-192): equal row counts
  if (rows1 != rows2) return GH_E_INVALID_DATA;
  // CommitmentDimensions::validate (linear.rs:45-59), and the key's 32-bit dimensions
  if (rows1 == 0 || cols1 == 0) return GH_E_INVALID_PARAMETERS;
  if (rows1 > 0xffffffffull || cols1 + cols2 > 0xffffffffull) return GH_E_INVALID_PARAMETERS;
  TRY(key_create_dev(ring, (uint32_t)rows1, msg, (uint32_t)cols1, open, (uint32_t)cols2, domain,
                     out));
  if (consumed) *consumed = total;
  return GH_OK;
}

}  // extern "C"

// ---- JL rows for aggregation --------------------------------------------------

namespace {

// shared argument checks of the JL entry points (gh_jl_project's limits)
int jl_args(gh_ring* ring, const uint8_t* seed, uint32_t rows, uint64_t cols,
            uint32_t num_parts) {
  if (!ring || !seed) return GH_E_INVALID_ARGUMENT;
  if (num_parts > 65535) return GH_E_INVALID_ARGUMENT;  // one grid row per part
  if (rows == 0 || rows > 256) return GH_E_INVALID_ARGUMENT;
  if (ring->d != 64 || cols % ring->d != 0) return GH_E_INVALID_ARGUMENT;
  if (cols > 0xffffffffull / ((rows + 3) / 4)) return GH_E_INVALID_ARGUMENT;
  return GH_OK;
}

}  // namespace

extern "C" {

GH_API int gh_jl_extract_flat(gh_ring* ring, const uint8_t seed[32], uint32_t rows,
                              uint32_t num_parts, uint64_t cols, int8_t* flat, int mem) {
  TRY(jl_args(ring, seed, rows, cols, num_parts));
  TRY(check_mem(mem));
  const size_t bytes = (size_t)rows * num_parts * cols;
  if (bytes == 0) return GH_OK;
  if (!flat) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (mem == GH_MEM_DEVICE) return launch_jl_flat(ring, seed, rows, num_parts, cols, flat);
  void* dev;
  TRY(ctx_staging(ctx, bytes, &dev));
  TRY(launch_jl_flat(ring, seed, rows, num_parts, cols, (int8_t*)dev));
  TRY(d2h(ctx, flat, dev, bytes));
  return sync(ctx);
}

GH_API int gh_jl_conjugate(gh_ring* ring, const uint8_t seed[32], const int8_t* flat,
                           uint32_t rows, uint32_t num_parts, uint64_t cols, void* polys,
                           int mem) {
  if (!ring || (!flat && !seed)) return GH_E_INVALID_ARGUMENT;
  static const uint8_t kNoSeed[32] = {0};
  TRY(jl_args(ring, flat ? kNoSeed : seed, rows, cols, num_parts));
  TRY(check_mem(mem));
  const uint64_t n = cols / ring->d;
  const size_t f_bytes = (size_t)rows * num_parts * cols;
  const size_t o_bytes = (size_t)rows * num_parts * n * ring->d * ring->word_bytes;
  if (o_bytes == 0) return GH_OK;
  if (!polys) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (mem == GH_MEM_DEVICE)
    return flat ? launch_jl_conj_flat(ring, flat, rows, num_parts, n, polys)
                : launch_jl_conj_seed(ring, seed, rows, num_parts, n, polys);
  void* dev;
  TRY(ctx_staging(ctx, o_bytes + (flat ? f_bytes : 0), &dev));
  char* out = (char*)dev;
  if (flat) {
    int8_t* fd = (int8_t*)(out + o_bytes);
    TRY(h2d(ctx, fd, flat, f_bytes));
    TRY(launch_jl_conj_flat(ring, fd, rows, num_parts, n, out));
  } else {
    TRY(launch_jl_conj_seed(ring, seed, rows, num_parts, n, out));
  }
  TRY(d2h(ctx, polys, out, o_bytes));
  return sync(ctx);
}

GH_API int gh_jl_evals(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
                       uint32_t num_parts, uint64_t cols, void* evals, int mem) {
  TRY(jl_args(ring, seed, rows, cols, num_parts));
  TRY(check_mem(mem));
  if (rows != 256 || !evals) return GH_E_INVALID_ARGUMENT;
  if (num_parts && cols && !parts) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const size_t e_bytes = (size_t)256 * ring->d * wb;
  const uint64_t n = cols / ring->d;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (num_parts == 0 || n == 0) {
    if (mem == GH_MEM_HOST) {
      std::memset(evals, 0, e_bytes);
      return GH_OK;
    }
    GH_HIP(hipMemsetAsync(evals, 0, e_bytes, ctx->stream));
    return GH_OK;
  }
  if (mem == GH_MEM_DEVICE) return launch_jl_evals(ring, seed, parts, num_parts, n, evals);
  const size_t in_bytes = (size_t)num_parts * cols * wb;
  TRY(check_canonical_host(ring, parts, (size_t)num_parts * cols));
  void* dev;
  TRY(ctx_staging(ctx, in_bytes + e_bytes, &dev));
  TRY(h2d(ctx, dev, parts, in_bytes));
  TRY(launch_jl_evals(ring, seed, dev, num_parts, n, (char*)dev + in_bytes));
  TRY(d2h(ctx, evals, (char*)dev + in_bytes, e_bytes));
  return sync(ctx);
}

GH_API int gh_jl_aggregate(gh_ring* ring, const uint8_t seed[32], uint32_t rows,
                           const void* scale, uint32_t num_parts, uint64_t cols, void* phi,
                           int mem) {
  TRY(jl_args(ring, seed, rows, cols, num_parts));
  TRY(check_mem(mem));
  if (rows != 256 || !scale) return GH_E_INVALID_ARGUMENT;
  const uint64_t n = cols / ring->d;
  const size_t words = (size_t)num_parts * n * ring->d;
  if (words == 0) return GH_OK;
  if (!phi) return GH_E_INVALID_ARGUMENT;
  gh_ctx* ctx = ring->ctx;
  const size_t wb = ring->word_bytes;
  const size_t s_bytes = (size_t)256 * ring->d * wb, p_bytes = words * wb;
  std::lock_guard<std::mutex> lk(ctx->mu);
  TRY(set_device(ctx));
  if (mem == GH_MEM_DEVICE) return launch_jl_aggregate(ring, seed, scale, num_parts, n, phi);
  TRY(check_canonical_host(ring, scale, (size_t)256 * ring->d));
  TRY(check_canonical_host(ring, phi, words));
  void* dev;
  TRY(ctx_staging(ctx, s_bytes + p_bytes, &dev));
  char* ds = (char*)dev;
  char* dp = ds + s_bytes;
  TRY(h2d(ctx, ds, scale, s_bytes));
  TRY(h2d(ctx, dp, phi, p_bytes));
  TRY(launch_jl_aggregate(ring, seed, ds, num_parts, n, dp));
  TRY(d2h(ctx, phi, dp, p_bytes));
  return sync(ctx);
}

}  // extern "C"
