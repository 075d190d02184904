// Serving loop with continuous batching (gh_server_* in include/gridhip.h).
//
// One worker thread per server owns the device side:
//   - submitters (any thread) validate a request, compute its HBM footprint and
//     queue it;
//   - the worker admits queued requests in arrival order while the arena has
//     room and fewer than max_inflight are admitted, and launches everything it
//     admitted at once as a wave on a free slot (stream): one segmented witness
//     transform per ring, one grouped MAC + batched finalize per key, one Gram
//     sweep + finalize per ring;
//   - a wave's completion event releases its requests' arena blocks, copies
//     host-memory outputs out of the slot's pinned staging and wakes waiters.
// With two or more slots the next wave is enqueued while the previous one is
// still on the device, and its kernels start as the previous wave's tail
// drains. Requests that arrive meanwhile are admitted into the next wave.
//
// The arithmetic is the batched path of gh_commit_batch and the Gram kernels;
// every partial sum is an exact residue, so a result does not depend on which
// requests share its wave or on the chunking.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gh_internal.h"

using namespace gh;

namespace {

using Clock = std::chrono::steady_clock;

inline size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// First-fit allocator over offsets [0, size) with coalescing free list.
class Arena {
 public:
  void init(size_t size) {
    free_.clear();
    free_[0] = size;
    size_ = size;
    used_ = 0;
  }
  size_t alloc(size_t bytes) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        return off;
      }
    }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      free_.erase(next);
    }
    free_[off] = bytes;
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

struct SrvReq {
  uint64_t ticket = 0;
  gh_request rq{};
  // arena block and the offsets of its parts
  size_t foot = 0, off = 0;
  size_t o_in = 0, o_sntt = 0, o_cpart = 0, o_tout = 0, o_gpart = 0, o_gout = 0;
  uint32_t nch = 0;                  // commit: key-column chunks per part
  uint32_t g_chunk = 0, g_nchunks = 0;
  size_t t_bytes = 0, g_bytes = 0;   // output sizes
  size_t stage_t = 0, stage_g = 0;   // host-memory outputs: offsets in the slot's pinned buffer
  size_t stage_f = 0;                // wire witness: its flag in the slot's pinned buffer
  bool bad = false;                  // wire witness held an element >= q
  int status = GH_OK;
  bool done = false;
  Clock::time_point t_sub, t_done;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* ddesc = nullptr;  // device copy of the wave's descriptors
  size_t ddesc_bytes = 0;
  void* hbuf = nullptr;   // pinned: descriptor image, then host-output staging
  size_t hbuf_bytes = 0;
  std::vector<SrvReq*> reqs;
  bool busy = false;
};

}  // namespace

struct gh_server {
  gh_ctx* ctx = nullptr;
  gh_server_config cfg{};
  uint32_t cc[2] = {256, 256};  // batched MAC chunking for q32 and Goldilocks keys, fixed
  uint32_t chunk(const gh_ring* ring) const { return cc[ring->kind == GH_RING_Q32 ? 0 : 1]; }
  int group = 8, variant = 1;
  char* base = nullptr;  // arena
  Arena arena;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvReq*> pending;
  std::unordered_map<uint64_t, SrvReq*> table;
  uint64_t next_ticket = 1;
  uint32_t admitted = 0;
  bool stop = false;
  gh_server_stats st{};
  std::thread worker;
};

namespace {

// Sizes the request's arena block; GH_OK or an error for an impossible request.
int plan_request(gh_server* s, SrvReq* q) {
  const gh_request& r = q->rq;
  const size_t wb = r.ring->word_bytes;
  const uint64_t m = (uint64_t)r.r * r.n * 64;  // witness words
  size_t at = 0;
  if (r.mem == GH_MEM_HOST) {  // device copy of the host witness (wire: its encoding)
    q->o_in = at;
    at += al(r.wire ? r.witness_len : m * wb);
  }
  q->o_sntt = at;
  at += al(m * wb);
  if (r.ops & GH_OP_COMMIT) {
    const uint32_t rows = r.key->rows;
    const uint32_t cc = s->chunk(r.ring);
    q->nch = (r.n + cc - 1) / cc;
    q->o_cpart = at;
    at += al((size_t)r.r * q->nch * rows * 64 * wb);
    q->t_bytes = (size_t)r.r * rows * 64 * wb;
    if (r.mem == GH_MEM_HOST) {
      q->o_tout = at;
      at += al(q->t_bytes);
    }
  }
  if (r.ops & GH_OP_GRAM) {
    const uint64_t npairs = (uint64_t)r.r * (r.r + 1) / 2;
    gram_chunking(s->ctx->num_cus, r.r, r.n, &q->g_chunk, &q->g_nchunks);
    q->o_gpart = at;
    at += al((size_t)q->g_nchunks * npairs * 64 * wb);
    q->g_bytes = (size_t)npairs * 64 * wb;
    if (r.mem == GH_MEM_HOST) {
      q->o_gout = at;
      at += al(q->g_bytes);
    }
  }
  q->foot = at;
  return at > s->arena.size() ? GH_E_OOM : GH_OK;
}

uint64_t get_le(const uint8_t* p, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}
void put_le(uint8_t* p, uint64_t v, int n) {
  for (int i = 0; i < n; i++) p[i] = (uint8_t)(v >> (8 * i));
}

// bytes of one encoded polynomial
uint64_t wire_poly(const gh_ring* g) { return (uint64_t)g->d * g->wire_bytes; }
// byte offset of part i's polynomials in a wire witness of parts of n polynomials
uint64_t wire_part(const gh_request& r, uint32_t i) {
  return 4 + (uint64_t)i * (4 + r.n * wire_poly(r.ring)) + 4;
}
// t_vecs as Vec<RingVec<poly>> (labrador/src/main_protocol/mod.rs:702-726)
uint64_t wire_t_bytes(const gh_request& r) {
  return 4 + (uint64_t)r.r * (8 + (uint64_t)r.key->rows * wire_poly(r.ring));
}
// the Gram polynomials as Vec<poly> (labrador/src/main_protocol/garbage.rs:76-93)
uint64_t wire_g_bytes(const gh_request& r) {
  return 4 + (uint64_t)r.r * (r.r + 1) / 2 * wire_poly(r.ring);
}

// Structure of a wire witness: LabradorWitness::parts, a Vec<Vec<poly>>
// (labrador/src/relation.rs:113-115; serialize/src/lib.rs:210-245), u32 part
// count, then per part a u32 length and its polynomials. Parts must share one
// length. Fills r and n (0 in the request) or checks them. Elements are
// checked on the device.
int parse_wire_witness(gh_request* r) {
  if (r->wire != 1 || r->mem != GH_MEM_HOST) return GH_E_INVALID_ARGUMENT;
  const gh_ring* g = r->ring;
  if (g->wire_bytes != g->word_bytes) return GH_E_INVALID_PARAMETERS;
  const uint8_t* p = (const uint8_t*)r->witness;
  const uint64_t len = r->witness_len, pb = wire_poly(g);
  if (len < 4) return GH_E_UNEXPECTED_END;
  const uint32_t parts = (uint32_t)get_le(p, 4);
  uint64_t off = 4, n = 0;
  for (uint32_t i = 0; i < parts; i++) {
    if (len - off < 4) return GH_E_UNEXPECTED_END;
    const uint32_t ni = (uint32_t)get_le(p + off, 4);
    if (i == 0) n = ni;
    else if (ni != n) return GH_E_DIMENSION_MISMATCH;
    off += 4;
    if ((len - off) / pb < ni) return GH_E_UNEXPECTED_END;
    off += ni * pb;
  }
  if (off != len) return GH_E_INVALID_DATA;  // deserialize_exact: trailing bytes
  if ((r->r && r->r != parts) || (r->n && r->n != n)) return GH_E_DIMENSION_MISMATCH;
  r->r = parts;
  r->n = (uint32_t)n;
  return GH_OK;
}

int validate(gh_server* s, gh_request* r) {
  if (r->ops == 0 || (r->ops & ~(uint32_t)(GH_OP_COMMIT | GH_OP_GRAM))) return GH_E_INVALID_ARGUMENT;
  if (r->mem != GH_MEM_HOST && r->mem != GH_MEM_DEVICE) return GH_E_INVALID_ARGUMENT;
  if (!r->ring || !r->witness || r->ring->ctx != s->ctx) return GH_E_INVALID_ARGUMENT;
  if (r->wire) {
    const int st = parse_wire_witness(r);
    if (st) return st;
  }
  if (r->r == 0 || r->n == 0) return GH_E_DIMENSION_MISMATCH;
  if ((uint64_t)r->r * r->n > 0xffffffffull || r->r > 0xffff) return GH_E_INVALID_ARGUMENT;
  if (r->ops & GH_OP_COMMIT) {
    if (!r->key || r->key->ring != r->ring || !r->t_out) return GH_E_INVALID_ARGUMENT;
    if (r->n > r->key->msg_len + r->key->open_len) return GH_E_DIMENSION_MISMATCH;
  }
  if ((r->ops & GH_OP_GRAM) && !r->g_out) return GH_E_INVALID_ARGUMENT;
  if (r->wire) {
    if ((r->ops & GH_OP_COMMIT) && r->t_cap < wire_t_bytes(*r)) return GH_E_INVALID_ARGUMENT;
    if ((r->ops & GH_OP_GRAM) && r->g_cap < wire_g_bytes(*r)) return GH_E_INVALID_ARGUMENT;
    return GH_OK;
  }
  if (r->mem == GH_MEM_HOST)
    return check_canonical_host(r->ring, r->witness, (size_t)r->r * r->n * 64);
  return GH_OK;
}

template <class T>
size_t put(std::vector<char>& img, const std::vector<T>& v) {
  const size_t at = img.size();
  img.resize(at + al(v.size() * sizeof(T)));
  if (!v.empty()) std::memcpy(img.data() + at, v.data(), v.size() * sizeof(T));
  return at;
}

int grow_slot(Slot& sl, size_t dbytes, size_t hbytes) {
  if (sl.ddesc_bytes < dbytes) {
    if (sl.ddesc) GH_HIP(hipFree(sl.ddesc));
    sl.ddesc = nullptr;
    sl.ddesc_bytes = 0;
    const size_t sz = std::max(al(dbytes), (size_t)1 << 20);
    GH_HIP(hipMalloc(&sl.ddesc, sz));
    sl.ddesc_bytes = sz;
  }
  if (sl.hbuf_bytes < hbytes) {
    if (sl.hbuf) GH_HIP(hipHostFree(sl.hbuf));
    sl.hbuf = nullptr;
    sl.hbuf_bytes = 0;
    const size_t sz = std::max(al(hbytes), (size_t)1 << 20);
    GH_HIP(hipHostMalloc(&sl.hbuf, sz, hipHostMallocDefault));
    sl.hbuf_bytes = sz;
  }
  return GH_OK;
}

// Enqueues one wave on the slot's stream. Called by the worker without the
// server lock; the requests belong to the worker until retired.
int launch_wave(gh_server* s, Slot& sl, const std::vector<SrvReq*>& wave) {
  struct RingPlan {
    gh_ring* ring;
    std::vector<BatchDesc> xd;
    std::vector<BatchSeg> seg;
    std::vector<GramDesc> gd;
    std::vector<GramWork> gw;
    std::vector<GramFin> gf;
    std::vector<WireSeg> ws;  // wire witnesses: one run per part
    std::vector<WireChunk> wc;
    size_t o_xd = 0, o_seg = 0, o_gd = 0, o_gw = 0, o_gf = 0, o_ws = 0, o_wc = 0;
  };
  struct KeyPlan {
    gh_key* key;
    std::vector<BatchDesc> md;
    size_t o_md = 0;
  };
  std::vector<RingPlan> rp;
  std::vector<KeyPlan> kp;
  auto ring_plan = [&](gh_ring* g) -> RingPlan& {
    for (RingPlan& p : rp)
      if (p.ring == g) return p;
    rp.push_back(RingPlan{g, {}, {}, {}, {}, {}, {}, {}});
    return rp.back();
  };
  auto key_plan = [&](gh_key* k) -> KeyPlan& {
    for (KeyPlan& p : kp)
      if (p.key == k) return p;
    kp.push_back(KeyPlan{k, {}});
    return kp.back();
  };
  size_t stage = 0;  // host-output staging bytes
  uint32_t nwire = 0;
  for (SrvReq* q : wave) {
    const gh_request& r = q->rq;
    const size_t wb = r.ring->word_bytes;
    char* blk = s->base + q->off;
    char* sntt = blk + q->o_sntt;
    uint64_t src = r.mem == GH_MEM_HOST ? (uint64_t)(uintptr_t)(blk + q->o_in)
                                        : (uint64_t)(uintptr_t)r.witness;
    RingPlan& p = ring_plan(r.ring);
    if (r.wire) {  // parts unpacked into the NTT buffer, then transformed in place
      for (uint32_t i = 0; i < r.r; i++)
        p.ws.push_back(WireSeg{(uint64_t)(uintptr_t)(blk + q->o_in + wire_part(r, i)),
                               (uint64_t)(uintptr_t)(sntt + (size_t)i * r.n * 64 * wb),
                               (uint64_t)r.n * 64, nwire, 0});
      q->stage_f = nwire++;
      src = (uint64_t)(uintptr_t)sntt;
    }
    const uint32_t len = r.r * r.n;
    const uint32_t xi = (uint32_t)p.xd.size();
    p.xd.push_back(BatchDesc{(uint64_t)(uintptr_t)sntt, src, 0, 0, len, 0});
    for (uint32_t f = 0; f < len; f += 64) p.seg.push_back(BatchSeg{xi, f});
    if (r.ops & GH_OP_COMMIT) {
      KeyPlan& k = key_plan(r.key);
      const uint32_t rows = r.key->rows;
      char* tout = r.mem == GH_MEM_HOST ? blk + q->o_tout : (char*)r.t_out;
      for (uint32_t i = 0; i < r.r; i++)
        k.md.push_back(BatchDesc{(uint64_t)(uintptr_t)(sntt + (size_t)i * r.n * 64 * wb), 0,
                                 (uint64_t)(uintptr_t)(tout + (size_t)i * rows * 64 * wb),
                                 (uint64_t)(uintptr_t)(blk + q->o_cpart +
                                                       (size_t)i * q->nch * rows * 64 * wb),
                                 r.n, 0});
      if (r.mem == GH_MEM_HOST) {
        q->stage_t = stage;
        stage += al(q->t_bytes);
      }
    }
    if (r.ops & GH_OP_GRAM) {
      const uint32_t gi = (uint32_t)p.gd.size();
      char* gout = r.mem == GH_MEM_HOST ? blk + q->o_gout : (char*)r.g_out;
      p.gd.push_back(GramDesc{(uint64_t)(uintptr_t)sntt, (uint64_t)(uintptr_t)(blk + q->o_gpart),
                              (uint64_t)(uintptr_t)gout, r.r, r.n, q->g_chunk, q->g_nchunks});
      const uint32_t tiles = gram_tiles(r.r);
      for (uint32_t c = 0; c < q->g_nchunks; c++)
        for (uint32_t t = 0; t < tiles; t++) p.gw.push_back(GramWork{gi, t, c});
      const uint32_t npairs = r.r * (r.r + 1) / 2;
      for (uint32_t e = 0; e < npairs; e++) p.gf.push_back(GramFin{gi, e});
      if (r.mem == GH_MEM_HOST) {
        q->stage_g = stage;
        stage += al(q->g_bytes);
      }
    }
  }
  for (KeyPlan& k : kp)
    std::stable_sort(k.md.begin(), k.md.end(),
                     [](const BatchDesc& a, const BatchDesc& b) { return a.len > b.len; });
  // descriptor image: every array 256-byte aligned, then copied in one transfer
  std::vector<char> img;
  for (RingPlan& p : rp) {
    p.o_xd = put(img, p.xd);
    p.o_seg = put(img, p.seg);
    p.o_gd = put(img, p.gd);
    p.o_gw = put(img, p.gw);
    p.o_gf = put(img, p.gf);
  }
  for (KeyPlan& k : kp) k.o_md = put(img, k.md);
  for (RingPlan& p : rp) {
    wire_chunks(p.ws, &p.wc);
    p.o_ws = put(img, p.ws);
    p.o_wc = put(img, p.wc);
  }
  const size_t o_flags = put(img, std::vector<uint32_t>(nwire, 0u));
  const size_t dbytes = img.size();
  const size_t stage_flags = stage;  // wire flags come back after the outputs
  stage += al((size_t)nwire * sizeof(uint32_t));
  int st = grow_slot(sl, dbytes, dbytes + stage);
  if (st) return st;
  char* hb = (char*)sl.hbuf;
  char* db = (char*)sl.ddesc;
  std::memcpy(hb, img.data(), dbytes);
  for (SrvReq* q : wave) {  // host-memory outputs land after the descriptor image
    q->stage_t += dbytes;
    q->stage_g += dbytes;
    if (q->rq.wire) q->stage_f = dbytes + stage_flags + q->stage_f * sizeof(uint32_t);
  }

  hipStream_t stream = sl.stream;
  GH_HIP(hipEventRecord(sl.ev0, stream));
  for (SrvReq* q : wave) {
    if (q->rq.mem != GH_MEM_HOST) continue;
    const size_t bytes = q->rq.wire ? q->rq.witness_len
                                    : (size_t)q->rq.r * q->rq.n * 64 * q->rq.ring->word_bytes;
    GH_HIP(hipMemcpyAsync(s->base + q->off + q->o_in, q->rq.witness, bytes, hipMemcpyHostToDevice,
                          stream));
  }
  GH_HIP(hipMemcpyAsync(db, hb, dbytes, hipMemcpyHostToDevice, stream));
  for (RingPlan& p : rp) {
    if (p.wc.empty()) continue;
    if (p.wc.size() > 0xffffffffull) return GH_E_INVALID_ARGUMENT;
    st = enqueue_wire_unpack(p.ring, stream, (const WireSeg*)(db + p.o_ws),
                             (const WireChunk*)(db + p.o_wc), (uint32_t)p.wc.size(),
                             (uint32_t*)(db + o_flags));
    if (st) return st;
  }
  for (RingPlan& p : rp) {
    st = enqueue_batch_xform(p.ring, stream, (const BatchDesc*)(db + p.o_xd),
                             (const BatchSeg*)(db + p.o_seg), (uint32_t)p.seg.size());
    if (st) return st;
  }
  for (KeyPlan& k : kp) {
    const BatchDesc* md = (const BatchDesc*)(db + k.o_md);
    const uint32_t count = (uint32_t)k.md.size();
    const uint32_t cc = s->chunk(k.key->ring);
    st = enqueue_batch_mac(k.key, stream, md, count, k.md[0].len, cc, s->group, s->variant);
    if (st) return st;
    st = enqueue_batch_finalize(k.key->ring, stream, md, count, k.key->rows, cc);
    if (st) return st;
  }
  for (RingPlan& p : rp) {
    if (p.gd.empty()) continue;
    st = enqueue_gram_batch(p.ring, stream, (const GramDesc*)(db + p.o_gd),
                            (const GramWork*)(db + p.o_gw), (uint32_t)p.gw.size(),
                            (const GramFin*)(db + p.o_gf), (uint32_t)p.gf.size());
    if (st) return st;
  }
  for (SrvReq* q : wave) {
    if (q->rq.mem != GH_MEM_HOST) continue;
    char* blk = s->base + q->off;
    if (q->rq.ops & GH_OP_COMMIT)
      GH_HIP(hipMemcpyAsync(hb + q->stage_t, blk + q->o_tout, q->t_bytes, hipMemcpyDeviceToHost,
                            stream));
    if (q->rq.ops & GH_OP_GRAM)
      GH_HIP(hipMemcpyAsync(hb + q->stage_g, blk + q->o_gout, q->g_bytes, hipMemcpyDeviceToHost,
                            stream));
  }
  if (nwire)
    GH_HIP(hipMemcpyAsync(hb + dbytes + stage_flags, db + o_flags, nwire * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, stream));
  GH_HIP(hipEventRecord(sl.ev1, stream));
  return GH_OK;
}

// Writes a wire request's outputs: headers, and the staged words, which are
// the encoded elements when the element width equals the word width.
void emit_wire(SrvReq* q, const char* hb) {
  const gh_request& r = q->rq;
  if (r.ops & GH_OP_COMMIT) {
    uint8_t* o = (uint8_t*)r.t_out;
    const uint32_t rows = r.key->rows;
    const size_t tb = (size_t)rows * 64 * r.ring->word_bytes;
    put_le(o, r.r, 4);
    size_t at = 4;
    for (uint32_t i = 0; i < r.r; i++) {
      put_le(o + at, rows, 8);
      std::memcpy(o + at + 8, hb + q->stage_t + i * tb, tb);
      at += 8 + tb;
    }
  }
  if (r.ops & GH_OP_GRAM) {
    uint8_t* o = (uint8_t*)r.g_out;
    put_le(o, (uint64_t)r.r * (r.r + 1) / 2, 4);
    std::memcpy(o + 4, hb + q->stage_g, q->g_bytes);
  }
}

// Completes a wave's requests (lock held by the caller).
void finish(gh_server* s, std::vector<SrvReq*>& reqs, int status) {
  const Clock::time_point now = Clock::now();
  for (SrvReq* q : reqs) {
    s->arena.release(q->off, q->foot);
    q->status = (status == GH_OK && q->bad) ? GH_E_INVALID_DATA : status;
    q->done = true;
    q->t_done = now;
    s->admitted--;
    if (q->status == GH_OK) {
      s->st.completed++;
      if (q->rq.ops & GH_OP_COMMIT) s->st.commits += q->rq.r;
      if (q->rq.ops & GH_OP_GRAM) s->st.grams++;
    } else {
      s->st.failed++;
    }
  }
  reqs.clear();
  s->st.mem_in_use = s->arena.used();
  s->cv_done.notify_all();
}

void run(gh_server* s) {
  hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  // Polling a wave's completion event gives microsecond latency but keeps a
  // core busy, so the worker polls only within kSpinWindow of its last
  // progress (a wave launched or completed) and past that sleeps in steps of
  // kSleepStep, woken at once by a new submission.
  constexpr auto kSpinWindow = std::chrono::microseconds(200);
  constexpr auto kSleepStep = std::chrono::microseconds(500);
  auto last_progress = std::chrono::steady_clock::now();
  for (;;) {
    bool progress = false;
    for (Slot& sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.ev1);
      if (e == hipErrorNotReady) continue;
      int status = e == hipSuccess ? GH_OK : GH_E_DEVICE;
      if (status == GH_OK) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, sl.ev0, sl.ev1) == hipSuccess) s->st.device_ms += ms;
        for (SrvReq* q : sl.reqs) {  // host-memory outputs leave the pinned staging
          if (q->rq.mem != GH_MEM_HOST) continue;
          if (q->rq.wire) {
            uint32_t flag;
            std::memcpy(&flag, (char*)sl.hbuf + q->stage_f, sizeof flag);
            q->bad = flag != 0;
            if (!q->bad) emit_wire(q, (const char*)sl.hbuf);
            continue;
          }
          if (q->rq.ops & GH_OP_COMMIT)
            std::memcpy(q->rq.t_out, (char*)sl.hbuf + q->stage_t, q->t_bytes);
          if (q->rq.ops & GH_OP_GRAM)
            std::memcpy(q->rq.g_out, (char*)sl.hbuf + q->stage_g, q->g_bytes);
        }
      }
      finish(s, sl.reqs, status);
      sl.busy = false;
      progress = true;
    }
    Slot* slot = nullptr;
    for (Slot& sl : s->slots)
      if (!sl.busy) {
        slot = &sl;
        break;
      }
    if (slot && !s->pending.empty()) {
      std::vector<SrvReq*> wave;
      while (!s->pending.empty() && wave.size() < s->cfg.max_wave &&
             s->admitted < s->cfg.max_inflight) {
        SrvReq* q = s->pending.front();
        const size_t off = s->arena.alloc(q->foot);
        if (off == SIZE_MAX) break;  // arrival order: the head waits for memory
        q->off = off;
        s->pending.pop_front();
        s->admitted++;
        wave.push_back(q);
      }
      if (!wave.empty()) {
        s->st.mem_in_use = s->arena.used();
        s->st.mem_peak = std::max<uint64_t>(s->st.mem_peak, s->st.mem_in_use);
        s->st.max_wave_seen = std::max<uint32_t>(s->st.max_wave_seen, (uint32_t)wave.size());
        s->st.waves++;
        slot->busy = true;
        lk.unlock();
        const int st = launch_wave(s, *slot, wave);
        lk.lock();
        slot->reqs = wave;
        if (st != GH_OK) {
          // nothing of a failed wave is waited on: drain what was enqueued first
          hipStreamSynchronize(slot->stream);
          finish(s, slot->reqs, st);
          slot->busy = false;
        }
        progress = true;
      }
    }
    bool busy = false;
    for (Slot& sl : s->slots) busy |= sl.busy;
    if (s->stop && s->pending.empty() && !busy) break;
    const auto now = std::chrono::steady_clock::now();
    if (progress) {
      last_progress = now;
    } else if (!busy) {
      s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
      last_progress = std::chrono::steady_clock::now();
    } else if (now - last_progress < kSpinWindow) {
      // a wave is on the device: poll its completion, staying responsive to
      // new submissions
      lk.unlock();
      for (int i = 0; i < 32; i++) std::this_thread::yield();
      lk.lock();
    } else {
      // a long wave: sleep between polls; a submission wakes the worker
      s->cv_work.wait_for(lk, kSleepStep);
    }
  }
}

void destroy_slots(gh_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.stream) hipStreamSynchronize(sl.stream);
    if (sl.ev0) hipEventDestroy(sl.ev0);
    if (sl.ev1) hipEventDestroy(sl.ev1);
    if (sl.ddesc) hipFree(sl.ddesc);
    if (sl.hbuf) hipHostFree(sl.hbuf);
    if (sl.stream) hipStreamDestroy(sl.stream);
  }
  s->slots.clear();
}

}  // namespace

extern "C" {

GH_API int gh_server_create(gh_ctx* ctx, const gh_server_config* cfg, gh_server** out) {
  if (!ctx || !out) return GH_E_INVALID_ARGUMENT;
  *out = nullptr;
  gh_server_config c = cfg ? *cfg : gh_server_config{};
  if (c.mem_bytes == 0) c.mem_bytes = (uint64_t)4 << 30;
  if (c.max_inflight == 0) c.max_inflight = 4096;
  if (c.max_wave == 0) c.max_wave = 1024;
  if (c.slots == 0) c.slots = 2;
  if (c.max_pending == 0) c.max_pending = 65536;
  if (c.slots > 4 || c.mem_bytes < 4096) return GH_E_INVALID_ARGUMENT;
  GH_HIP(hipSetDevice(ctx->device));
  gh_server* s = new (std::nothrow) gh_server();
  if (!s) return GH_E_OOM;
  s->ctx = ctx;
  s->cfg = c;
  s->cc[0] = tuned_batch_chunk(ctx->tune, GH_RING_Q32);
  s->cc[1] = tuned_batch_chunk(ctx->tune, GH_RING_GOLDILOCKS);
  s->group = ctx->tune.batch_group;
  s->variant = ctx->tune.mac_variant;
  if (hipMalloc((void**)&s->base, c.mem_bytes) != hipSuccess) {
    delete s;
    return GH_E_OOM;
  }
  s->arena.init(c.mem_bytes);
  s->slots.resize(c.slots);
  for (Slot& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&sl.ev0) != hipSuccess || hipEventCreate(&sl.ev1) != hipSuccess) {
      destroy_slots(s);
      hipFree(s->base);
      delete s;
      return GH_E_DEVICE;
    }
  }
  s->worker = std::thread(run, s);
  *out = s;
  return GH_OK;
}

GH_API int gh_server_destroy(gh_server* s) {
  if (!s) return GH_E_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();
  hipSetDevice(s->ctx->device);
  destroy_slots(s);
  hipFree(s->base);
  for (auto& kv : s->table) delete kv.second;
  delete s;
  return GH_OK;
}

GH_API int gh_server_submit(gh_server* s, const gh_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return GH_E_INVALID_ARGUMENT;
  SrvReq* q = new (std::nothrow) SrvReq();
  if (!q) return GH_E_OOM;
  q->rq = *req;
  int st = validate(s, &q->rq);
  if (st == GH_OK) st = plan_request(s, q);
  if (st) {
    delete q;
    return st;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete q;
      return GH_E_INVALID_ARGUMENT;
    }
    if (s->pending.size() >= s->cfg.max_pending) {
      delete q;
      return GH_E_BUSY;
    }
    q->ticket = s->next_ticket++;
    q->t_sub = Clock::now();
    s->table[q->ticket] = q;
    s->pending.push_back(q);
    s->st.submitted++;
    *ticket = q->ticket;
  }
  s->cv_work.notify_one();
  return GH_OK;
}

static void take(gh_server* s, SrvReq* q, int* status, double* latency_us) {
  if (status) *status = q->status;
  if (latency_us)
    *latency_us = std::chrono::duration<double, std::micro>(q->t_done - q->t_sub).count();
  s->table.erase(q->ticket);
  delete q;
}

GH_API int gh_server_wait(gh_server* s, uint64_t ticket, int* status, double* latency_us) {
  if (!s) return GH_E_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->table.find(ticket);
  if (it == s->table.end()) return GH_E_INVALID_ARGUMENT;
  SrvReq* q = it->second;
  s->cv_done.wait(lk, [&] { return q->done; });
  take(s, q, status, latency_us);
  return GH_OK;
}

GH_API int gh_server_poll(gh_server* s, uint64_t ticket, int* done, int* status,
                          double* latency_us) {
  if (!s || !done) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->table.find(ticket);
  if (it == s->table.end()) return GH_E_INVALID_ARGUMENT;
  SrvReq* q = it->second;
  *done = q->done ? 1 : 0;
  if (q->done) take(s, q, status, latency_us);
  return GH_OK;
}

GH_API int gh_server_query(gh_server* s, gh_server_stats* out) {
  if (!s || !out) return GH_E_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->pending = (uint32_t)s->pending.size();
  out->admitted = s->admitted;
  return GH_OK;
}

}  // extern "C"
