// Library-internal host structures shared by the C ABI and the launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/gridhip.h"
#include "gh_field.h"

namespace gh {
// Per-item descriptor of a batch (device copy next to the batch's buffers).
// All addresses are device addresses, so the items of one batch may live in
// unrelated allocations.
struct BatchDesc {
  uint64_t s_ntt;    // the item's NTT-domain witness, len x 64 words
  uint64_t src;      // the coefficient-domain witness (transform input)
  uint64_t dst;      // t_out: rows polynomials
  uint64_t partial;  // per-chunk partial sums, [chunk][row][slot]
  uint32_t len;      // witness polynomials (key columns used)
  uint32_t pad;
};
// a run of at most 64 polynomials of one item, transformed by one block
struct BatchSeg {
  uint32_t item;   // index into the descriptors of the transform
  uint32_t first;  // first polynomial of the segment within the item
};
// one Gram matrix of a batch: r parts of n NTT-domain polynomials
struct GramDesc {
  uint64_t s_ntt;    // r x n x 64 words
  uint64_t partial;  // [chunk][pair][slot]
  uint64_t dst;      // r(r+1)/2 polynomials, coefficient domain
  uint32_t r, n, chunk, nchunks;
};
struct GramWork {  // one wave of a Gram batch: (matrix, 4x4 pair tile, column chunk)
  uint32_t g, tile, chunk;
};
struct GramFin {  // one output polynomial of a Gram batch
  uint32_t g, pair;
};
// `count` field elements between an encoded byte stream and ring words (both
// device addresses). Unpacking raises flags[flag] on an element >= q; a
// segment with words == 0 is only checked.
struct WireSeg {
  uint64_t bytes;
  uint64_t words;
  uint64_t count;
  uint32_t flag;
  uint32_t pad;
};
constexpr uint32_t kWireChunk = 4096;  // elements of a segment per block
struct WireChunk {                     // one block's run of a segment
  uint64_t first;
  uint32_t seg;
  uint32_t pad;
};
// Layout of a commitment batch's device memory: [NTT-domain witnesses]
// [partials][descriptors][segments], each 256-byte aligned; descriptors sorted
// by witness length, longest first.
struct BatchLayout {
  std::vector<BatchDesc> desc;
  std::vector<BatchSeg> seg;
  size_t s_bytes = 0, p_bytes = 0, d_bytes = 0, g_bytes = 0;
  size_t total() const { return s_bytes + p_bytes + d_bytes + g_bytes; }
};
}  // namespace gh

struct gh_tuning {
  int mac_variant = 1;        // q32 MAC: 0 = dword loads, 1 = 16-byte loads
  int mac_waves = 4;          // waves per MAC block (4, 8, 16)
  int mac_blocks_per_cu = 2;  // MAC blocks (= partial sums) per compute unit
  int64_t quad_max = 1 << 17; // largest batch transformed four lanes per polynomial
  // q32 commit: 0 = transform and MAC kernels, 1 = one fused kernel whose
  // waves split a block's columns, 2 = one fused kernel whose waves split a
  // 16-row block's rows (rows % 16 == 0; four waves per block; else as 1)
  int fused = 2;
  int fused_depth = 2;        // fused kernel: key columns in flight per wave, x 8 / mac_waves
  int batch_chunk_cols = 256; // batched MAC: key columns per block (q32 keys)
  int batch_chunk_cols_gold = 256; // batched MAC: key columns per block (Goldilocks keys)
  int batch_group = 8;        // batched MAC: witnesses sharing each loaded key column (4, 8, 16)
  // device-memory gh_commit: 1 = hold the finalize back into the next
  // commitment's launch (one launch per commitment; results complete at the
  // next call on the context), 0 = finalize on the stream before returning
  int defer_finalize = 1;
  // held-back finalize, first stage (partial sums -> group sums), run in the
  // next launch: 0 = one task in each MAC block, n (1..8) = every task in
  // dedicated leading blocks of n tasks, -1 = such blocks, about num_cus / 2
  int stage1_tasks_per_block = -1;
  // Goldilocks gh_commit_batch and batch plans: 0 = transform, grouped MAC
  // and finalize kernels (faster), 1 = one launch (k_batch_fused_gold)
  int batch_fused = 0;
  // key columns per block of that launch (a multiple of 16)
  int batch_fused_cols = 256;
  // grouped MAC of gh_commit_batch and the server (groups of 4, rows % 16 ==
  // 0): 0 = the register-pipelined kernels, 2..4 = that many key columns
  // staged per wave in an LDS ring filled by global-to-LDS loads, each block
  // running through consecutive chunks of its group
  int batch_mac_ring = 4;
  // witnesses per group of the q32 ring MAC (4 or 8; Goldilocks rings use 4):
  // 8 halves the key reads (same-shape batches 94 k against 87 k commits/s)
  // but halves the blocks of small batches (inner commitments 53 against 41 us)
  int q32_ring_group = 4;
  // gh_negacyclic_mul: 0 = one pair per thread (the first operand's transform
  // parked in LDS), 1 = one pair per quad of lanes, -1 = the faster per field
  // (per quad, except per thread for q < 2^32 when quad_lds is 0)
  int negacyclic_quad = -1;
  // twisted transforms of gh_twisted_forward/inverse: 1 = four lanes per
  // polynomial at every size, 0 = forward four lanes up to quad_max and
  // inverse a polynomial per thread, -1 = 1 for Goldilocks and 0 for q < 2^32
  int xform_quad = 1;
  // quad transforms finish their 4-point DFT through LDS (1) or with lane
  // exchanges (0)
  int quad_lds = 1;
};
// most stage-1 tasks one dedicated block runs
constexpr uint32_t kStage1MaxTpb = 8;

// Tuned key columns per block of the batched MAC for keys over `ring`.
inline uint32_t tuned_batch_chunk(const gh_tuning& t, int ring_kind) {
  return (uint32_t)(ring_kind == GH_RING_Q32 ? t.batch_chunk_cols : t.batch_chunk_cols_gold);
}

// Finalizes of device-memory commitments (fused path) held back so that they
// run inside the next commitments' launches instead of as launches of their
// own: stage s1 is the last commitment, whose partial sums the next launch
// reduces to group sums (stage 1), and s2 the one before, whose group sums
// the next launch reduces and inverse-transforms into its t_out (stage 2).
// Sums of a stage lie at work + off, n polynomials per row, [row][n][slot].
// Every other call on the context issues both first (ctx_flush); a pipeline
// holds commitments of one ring. The workspace is carved into four areas of
// `cap` bytes: partial sums of alternate commitments in areas 0 and 1, their
// group sums in areas 2 and 3 (commitment c uses areas c % 2 and 2 + c % 2).
struct PendingStage {
  bool active = false;
  size_t off = 0;
  uint32_t n = 0, rows = 0;
  void* t_out = nullptr;
};
struct PendingFin {
  const gh_ring* ring = nullptr;
  PendingStage s1, s2;
  uint32_t flip = 0;  // area pair of the next commitment
  bool any() const { return s1.active || s2.active; }
};

struct gh_ctx {
  int device = 0;
  gh_tuning tune;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  // grow-only device workspace used by the calls of this context
  void* work = nullptr;
  size_t work_bytes = 0;
  // staging for host-memory calls
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // small fixed device buffer (flags)
  void* misc = nullptr;
  // this device's address of the no-allocation scratch array
  void* scratch = nullptr;
  // pinned host staging of batch descriptors: the upload of one batch reads it
  // asynchronously, so the next batch waits for `upload_done` before reuse
  gh::BatchLayout batch_layout;
  void* upload_host = nullptr;
  size_t upload_bytes = 0;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  // per-item completion counters of the one-launch Goldilocks batch; zero
  // between launches (the kernel resets what it counts)
  uint32_t* batch_done = nullptr;
  size_t batch_done_items = 0;
  // phase timing (gh_ctx_profile_*): event pairs recorded per commit
  bool prof_on = false;
  std::vector<hipEvent_t> prof_pool;
  size_t prof_used = 0;
  struct ProfRec { int phase; hipEvent_t a, b; };
  std::vector<ProfRec> prof_recs;
  PendingFin fin;
  std::mutex mu;
};

struct gh_ring {
  gh_ctx* ctx = nullptr;
  uint64_t q = 0;
  uint32_t d = 0;
  int kind = 0;
  uint32_t word_bytes = 0;
  uint32_t wire_bytes = 0;  // encoded element: ceil(bits(q - 1) / 8)
  uint64_t generator = 0, psi = 0, omega = 0, d_inv = 0;
  uint32_t qinv32 = 0;
  std::vector<uint64_t> host_tab;  // twiddle table image (layout in gh_field.h)
  void* dev_tab = nullptr;
};

struct gh_key {
  gh_ring* ring = nullptr;
  uint32_t rows = 0, msg_len = 0, open_len = 0;
  // A in NTT domain, layout [col][row][slot], col over msg then open columns
  void* dev_a = nullptr;
};

namespace gh {

uint64_t powmod(uint64_t a, uint64_t e, uint64_t q);
int plan_ring(uint64_t q, uint32_t d, gh_ring* ring);

// host inputs are checked on the host before upload: GH_OK or GH_E_NOT_CANONICAL
int check_canonical_host(const gh_ring* ring, const void* p, size_t words);
// device workspace of at least `bytes` bytes on the context (grow-only)
int ctx_workspace(gh_ctx* ctx, size_t bytes, void** out);
int ctx_staging(gh_ctx* ctx, size_t bytes, void** out);
// records an event on the context stream when profiling is on (else nullptr)
hipEvent_t prof_mark(gh_ctx* ctx);
void prof_span(gh_ctx* ctx, int phase, hipEvent_t a, hipEvent_t b);
// a start/stop event pair for a kernel launched with hipExtLaunchKernelGGL,
// registered under `phase` (both null when profiling is off)
void prof_pair(gh_ctx* ctx, int phase, hipEvent_t* start, hipEvent_t* stop);

// launchers (device pointers; enqueue on ctx->stream)
enum XformMode { kTwistedFwd = 0, kTwistedInv = 1, kCyclicFwd = 2, kCyclicInv = 3 };
int launch_xform(gh_ring* ring, int mode, const void* src, void* dst, size_t count);
int launch_negacyclic_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count);
int launch_pointwise_mul(gh_ring* ring, const void* a, const void* b, void* out, size_t count);
int launch_key_layout(gh_ring* ring, const void* src_rowmajor, uint32_t rows, uint32_t cols,
                      uint32_t col_offset, uint32_t total_cols, void* dst, int domain);
// defer: a commitment on the fused path leaves its finalize pending (see
// PendingFin); other paths flush pending ones and finalize in place
int launch_commit(gh_key* key, const void* s_dev, uint32_t cols, int domain, void* t_dev,
                  bool defer = false);
// issues a pending finalize on the context stream (context lock held)
int ctx_flush(gh_ctx* ctx);
// Bytes of the device array behind gh_commit_noalloc (partial sums of one
// commitment: blocks x rows x 64 words), and its address on the current device.
constexpr size_t kScratchBytes = (size_t)4 << 20;
int scratch_address(void** out);
// gh_commit_noalloc: fused path only (rows % 4 == 0, coefficient-domain
// witness), partial sums in `scratch`; GH_E_INVALID_ARGUMENT beyond its size
int launch_commit_noalloc(gh_key* key, const void* s_dev, uint32_t cols, void* t_dev,
                          void* scratch);
// items hold device pointers; every len in [1, key columns]
int launch_commit_batch(gh_key* key, const gh_commit_item* items, size_t count);
// one-launch Goldilocks batch over sorted descriptors (ranges of range_cols
// key columns); done: count zeroed counters, zero again when it completes
int enqueue_batch_fused_gold(gh_key* key, hipStream_t stream, const BatchDesc* desc,
                             uint32_t count, uint32_t maxlen, uint32_t range_cols,
                             uint32_t* done);

// Batch building blocks on an explicit stream. Descriptor arrays are device
// copies; MAC descriptors are sorted by len, longest first, and the host
// passes the longest len (the grid's chunk count follows from it).
int enqueue_batch_xform(gh_ring* ring, hipStream_t stream, const BatchDesc* desc,
                        const BatchSeg* seg, uint32_t nseg);
int enqueue_batch_mac(gh_key* key, hipStream_t stream, const BatchDesc* desc, uint32_t count,
                      uint32_t maxlen, uint32_t chunk_cols, int group, int variant);
int enqueue_batch_finalize(gh_ring* ring, hipStream_t stream, const BatchDesc* desc,
                           uint32_t count, uint32_t rows, uint32_t chunk_cols);
// column chunking of one Gram matrix (partial sums: nchunks x r(r+1)/2 polys)
void gram_chunking(int num_cus, uint32_t r, uint32_t n, uint32_t* chunk, uint32_t* nchunks);
uint32_t gram_tiles(uint32_t r);
int enqueue_gram_batch(gh_ring* ring, hipStream_t stream, const GramDesc* desc,
                       const GramWork* work, uint32_t nwork, const GramFin* fin, uint32_t nfin);
int launch_check_canonical(gh_ring* ring, const void* p, size_t words, int* bad_flag_dev);
int launch_gram(gh_ring* ring, const void* s_dev, uint32_t r, uint32_t n, int domain, void* g_dev);
int launch_decompose(gh_ring* ring, const void* in, size_t polys, uint64_t base, uint32_t limbs,
                     void* out);
int launch_garbage_h(gh_ring* ring, const void* phi, const void* s, uint32_t r, uint32_t n,
                     int domain, void* h_dev);
// s_scratch: r x n polys of device memory for the NTT-domain witness (domain COEFF)
int launch_amortize(gh_ring* ring, const void* c, const void* s, uint32_t r, uint32_t n,
                    int domain, void* z_dev, void* s_scratch);
// gh_measure_issue's kernels: kind 0..3, kIssueIters iterations per wave
int launch_issue_probe(gh_ctx* ctx, int kind, uint32_t blocks, int* out, hipEvent_t e0,
                       hipEvent_t e1);
// gh_measure_valu's kernels: kind 0..9, kValuIters iterations of eight chains
constexpr int kValuKinds = 10;
int launch_valu_probe(gh_ctx* ctx, int kind, uint32_t blocks, uint32_t* out, hipEvent_t e0,
                      hipEvent_t e1);
int launch_stream_probe(gh_ctx* ctx, const void* src, void* dst, size_t bytes, int copy,
                        hipEvent_t e0, hipEvent_t e1);
// Ring slice hooks (gh_slice.hip) on `len` device words: dst op= src, the
// scalar a canonical residue (kSliceScale: dst *= scalar, src unused;
// kSliceAxpy: dst += scalar * src)
enum SliceOp { kSliceAdd = 0, kSliceSub = 1, kSliceScale = 2, kSliceAxpy = 3 };
int launch_slice(gh_ring* ring, int op, void* dst, const void* src, uint64_t scalar, size_t len);
// out[0] = sum a[i] * b[i] mod q; partial: dot_partial_words(len) device words
size_t dot_partial_words(gh_ring* ring, size_t len);
int launch_dot(gh_ring* ring, const void* a, const void* b, size_t len, void* partial, void* out);
// canonical encodings (gh_wire.hip); chunk lists follow the segments
void wire_chunks(const std::vector<WireSeg>& segs, std::vector<WireChunk>* out);
int enqueue_wire_unpack(gh_ring* ring, hipStream_t stream, const WireSeg* segs,
                        const WireChunk* chunks, uint32_t nchunks, uint32_t* flags);
int enqueue_wire_pack(gh_ring* ring, hipStream_t stream, const WireSeg* segs,
                      const WireChunk* chunks, uint32_t nchunks);
// dst (row-major rows x cols polys) = key columns [col_offset, col_offset + cols)
int launch_key_rowmajor(gh_ring* ring, const void* key_a, uint32_t rows, uint32_t cols,
                        uint32_t col_offset, void* dst);
// JL rows for aggregation (gh_jl.hip); d = 64
int launch_jl_flat(gh_ring* ring, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                   uint64_t cols, int8_t* flat);
int launch_jl_conj_seed(gh_ring* ring, const uint8_t seed[32], uint32_t rows, uint32_t num_parts,
                        uint64_t n, void* out);
int launch_jl_conj_flat(gh_ring* ring, const int8_t* flat, uint32_t rows, uint32_t num_parts,
                        uint64_t n, void* out);
// Launch shape of an int8 matrix-core JL kernel over `total` witness
// polynomials: polynomials per chunk launch (chunks of them), blocks of the
// first (largest) chunk, polynomials per wave (projection) or per block
// (evaluations), and the int32 bound: each sum adds at most `terms` products
// of magnitude <= 128, bound = terms * 128 <= 2^31 - 1.
struct JlShape {
  uint64_t chunk_polys, per_unit, blocks, chunks, terms, bound;
};
void jl_proj_shape(uint64_t total, int num_cus, JlShape* s);
void jl_evals_shape(uint64_t total, int num_cus, JlShape* s);
// the int64 p_exact is exact below this many polynomials:
// |p| <= total * 64 * (q - 1) / 2 < 2^26 * 2^6 * 2^31 = 2^63
constexpr uint64_t kJlExactMaxPolys = 1ull << 26;
// witness polynomials the JL entry points accept: 65535 parts x 2^20
constexpr uint64_t kJlMaxPolys = 1ull << 36;
// p = Pi * s on the matrix cores; rows in [253, 256] (one keystream block per
// column); p_exact (q32 only, nullable): the exact centred sum
int launch_jl_proj(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
                   uint32_t num_parts, uint64_t n, void* p_modq, int64_t* p_exact);
// rows = 256; parts: num_parts x n polys, coefficient domain; out: 256 polys
int launch_jl_evals(gh_ring* ring, const uint8_t seed[32], const void* parts, uint32_t num_parts,
                    uint64_t n, void* out);
// phi[g] += sum_m scale_m * sigma_{-1}(pi_g^(m)) over the 256 JL rows (device pointers)
int launch_jl_aggregate(gh_ring* ring, const uint8_t seed[32], const void* scale,
                        uint32_t num_parts, uint64_t n, void* phi);
int launch_jl(gh_ring* ring, const uint8_t seed[32], uint32_t rows, const void* parts,
              uint32_t num_parts, uint64_t cols, void* p_modq, int64_t* p_exact, int8_t* flat);

}  // namespace gh

#define GH_HIP(call)                                  \
  do {                                                \
    hipError_t _e = (call);                           \
    if (_e != hipSuccess) {                           \
      return _e == hipErrorOutOfMemory ? GH_E_OOM : GH_E_DEVICE; \
    }                                                 \
  } while (0)
