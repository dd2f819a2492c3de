// Launchers of the gfx950 sweep kernels (ezlda_kernels.hip).  Plain device
// pointers and sizes; every launcher enqueues on the given stream and returns
// the hipError_t of its launch.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace ez {

// Sweep counter and persistent-sampler queue heads: each head on its own
// 128-byte line, so the workgroups claiming items from different queues do
// not serialize on one line.
static constexpr int kQueueStride = 32;
static constexpr int kCounterWords = kQueueStride * 9;
// Documents of at most this many tokens can have their DT rows built by
// sorting their topics (large K); longer ones use a K-sized histogram.
static constexpr uint32_t kDtShort = 512;

// Device image of one corpus + model.  Token ids are word-major positions;
// N < 2^32.  Sizes: V words, VD head words, D docs, K topics.
struct DevState {
    int32_t K = 0, V = 0, VD = 0, D = 0;
    uint32_t N = 0;
    float alpha = 0.f, beta = 0.f, vbeta = 0.f;
    uint64_t seed = 0;
    // tokens, word-major (token id = word rank, doc order, position)
    uint32_t *word_off = nullptr;   // V+1
    uint16_t *z = nullptr;          // N
    // tokens by doc slot: doc q (docs by descending length) owns slots
    // [soff[q], soff[q+1]), its tokens grouped by word rank, then position
    uint32_t *soff = nullptr;       // D+1
    uint32_t *slot_wrd = nullptr;   // N: word rank
    uint32_t *slot_tok = nullptr;   // N: token id of the slot
    uint16_t *z_doc = nullptr;      // N: z by slot, kept by the sampler for head-word slots
    uint32_t *aux_doc = nullptr;    // N: n_dk1 of the slot's word | its doc's row length (nnz) << 16
    // doc-topic rows
    uint2 *doc_info = nullptr;      // D, by doc q: x = DT row offset, y = nnz | len << 16
    int32_t D_long = 0;             // docs of more than kDtShort tokens (the first D_long in doc order)
    uint32_t *dt = nullptr;         // rows: u32 (topic << 2) | count << 16, or (compact) u16 topic | piece << 10
    bool compact = false;           // K <= 8192: 16-bit entries topic | piece << topic_bits
    uint32_t topic_bits = 16;       // max(10, ceil(log2 K)); pieces of <= 2^(16 - topic_bits) - 1
    // word-topic: dense head rows + sorted sparse tail rows
    int32_t *wt_head = nullptr;     // VD*K
    uint32_t *tail_off = nullptr;   // V-VD+1 (capacity min(n_w,K))
    uint32_t *wt_tail = nullptr;    // topic | count << 16
    uint32_t *tail_nnz = nullptr;   // V-VD
    int32_t *nk = nullptr;          // K
    float *inv = nullptr;           // K: 1/(n_k + V beta)
    ushort4 *wtop = nullptr;        // V: k1, k2, k3 (0xFFFF = none)
    uint32_t *zero_order = nullptr; // K: topics by beta * inv, largest first (tail words' zero counts)
    // work items
    uint4 *wt_segs = nullptr;       // head-word histogram segments (w, beg, end)
    int32_t n_wt_segs = 0;
    uint4 *smp_segs = nullptr;      // sampler items (w, first run, end run, draw-key offset)
    int32_t n_smp_segs = 0;
    uint4 *runs = nullptr;          // (first token, first slot, doc's DT row offset, tokens | doc length << 16):
                                    // a word's tokens in one doc
    int64_t n_runs = 0;
    int32_t n_queues = 1;           // persistent sampler: item queues (<= 8)
    uint32_t *counters = nullptr;   // kCounterWords: [0] sweep, queue q's head at [kQueueStride * (q + 1)]
    unsigned long long *stats = nullptr;  // [0] full draws, [1] DT entries scanned
    double *ll_part = nullptr;      // n_smp_segs
    double *ll_out = nullptr;       // 1
    int32_t block_dim = 128;
    int32_t grid_dim = 0;           // 0 = one workgroup per sampler item
    int32_t n_cu = 256;             // compute units of the device
    int32_t lds_cu = 160 * 1024;    // LDS bytes per compute unit
};

hipError_t launch_zero_counts(const DevState &s, hipStream_t st);
hipError_t launch_build_wt(const DevState &s, hipStream_t st);
// Adds the counts of s.z (n_k, dense head rows, tail rows) without zeroing.
hipError_t launch_accumulate_wt(const DevState &s, hipStream_t st);
hipError_t launch_topk(const DevState &s, hipStream_t st);
// DT rows, doc_info and the per-slot records of the sampler (n_dk1 | nnz << 16
// in s.aux_doc); with n_dk2 set, (n_dk1 | n_dk2 << 16) instead (the reference's
// per-token pair, which the sampler does not need).
hipError_t launch_build_dt(const DevState &s, hipStream_t st, bool n_dk2 = false);
hipError_t launch_z_to_doc(const DevState &s, hipStream_t st);
// Sampler stages (mode bits): the items of head words, of tail words, and
// the max-topic shortcut alone (tokens it accepts get their top topic; the
// others keep theirs for a later full draw).  A token's draw depends only on
// the counts of the sweep, so the stages give the fused sampler's topics in
// any order.
static constexpr uint32_t kSampleHead = 1u, kSampleTail = 2u, kShortcutOnly = 4u;
hipError_t launch_sample(const DevState &s, bool shortcut, hipStream_t st,
                         uint32_t mode = kSampleHead | kSampleTail);
// Dense head rows of the current z (zeroed first) and n_k alone.
hipError_t launch_head_rows(const DevState &s, hipStream_t st);
hipError_t launch_topic_hist(const DevState &s, hipStream_t st);
hipError_t launch_advance_sweep(const DevState &s, hipStream_t st);
hipError_t launch_loglik(const DevState &s, hipStream_t st);
// Log-likelihood partials the allocation-free entry point can hold (one per
// sampler item); its static device array holds these plus the result.
static constexpr size_t kLLScratchItems = (size_t)1 << 20;
// Device address of that array on the current device (hipGetSymbolAddress).
hipError_t scratch_address(double **p);
hipError_t launch_set_u32(uint32_t *p, uint32_t v, hipStream_t st);

hipError_t launch_selftest_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, int n, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
// n_reads scattered 64-byte block reads (four 16-byte loads per lane), each
// a distinct block of the largest power-of-two block count within bytes of
// src (n_reads may not exceed it); sink (>= 64 words) is written only on an
// impossible value.
hipError_t launch_gather64(const void *src, size_t bytes, uint64_t n_reads, uint32_t *sink, hipStream_t st);

// Fold-in against a frozen model: per-word phi rows, Q running sums inside
// 64-topic slabs (qs, V x K) and over slab totals (qacc, V x ceil(K/64)).
hipError_t launch_foldin_tables(const DevState &s, float *phi, float *qs, float *qacc, hipStream_t st);
// Buffers of the word-major fold-in step.
struct FoldinBatch {
    uint32_t *rows = nullptr;       // row pool, parallel to the token pool: topic | count << 16
    uint32_t *row_nnz = nullptr;    // per slot
    uint32_t *wcount = nullptr;     // V: tokens per word in the step
    uint32_t *wcursor = nullptr;    // V: bucket cursors
    uint32_t *bsum = nullptr;       // 3 x ceil(V / 1024): per-block big-word tokens, big items, other tokens
    uint32_t *tokw = nullptr;       // resident tokens: slot << 16 | position, by word (big words first)
    uint4 *items = nullptr;         // foldin_max_items(V, tokens) big items: (word, first, end, 0)
    uint32_t *ctr = nullptr;        // big items, big-word tokens, other tokens, 0
    uint16_t *zout = nullptr;       // finishing documents' topics, packed
    int grid = 0;                   // persistent sampling workgroups
};
size_t foldin_lds_bytes(int K);
uint32_t foldin_max_items(int V, uint32_t max_tokens);
// Copies admitted documents (source offset, pool offset, tokens, 0) from an
// upload buffer into the word pool.
hipError_t launch_foldin_admit(const uint4 *adm, int nadm, const uint32_t *src, uint32_t *wpool, hipStream_t st);
// One fold-in sweep of every slot (request id, pool offset (a multiple of 8),
// tokens, sweep); fin[j] = (output index or ~0, topic offset) for the
// documents that finish with this sweep; ntok = their tokens.  nslots <=
// 65535 documents of <= 65535 tokens.
// The plain path: every sweep of one document (n tokens, word ranks) in one
// launch of one workgroup; z receives its n topics.
hipError_t launch_foldin_doc(const uint32_t *words, uint32_t n, uint32_t rid, int sweeps, int K, float alpha,
                             const float *phi, const float *qs, const float *qacc, uint64_t seed, uint16_t *z,
                             hipStream_t st);
hipError_t launch_foldin_step(const FoldinBatch &fb, const uint4 *slots, const uint2 *fin, int nslots, uint32_t ntok,
                              int V, int K, float alpha, const float *phi, const float *qs, const float *qacc,
                              const uint32_t *wpool, uint16_t *zpool, uint64_t seed, hipStream_t st);

// Bytes of dynamic LDS the sampler needs for K topics and block size bd.
size_t sample_lds_bytes(int K, int bd);
// Static LDS of the sampler kernels: must be 0 (phi is read at LDS address
// 0); (size_t)-1 if it cannot be queried.
size_t sampler_static_lds();
// Largest LDS any launch of the library asks for at this K and block size.
size_t max_lds_bytes(int K, int bd);

}  // namespace ez
