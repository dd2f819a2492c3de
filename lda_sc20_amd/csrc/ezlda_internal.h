// Context layout shared by the runtime and the fold-in server.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ezlda.h"
#include "ezlda_device.h"

// Out-of-core execution (num_chunks > 1): the documents, in the layout's doc
// order, are cut into chunks of about equal token counts.  Each chunk keeps
// its token state and word-major layout in one pinned host image; a sweep
// passes every chunk through one of two device image buffers.
struct EzChunk {
    uint32_t q0 = 0, D = 0, N = 0;  // first doc (doc order), docs, tokens
    uint32_t slot0 = 0;             // first doc slot of the corpus layout
    uint32_t n_items = 0, n_wt_segs = 0, n_queues = 1;
    int32_t D_long = 0;             // docs of more than ez::kDtShort tokens (the chunk's first)
    uint64_t n_runs = 0, dt_entries = 0;
    // image: z | z_doc | doc_info (copied back after each sweep), then
    // soff, slot_wrd, slot_tok, items, runs, WT segments
    size_t off_doc_info = 0, back_bytes = 0, off_soff = 0, off_slot_wrd = 0, off_slot_tok = 0, off_items = 0,
           off_runs = 0, off_wt_segs = 0, bytes = 0;
    uint8_t *img = nullptr;  // pinned
};

struct ezlda_ctx {
    ezlda_args a{};
    ez::DevState s{};
    hipStream_t st = nullptr;
    std::string err;
    bool loaded = false;
    bool counts_fresh = false;
    uint32_t sweep = 0;
    int64_t dev_bytes = 0;
    int64_t dt_cap = 0, tail_cap = 0, dt_cap_canon = 0;
    std::vector<void *> allocs;         // corpus-level device arrays
    std::vector<void *> layout_allocs;  // arrays the tuning arguments decide
    int64_t layout_bytes = 0;
    // host mirrors for result mapping
    std::vector<uint32_t> doc_tok;   // doc-major position -> token id
    std::vector<int32_t> rank_of;    // input word -> rank
    std::vector<uint32_t> word_off;  // V+1 by rank
    std::vector<uint32_t> doc_off;   // D+1
    std::vector<int32_t> dord;       // doc order q -> input doc (length desc, id asc)
    std::vector<uint32_t> soff;      // D+1: slots of doc q
    std::vector<uint32_t> slot_tok;  // N: slot -> token id
    std::vector<uint32_t> row_off;   // D: DT row offset of doc q (the runs carry it)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int n_xcd = 1;      // XCDs of the device (round-robin workgroup dispatch)
    int doc_tiles = 0;  // doc tiles used for head-word items (0 = none)
    // out-of-core execution
    std::vector<EzChunk> chunks;       // empty: the corpus is resident
    std::vector<uint32_t> chunk_slot0; // first slot of each chunk (lookup)
    std::vector<uint32_t> tok_slot_g;  // token id -> corpus slot
    uint8_t *cimg[2] = {nullptr, nullptr};  // device image buffers
    uint32_t *cdt[2] = {nullptr, nullptr};  // DT rows of the chunk in flight
    uint32_t *caux[2] = {nullptr, nullptr};
    uint32_t *cctr[2] = {nullptr, nullptr};  // per-buffer sweep counter and item queues
    double *cll[2] = {nullptr, nullptr};     // per-item log-likelihood partials
    double *cllo[2] = {nullptr, nullptr};
    int32_t *wt_next = nullptr, *nk_next = nullptr;  // counts of the sweep's new topics
    hipStream_t st2 = nullptr;  // second chunk stream
    // the device's static log-likelihood array (ezlda_loglik_noalloc), looked
    // up once per device when the first context on it is created
    double *ll_scratch = nullptr;
    std::string scratch_err;
};

namespace ezc {
void set_error(const std::string &m);
// Inside a catch block: the status for the exception in flight (std::bad_alloc
// -> EZLDA_ENOMEM, anything else -> EZLDA_EINVAL), with its message recorded
// for ezlda_last_error(c) (c may be null) or ezlda_server_last_error(s).  No
// exception leaves an extern "C" entry point.
int exception_status(ezlda_ctx *c);
int server_exception_status(ezlda_server *s);
// Rebuilds ctx's counts from its current z and 1/(n_k + V beta), synchronously.
int refresh_counts(ezlda_ctx *c);
// The chunked folder at root (ezlda_corpus.cpp): per doc-major token,
// chunk << 32 | its TL.txt position, and its word (rank); tokens per chunk.
bool chunked_layout(const std::string &root, std::vector<uint64_t> &pos, std::vector<int32_t> &words,
                    std::vector<int64_t> &tll);
// Rewrites the folder's TL.txt files (and a current binary sidecar) with the
// doc-major topics z (0-based, < K).
bool rewrite_tl(const std::string &root, const std::vector<uint64_t> &pos, const std::vector<int64_t> &tll,
                const uint16_t *z, int K);
}  // namespace ezc
