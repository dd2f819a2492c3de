// Context layout shared by the runtime and the fold-in server.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ezlda.h"
#include "ezlda_device.h"

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
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int n_xcd = 1;      // XCDs of the device (round-robin workgroup dispatch)
    int doc_tiles = 0;  // doc tiles used for head-word items (0 = none)
};

namespace ezc {
void set_error(const std::string &m);
// Rebuilds ctx's counts from its current z and 1/(n_k + V beta), synchronously.
int refresh_counts(ezlda_ctx *c);
}  // namespace ezc
