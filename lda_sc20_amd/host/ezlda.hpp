// C++ host surface over the C ABI (include/ezlda.h): RAII handles that
// throw ezlda::Error (the library's status and message) where the reference's
// void launchers leave errors unchecked.  Header-only; link with -lezlda.
// The `ezlda` command (ezlda_cli.cpp) is written against it.
//
// Launcher mapping (reference -> this build):
//   UpdateWTDenKernel, UpdateWTKernel, WTAdditionKernel (WTDenUpdateKernel.cuh:25,
//   WTUpdateKernel.cuh:23, WTAddKernel.cuh:24) -> Context::build_wt
//   UpdateWTDenRowSumKernel (WTDenUpdateKernel.cuh:26) -> Context::wt_rowsum
//   MaxTopicKernel (SamplingKernel.cuh:12)   -> Context::topk
//   UpdateDTKernel (DTUpdateKernel.cuh:13)   -> Context::build_dt
//   UpdateProbKernelD (SamplingKernel.cuh:15) -> Context::shortcut
//   SampleKernelD (SamplingKernel.cuh:10)    -> Context::sample_head
//   SampleKernel (SamplingKernel.cuh:11)     -> Context::sample_tail
//   the head commit (mainKernel.cu:355-356)  -> Context::commit_head
//   (the last four fused)                    -> Context::sample
//   PerplexityKernel (SamplingKernel.cuh:13) -> Context::loglik
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ezlda.h"

namespace ezlda {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const ezlda_ctx *ctx) {
    if (rc != EZLDA_OK) throw Error(rc, ezlda_last_error(ctx));
}

class Corpus {
public:
    explicit Corpus(ezlda_corpus *h) : h_(h) {
        if (!h_) throw Error(EZLDA_EIO, ezlda_last_error(nullptr));
    }
    Corpus(const Corpus &) = delete;
    Corpus &operator=(const Corpus &) = delete;
    Corpus(Corpus &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    ~Corpus() { ezlda_corpus_free(h_); }

    static Corpus generate(const ezlda_gen_params &p) { return Corpus(ezlda_gen_corpus(&p)); }
    static Corpus read_docword(const std::string &path) { return Corpus(ezlda_read_docword(path.c_str())); }
    static Corpus read_chunked(const std::string &prefix, int chunks = 0) {
        return Corpus(ezlda_read_chunked(prefix.c_str(), chunks));
    }
    void write_chunked(const std::string &prefix, int K, int chunks, uint64_t seed,
                       const uint16_t *z0 = nullptr) const {
        check(ezlda_write_chunked(prefix.c_str(), h_, z0, K, chunks, seed), nullptr);
    }

    int64_t tokens() const { return ezlda_corpus_num_tokens(h_); }
    int32_t docs() const { return ezlda_corpus_num_docs(h_); }
    int32_t words() const { return ezlda_corpus_num_words(h_); }
    const int64_t *doc_off() const { return ezlda_corpus_doc_off(h_); }
    const int32_t *word_ids() const { return ezlda_corpus_words(h_); }
    const uint16_t *topics() const { return ezlda_corpus_topics(h_); }

private:
    ezlda_corpus *h_;
};

inline ezlda_args default_args() {
    ezlda_args a;
    ezlda_default_args(&a);
    return a;
}

class Context {
public:
    explicit Context(const ezlda_args &a) : h_(ezlda_create(&a)) {
        if (!h_) throw Error(EZLDA_EHIP, ezlda_last_error(nullptr));
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    ~Context() { ezlda_destroy(h_); }

    void load(const Corpus &c) {
        check(ezlda_load_docs(h_, c.words(), c.docs(), c.doc_off(), c.word_ids(), c.topics()), h_);
    }
    void load_chunked(const std::string &prefix, int chunks = 0) {
        check(ezlda_load_chunked(h_, prefix.c_str(), chunks), h_);
    }
    ezlda_info info() const {
        ezlda_info i;
        check(ezlda_get_info(h_, &i), h_);
        return i;
    }
    void build_wt() { check(ezlda_build_wt(h_), h_); }
    void topk() { check(ezlda_topk(h_), h_); }
    void build_dt() { check(ezlda_build_dt(h_), h_); }
    void sample() { check(ezlda_sample(h_), h_); }
    void shortcut() { check(ezlda_shortcut(h_), h_); }
    void sample_head() { check(ezlda_sample_head(h_), h_); }
    void sample_tail() { check(ezlda_sample_tail(h_), h_); }
    void commit_head() { check(ezlda_commit_head(h_), h_); }
    void wt_rowsum() { check(ezlda_wt_rowsum(h_), h_); }
    void sweep(int n = 1) { check(ezlda_sweep(h_, n), h_); }
    void sweep_graph(int n = 1) { check(ezlda_sweep_graph(h_, n), h_); }
    double loglik() {
        double ll = 0.0;
        check(ezlda_loglik(h_, &ll), h_);
        return ll;
    }
    // Wall ms of n back-to-back sweeps; stage_ms gets EZLDA_NUM_STAGES values.
    double time_sweeps(int n, bool graph, double *stage_ms = nullptr) {
        double ms = 0.0, st[EZLDA_NUM_STAGES];
        check(ezlda_time_sweeps(h_, n, graph ? 1 : 0, &ms, stage_ms ? stage_ms : st), h_);
        return ms;
    }
    std::vector<uint16_t> topics() {
        std::vector<uint16_t> z((size_t)info().num_tokens);
        check(ezlda_get_topics(h_, z.data()), h_);
        return z;
    }
    // Layout arguments only (block_dim, grid_dim, token_segment, dense_words,
    // dense_threshold, doc_tiles, run_cap); topics do not depend on them.
    void set_layout(const ezlda_args &a) { check(ezlda_set_layout(h_, &a), h_); }
    std::vector<ezlda_tune_trial> autotune(int sweeps_per_trial, ezlda_args *chosen, double *ms_per_sweep) {
        std::vector<ezlda_tune_trial> t(64);
        int32_t n = 0;
        check(ezlda_autotune(h_, sweeps_per_trial, chosen, ms_per_sweep, t.data(), (int32_t)t.size(), &n), h_);
        t.resize(std::min<size_t>((size_t)n, t.size()));
        return t;
    }
    void export_model(const std::string &dir) { check(ezlda_export_model(h_, dir.c_str()), h_); }
    void save_checkpoint(const std::string &p) { check(ezlda_save_checkpoint(h_, p.c_str()), h_); }
    void load_checkpoint(const std::string &p) { check(ezlda_load_checkpoint(h_, p.c_str()), h_); }
    // The chunked folder's TL.txt rewritten with the current topics, plus
    // (seed, sweep); load_tl resumes from it.
    void save_tl(const std::string &prefix) { check(ezlda_save_tl(h_, prefix.c_str()), h_); }
    void load_tl(const std::string &prefix) { check(ezlda_load_tl(h_, prefix.c_str()), h_); }
    ezlda_ctx *handle() { return h_; }

private:
    ezlda_ctx *h_;
};

// Fold-in server over a trained context (the model is frozen at creation).
class Server {
public:
    Server(Context &model, const ezlda_server_args &a) : h_(ezlda_server_create(model.handle(), &a)) {
        if (!h_) throw Error(EZLDA_EHIP, ezlda_last_error(nullptr));
    }
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;
    ~Server() { ezlda_server_destroy(h_); }

    void submit(uint32_t rid, const std::vector<int32_t> &words) {
        chk(ezlda_server_submit(h_, rid, words.data(), (int32_t)words.size()));
    }
    // Documents d = 0..ids.size()-1: words [doc_off[d], doc_off[d+1]); all or none.
    void submit_many(const std::vector<uint32_t> &ids, const std::vector<int64_t> &doc_off,
                     const std::vector<int32_t> &words) {
        if (doc_off.size() != ids.size() + 1) throw Error(EZLDA_EINVAL, "doc_off needs ids.size() + 1 entries");
        chk(ezlda_server_submit_many(h_, (int32_t)ids.size(), ids.data(), doc_off.data(), words.data()));
    }
    // Topics and topic counts of a finished request; latency in ms.
    double wait(uint32_t rid, std::vector<uint16_t> *z, std::vector<int32_t> *counts, int32_t timeout_ms = -1) {
        double lat = 0.0;
        chk(ezlda_server_wait(h_, rid, z ? z->data() : nullptr, counts ? counts->data() : nullptr, &lat, timeout_ms));
        return lat;
    }
    void drain() { chk(ezlda_server_drain(h_)); }
    ezlda_server_stats stats() {
        ezlda_server_stats st;
        chk(ezlda_server_get_stats(h_, &st));
        return st;
    }
    void reset_stats() { chk(ezlda_server_reset_stats(h_)); }
    void plain(uint32_t rid, const std::vector<int32_t> &words, std::vector<uint16_t> *z,
               std::vector<int32_t> *counts) {
        chk(ezlda_foldin_plain(h_, rid, words.data(), (int32_t)words.size(), z ? z->data() : nullptr,
                               counts ? counts->data() : nullptr));
    }

private:
    void chk(int rc) {
        if (rc != EZLDA_OK) throw Error(rc, ezlda_server_last_error(h_));
    }
    ezlda_server *h_;
};


}  // namespace ezlda
