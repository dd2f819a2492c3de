// ezlda -- command-line driver over the C ABI.
//
//   ezlda train --prefix DIR [--iters 100] [--K 1024] ...
//   ezlda gen --out DIR --V 10000 --D 10000 --K-true 64 [--K 64] ...
//   ezlda preprocess --docword FILE --out DIR [--K 1024] [--chunks 8]
//   ezlda bench [--V ...] [--sweeps 20]
//
// `train` is the run-time form of EZLDA's main() (EZLDA/mainKernel.cu:30-386),
// whose dataset path, chunk count and sweep count are compile-time constants
// there (mainKernel.cu:50-56, Argument.cuh:3-9): it loads the chunked folder
// that preprocess writes, runs the sweeps, appends one per-sweep time in ms per
// line to DIR/SamplingDRecord.txt (mainKernel.cu:59,376) and prints the
// perplexity.  `preprocess` replaces preprocess/streamTestBalance.cpp, whose
// input and output paths are hard-coded (:90,:176,:334).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "ezlda.hpp"

namespace {

struct Opts {
    std::map<std::string, std::string> kv;
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    std::string str(const std::string &k, const std::string &d = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    double num(const std::string &k, double d) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atof(it->second.c_str());
    }
    long long integer(const std::string &k, long long d) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::atoll(it->second.c_str());
    }
};

const char *kUsage =
    "usage: ezlda <train|foldin|gen|preprocess|bench> [options]\n"
    "  train      --prefix DIR [--chunks N] [--iters N] [--K K] [--alpha A] [--beta B]\n"
    "             [--block-dim N] [--grid-dim N] [--streams N] [--token-segment N] [--seed S]\n"
    "             [--dense-threshold N] [--dense-words N] [--doc-tiles N] [--run-cap N] [--plain]\n"
    "             [--topic-bits 16] [--graph] [--autotune] [--ll-every N] [--checkpoint FILE]\n"
    "             [--resume FILE] [--checkpoint-tl] [--resume-tl] [--record FILE] [--export DIR]\n"
    "             (--checkpoint-tl rewrites DIR's chunk TL.txt files with the final topics plus\n"
    "             DIR/ezlda_state.txt; --resume-tl continues from them)\n"
    "  foldin     train options, plus --requests DOCWORD [--foldin-sweeps 20] [--slots N]\n"
    "             [--budget-mb N] [--out FILE]: folds the documents of a docword file into the\n"
    "             trained model and writes each one's topic counts\n"
    "  gen        --out DIR --V N --D N --K-true N [--K K] [--chunks N] [--len poisson|lognormal|pareto]\n"
    "             [--len-mean X] [--len-sigma X] [--pareto-shape X] [--pareto-xm X] [--len-min N]\n"
    "             [--len-max N] [--zipf S] [--phi-conc C] [--theta-conc C] [--seed S]\n"
    "  preprocess --docword FILE --out DIR [--K K] [--chunks N] [--seed S]\n"
    "  bench      synthetic in-memory run: gen options plus [--K K] [--sweeps N] [--warmup N]\n";

Opts parse(int argc, char **argv, int first) {
    Opts o;
    for (int i = first; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) throw ezlda::Error(EZLDA_EINVAL, "unexpected argument " + a);
        a = a.substr(2);
        const bool flag = (i + 1 >= argc) || std::strncmp(argv[i + 1], "--", 2) == 0;
        o.kv[a] = flag ? "1" : argv[++i];
    }
    return o;
}

ezlda_args args_from(const Opts &o) {
    ezlda_args a = ezlda::default_args();
    a.K = (int32_t)o.integer("K", a.K);
    a.alpha = (float)o.num("alpha", 50.0 / a.K);  // Argument.cuh:9
    a.beta = (float)o.num("beta", a.beta);
    a.block_dim = (int32_t)o.integer("block-dim", a.block_dim);
    a.grid_dim = (int32_t)o.integer("grid-dim", a.grid_dim);
    a.num_chunks = (int32_t)o.integer("chunks", a.num_chunks);
    a.num_streams = (int32_t)o.integer("streams", a.num_streams);
    a.token_segment = (int32_t)o.integer("token-segment", a.token_segment);
    a.seed = (uint64_t)o.integer("seed", (long long)a.seed);
    a.dense_threshold = (int32_t)o.integer("dense-threshold", a.dense_threshold);
    a.dense_words = (int32_t)o.integer("dense-words", a.dense_words);
    a.doc_tiles = (int32_t)o.integer("doc-tiles", a.doc_tiles);
    a.run_cap = (int32_t)o.integer("run-cap", a.run_cap);
    a.shortcut = o.has("plain") ? 0 : 1;
    a.topic_bits = (int32_t)o.integer("topic-bits", a.topic_bits);  // checked by ezlda_create
    return a;
}

ezlda_gen_params gen_from(const Opts &o) {
    ezlda_gen_params p;
    std::memset(&p, 0, sizeof p);
    p.V = (int32_t)o.integer("V", 10000);
    p.D = (int32_t)o.integer("D", 10000);
    p.K_true = (int32_t)o.integer("K-true", 64);
    const std::string len = o.str("len", "poisson");
    p.len_dist = len == "poisson" ? 0 : len == "lognormal" ? 1 : len == "pareto" ? 2 : -1;
    if (p.len_dist < 0) throw ezlda::Error(EZLDA_EINVAL, "--len must be poisson, lognormal or pareto");
    p.len_mean = o.num("len-mean", 100.0);
    p.len_sigma = o.num("len-sigma", 0.8);
    p.pareto_shape = o.num("pareto-shape", 1.2);
    p.pareto_xm = o.num("pareto-xm", 1.0);
    p.len_min = (int32_t)o.integer("len-min", 1);
    p.len_max = (int32_t)o.integer("len-max", 65535);
    p.zipf_s = o.num("zipf", 0.0);
    p.phi_conc = o.num("phi-conc", 0.01);
    p.theta_conc = o.num("theta-conc", 0.1);
    p.seed = (uint64_t)o.integer("seed", 1001);
    return p;
}

// Trains on --prefix (the body of main(), mainKernel.cu:30-386); the
// context is left with the trained model.
int run_training(const Opts &o, ezlda::Context &ctx) {
    if (!o.has("prefix")) throw ezlda::Error(EZLDA_EINVAL, "train needs --prefix DIR");
    const std::string prefix = o.str("prefix");
    const int iters = (int)o.integer("iters", 100);  // mainKernel.cu:51
    const int ll_every = (int)o.integer("ll-every", 0);
    const bool graph = o.has("graph");
    auto t0 = std::chrono::steady_clock::now();
    if (o.has("resume-tl"))
        ctx.load_tl(prefix);  // TL.txt topics and the saved sweep
    else
        ctx.load_chunked(prefix, (int)o.integer("chunks", 0));
    const ezlda_info inf = ctx.info();
    std::printf("loaded %s: %lld tokens, %d docs, %d words (%d dense), K=%d, %.2f GB on device, %.2f s\n",
                prefix.c_str(), (long long)inf.num_tokens, inf.num_docs, inf.num_words, inf.num_head_words, inf.K,
                inf.device_bytes / 1e9,
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    if (o.has("resume")) {
        ctx.load_checkpoint(o.str("resume"));
        std::printf("resumed at sweep %u\n", ctx.info().sweep);
    }
    if (o.has("autotune")) {
        ezlda_args best;
        double ms = 0.0;
        const auto tr = ctx.autotune((int)o.integer("tune-sweeps", 3), &best, &ms);
        for (const auto &t : tr)
            std::printf("tune: dense_words %d run_cap %d doc_tiles %d block_dim %d token_segment %d grid_dim %d: "
                        "%.3f ms/sweep\n", t.args.dense_words, t.args.run_cap, t.args.doc_tiles, t.args.block_dim,
                        t.args.token_segment, t.args.grid_dim, t.ms_per_sweep);
        std::printf("tuned layout: dense_words %d run_cap %d doc_tiles %d block_dim %d token_segment %d grid_dim %d "
                    "(%.3f ms/sweep; the %zu trials ran %u sweeps)\n", best.dense_words, best.run_cap,
                    best.doc_tiles, best.block_dim, best.token_segment, best.grid_dim, ms, tr.size(),
                    ctx.info().sweep);
    }
    std::ofstream rec(o.str("record", prefix + "/SamplingDRecord.txt"), std::ios::app);
    double total = 0.0;
    for (int it = 0; it < iters; ++it) {
        const double ms = ctx.time_sweeps(1, graph);
        total += ms;
        rec << ms << "\n";
        if (ll_every > 0 && (it + 1) % ll_every == 0) {
            const double ll = ctx.loglik();
            std::printf("sweep %d: %.3f ms, log-likelihood %.6e, perplexity %.3f\n", it + 1, ms, ll,
                        std::exp(-ll / (double)inf.num_tokens));
        }
    }
    const double ll = ctx.loglik();
    std::printf("%d sweeps in %.1f ms (%.3e tokens/s); log-likelihood %.6e; perplexity %.4f\n", iters, total,
                inf.num_tokens * (double)iters / (total / 1e3), ll, std::exp(-ll / (double)inf.num_tokens));
    if (o.has("checkpoint")) ctx.save_checkpoint(o.str("checkpoint"));
    if (o.has("checkpoint-tl")) {
        ctx.save_tl(prefix);
        std::printf("topics of sweep %u written to %s/chunk*/TL.txt\n", ctx.info().sweep, prefix.c_str());
    }
    if (o.has("export")) ctx.export_model(o.str("export"));
    return 0;
}

int cmd_train(const Opts &o) {
    ezlda::Context ctx(args_from(o));
    return run_training(o, ctx);
}

// Fold-in of a docword file's documents (word ids as in the training
// corpus) against the trained model: one line per document, "doc n_d
// topic:count ...", topics 0-based, by count.
int cmd_foldin(const Opts &o) {
    if (!o.has("requests")) throw ezlda::Error(EZLDA_EINVAL, "foldin needs --requests DOCWORD");
    ezlda::Context ctx(args_from(o));
    run_training(o, ctx);
    ezlda::Corpus req = ezlda::Corpus::read_docword(o.str("requests"));
    const int K = ctx.info().K;
    ezlda_server_args sa;
    std::memset(&sa, 0, sizeof sa);
    sa.max_slots = (int32_t)o.integer("slots", 1024);
    sa.sweeps = (int32_t)o.integer("foldin-sweeps", 20);
    sa.hbm_budget_bytes = (int64_t)o.integer("budget-mb", 1024) << 20;
    sa.seed = (uint64_t)o.integer("seed", 2002) + 7;
    ezlda::Server srv(ctx, sa);
    const int D = req.docs();
    const int64_t *off = req.doc_off();
    const int32_t *w = req.word_ids();
    {  // one batch: every document or none
        std::vector<uint32_t> ids((size_t)D);
        for (int d = 0; d < D; ++d) ids[d] = (uint32_t)d;
        srv.submit_many(ids, std::vector<int64_t>(off, off + D + 1), std::vector<int32_t>(w, w + off[D]));
    }
    srv.drain();
    std::FILE *f = o.has("out") ? std::fopen(o.str("out").c_str(), "w") : stdout;
    if (!f) throw ezlda::Error(EZLDA_EIO, "cannot open " + o.str("out"));
    std::vector<int32_t> cnt(K);
    std::vector<uint16_t> z;
    for (int d = 0; d < D; ++d) {
        z.assign((size_t)std::max<int64_t>(1, off[d + 1] - off[d]), 0);
        srv.wait((uint32_t)d, &z, &cnt);
        std::vector<std::pair<int32_t, int>> nz;
        for (int k = 0; k < K; ++k)
            if (cnt[k]) nz.push_back({-cnt[k], k});
        std::sort(nz.begin(), nz.end());
        std::fprintf(f, "%d %lld", d, (long long)(off[d + 1] - off[d]));
        for (auto &x : nz) std::fprintf(f, " %d:%d", x.second, -x.first);
        std::fputc('\n', f);
    }
    if (f != stdout) std::fclose(f);
    const ezlda_server_stats st = srv.stats();
    std::fprintf(stderr, "folded in %lld docs: %.1f docs/s, latency p50 %.2f ms p99 %.2f ms, %lld launches\n",
                 (long long)st.completed, st.docs_per_s, st.p50_ms, st.p99_ms, (long long)st.steps);
    return 0;
}

int cmd_gen(const Opts &o) {
    if (!o.has("out")) throw ezlda::Error(EZLDA_EINVAL, "gen needs --out DIR");
    const ezlda_gen_params p = gen_from(o);
    ezlda::Corpus c = ezlda::Corpus::generate(p);
    const int K = (int)o.integer("K", p.K_true);
    c.write_chunked(o.str("out"), K, (int)o.integer("chunks", 1), (uint64_t)o.integer("seed", 1001) + 1000);
    std::printf("wrote %s: %lld tokens, %d docs, V=%d, K=%d\n", o.str("out").c_str(), (long long)c.tokens(), c.docs(),
                c.words(), K);
    return 0;
}

int cmd_preprocess(const Opts &o) {
    if (!o.has("docword") || !o.has("out")) throw ezlda::Error(EZLDA_EINVAL, "preprocess needs --docword and --out");
    ezlda::Corpus c = ezlda::Corpus::read_docword(o.str("docword"));
    const int K = (int)o.integer("K", 1024);          // streamTestBalance.cpp:86
    const int chunks = (int)o.integer("chunks", 8);   // streamTestBalance.cpp:106
    c.write_chunked(o.str("out"), K, chunks, (uint64_t)o.integer("seed", 1));
    std::printf("wrote %s: %lld tokens, %d docs, V=%d, %d chunks\n", o.str("out").c_str(), (long long)c.tokens(),
                c.docs(), c.words(), chunks);
    return 0;
}

int cmd_bench(const Opts &o) {
    Opts g = o;
    if (!g.has("V")) g.kv["V"] = "102660";
    if (!g.has("D")) g.kv["D"] = "300000";
    if (!g.has("K-true")) g.kv["K-true"] = "1024";
    if (!g.has("len")) g.kv["len"] = "lognormal";
    if (!g.has("len-mean")) g.kv["len-mean"] = "330";
    if (!g.has("zipf")) g.kv["zipf"] = "1.07";
    if (!g.has("seed")) g.kv["seed"] = "1002";
    ezlda::Corpus c = ezlda::Corpus::generate(gen_from(g));
    ezlda_args a = args_from(o);
    if (!o.has("seed")) a.seed = 2002;
    if (!o.has("dense-words") && !o.has("dense-threshold")) a.dense_words = (c.words() + 99) / 100;
    ezlda::Context ctx(a);
    ctx.load(c);
    ctx.sweep((int)o.integer("warmup", 10));
    const int n = (int)o.integer("sweeps", 20);
    double st[EZLDA_NUM_STAGES];
    const double ms = ctx.time_sweeps(n, false, st);
    std::printf("{\"tokens\": %lld, \"K\": %d, \"sweeps\": %d, \"ms_per_sweep\": %.3f, \"tokens_per_s\": %.4e, "
                "\"stage_ms\": [%.3f, %.3f, %.3f, %.3f]}\n",
                (long long)c.tokens(), a.K, n, ms / n, c.tokens() * (double)n / (ms / 1e3), st[0] / n, st[1] / n,
                st[2] / n, st[3] / n);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || !std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
        std::fputs(kUsage, argc < 2 ? stderr : stdout);
        return argc < 2 ? 2 : 0;
    }
    const std::string cmd = argv[1];
    try {
        const Opts o = parse(argc, argv, 2);
        if (cmd == "train") return cmd_train(o);
        if (cmd == "foldin") return cmd_foldin(o);
        if (cmd == "gen") return cmd_gen(o);
        if (cmd == "preprocess") return cmd_preprocess(o);
        if (cmd == "bench") return cmd_bench(o);
        std::fprintf(stderr, "unknown command %s\n%s", cmd.c_str(), kUsage);
        return 2;
    } catch (const ezlda::Error &e) {
        std::fprintf(stderr, "ezlda %s: error %d: %s\n", cmd.c_str(), e.code(), e.what());
        return 1;
    }
}
