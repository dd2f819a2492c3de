// Host-side corpus sources for include/ezlda.h:
//   * a synthetic corpus from the LDA generative process (Philox-keyed, so
//     the same parameters give the same corpus on any machine and any
//     thread count);
//   * the UCI "docword" bag-of-words file that preprocess/ reads
//     (preprocess/streamTestBalance.cpp:90-145);
//   * the chunked text folder preprocess/ writes and EZLDA reads
//     (streamTestBalance.cpp:276-547; readers DataChunk.cu:50-100,
//     Doc.cu:45-108, DT.cu:26-100, WT.cu:133-178, WTChunk.cu:38-52).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <numeric>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "../../include/ezlda.h"
#include "ezlda_philox.h"

struct ezlda_corpus {
    int32_t V = 0, D = 0;
    std::vector<int64_t> doc_off{0};
    std::vector<int32_t> words;
    std::vector<uint16_t> topics;  // initial topics, 0-based (chunked input only)
};

namespace ezc {
void set_error(const std::string &m);
int exception_status(ezlda_ctx *c);
}

namespace {

void err(const std::string &m) { ezc::set_error(m); }

// Philox stream for one generator entity (a topic or a doc).
struct Rng {
    uint64_t seed;
    uint32_t entity, kind, ctr = 0;
    uint32_t buf[4] = {0, 0, 0, 0};
    int pos = 4;
    Rng(uint64_t s, uint32_t e, uint32_t k) : seed(s), entity(e), kind(k) {}
    uint32_t next() {
        if (pos == 4) {
            ez_u4 c{entity, kind, ctr++, EZ_TAG_GEN};
            ez_u4 o = ez_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            buf[0] = o.x; buf[1] = o.y; buf[2] = o.z; buf[3] = o.w;
            pos = 0;
        }
        return buf[pos++];
    }
    // (0,1), 53 bits
    double uniform() {
        const uint64_t a = next() >> 5, b = next() >> 6;
        return ((double)(a * 67108864ull + b) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // log of a Gamma(a, 1) draw (Marsaglia & Tsang 2000; a < 1 boosted by
    // U^(1/a), kept in log space so tiny shapes do not underflow).
    double log_gamma(double a) {
        double boost = 0.0;
        if (a < 1.0) {
            boost = std::log(uniform()) / a;
            a += 1.0;
        }
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x || std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v)))
                return std::log(d * v) + boost;
        }
    }
    int poisson(double lam) {
        if (lam > 500.0) return (int)std::lround(std::max(0.0, lam + std::sqrt(lam) * normal()));
        const double L = std::exp(-lam);
        int k = 0;
        double p = 1.0;
        do {
            ++k;
            p *= uniform();
        } while (p > L);
        return k - 1;
    }
};

// Dirichlet(conc) draw into out (normalised), via log-gammas.
void dirichlet(Rng &r, const std::vector<double> &conc, std::vector<double> &out) {
    const size_t n = conc.size();
    out.resize(n);
    double mx = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
        out[i] = r.log_gamma(conc[i]);
        mx = std::max(mx, out[i]);
    }
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        out[i] = std::exp(out[i] - mx);
        s += out[i];
    }
    for (size_t i = 0; i < n; ++i) out[i] /= s;
}

int threads_for(int requested) {
    int n = requested > 0 ? requested : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

template <class F>
void pfor(int n, int T, F f) {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([=] {
            for (int i = t; i < n; i += T) f(i);
        });
    for (auto &x : th) x.join();
}

bool read_ints(const std::string &path, std::vector<int64_t> &out, size_t expect) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        err("cannot open " + path);
        return false;
    }
    out.clear();
    int64_t x;
    while (f >> x) out.push_back(x);
    if (!f.eof()) {
        err("malformed number in " + path);
        return false;
    }
    if (expect != (size_t)-1 && out.size() != expect) {
        err(path + ": expected " + std::to_string(expect) + " values, found " + std::to_string(out.size()));
        return false;
    }
    return true;
}

bool make_dir(const std::string &p) {
    if (mkdir(p.c_str(), 0755) == 0 || errno == EEXIST) return true;
    err("cannot create directory " + p);
    return false;
}

// ---- binary sidecar of the chunked folder
// <prefix>/corpus.ezb holds the doc-major corpus exactly as the text reader
// returns it (docs of chunk 0 first; word ids are ranks; topics 0-based),
// with the size and modification time of every chunk's TL.txt,
// word2DocMap.txt and docCountOffset.txt: the reader takes it only while all
// of them still match, so a rewritten or edited text file falls back to the
// text path.  Layout (little-endian): "EZLDACB1", u32 version 1, u32 chunks,
// i64 V, D, N, per chunk and file i64 size and mtime (ns), i64 doc_off[D+1],
// i32 words[N], u16 topics[N].
constexpr char kSideMagic[8] = {'E', 'Z', 'L', 'D', 'A', 'C', 'B', '1'};
constexpr const char *kSideFiles[3] = {"TL.txt", "word2DocMap.txt", "docCountOffset.txt"};

bool file_stamp(const std::string &path, int64_t out[2]) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return false;
    out[0] = (int64_t)st.st_size;
    out[1] = (int64_t)st.st_mtim.tv_sec * 1000000000ll + (int64_t)st.st_mtim.tv_nsec;
    return true;
}

bool chunk_stamps(const std::string &root, int chunks, std::vector<int64_t> &st) {
    st.assign((size_t)chunks * 6, 0);
    for (int ch = 0; ch < chunks; ++ch)
        for (int f = 0; f < 3; ++f)
            if (!file_stamp(root + "/chunk" + std::to_string(ch) + "/" + kSideFiles[f], &st[(size_t)ch * 6 + 2 * f]))
                return false;
    return true;
}

std::string sidecar_path(const std::string &root) { return root + "/corpus.ezb"; }

// Writes the sidecar of a complete folder; false (and the error set) on an
// I/O failure.
bool write_sidecar(const std::string &root, int chunks, const ezlda_corpus &c) {
    std::vector<int64_t> st;
    if (!chunk_stamps(root, chunks, st)) {
        err("cannot stat the chunk files of " + root);
        return false;
    }
    const std::string path = sidecar_path(root), tmp = path + ".tmp";
    std::FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        err("cannot write " + tmp);
        return false;
    }
    const uint32_t hdr[2] = {1u, (uint32_t)chunks};
    const int64_t dims[3] = {c.V, c.D, c.doc_off.back()};
    bool ok = std::fwrite(kSideMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(dims, 8, 3, f) == 3 &&
              std::fwrite(st.data(), 8, st.size(), f) == st.size() &&
              std::fwrite(c.doc_off.data(), 8, c.doc_off.size(), f) == c.doc_off.size() &&
              std::fwrite(c.words.data(), 4, c.words.size(), f) == c.words.size() &&
              std::fwrite(c.topics.data(), 2, c.topics.size(), f) == c.topics.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        err("cannot write " + path);
        return false;
    }
    return true;
}

// The sidecar's corpus when it exists, describes this folder (V, D, N and
// chunk count as the text files give them) and every stamped file is
// unchanged; else nullptr (no error: the caller reads the text).
ezlda_corpus *read_sidecar(const std::string &root, int chunks, int64_t V, int64_t D, int64_t N) {
    std::FILE *f = std::fopen(sidecar_path(root).c_str(), "rb");
    if (!f) return nullptr;
    ezlda_corpus *c = nullptr;
    char magic[8];
    uint32_t hdr[2];
    int64_t dims[3];
    std::vector<int64_t> st((size_t)chunks * 6), now;
    if (std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kSideMagic, 8) == 0 && std::fread(hdr, 4, 2, f) == 2 &&
        hdr[0] == 1u && hdr[1] == (uint32_t)chunks && std::fread(dims, 8, 3, f) == 3 && dims[0] == V &&
        dims[1] == D && dims[2] == N && std::fread(st.data(), 8, st.size(), f) == st.size() &&
        chunk_stamps(root, chunks, now) && now == st) {
        c = new ezlda_corpus();
        c->V = (int32_t)V;
        c->D = (int32_t)D;
        c->doc_off.resize((size_t)D + 1);
        c->words.resize((size_t)N);
        c->topics.resize((size_t)N);
        const bool ok = std::fread(c->doc_off.data(), 8, c->doc_off.size(), f) == c->doc_off.size() &&
                        std::fread(c->words.data(), 4, c->words.size(), f) == c->words.size() &&
                        std::fread(c->topics.data(), 2, c->topics.size(), f) == c->topics.size() &&
                        c->doc_off[0] == 0 && c->doc_off[(size_t)D] == N;
        bool sane = ok;
        for (size_t d = 0; sane && d < (size_t)D; ++d) sane = c->doc_off[d] <= c->doc_off[d + 1];
        for (size_t i = 0; sane && i < (size_t)N; ++i) sane = c->words[i] >= 0 && c->words[i] < V;
        if (!sane) {
            delete c;
            c = nullptr;
        }
    }
    std::fclose(f);
    return c;
}

// The chunked folder's text files read into a doc-major corpus (chunk 0's
// docs first, local doc order; word ids are ranks; topics from TL.txt made
// 0-based).  pos (nullable) receives, per doc-major token, chunk << 32 | its
// position in the chunk's TL.txt.  tll, dl: per-chunk tokens and docs.
ezlda_corpus *read_chunked_text(const std::string &root, int num_chunks, int64_t V, const std::vector<int64_t> &tll,
                                const std::vector<int64_t> &dl, std::vector<uint64_t> *pos) {
    auto *c = new ezlda_corpus();
    c->V = (int32_t)V;
    if (pos) pos->clear();
    for (int ch = 0; ch < num_chunks; ++ch) {
        const std::string dir = root + "/chunk" + std::to_string(ch);
        const int64_t Nc = tll[ch], Dc = dl[ch];
        if (Nc < 0 || Dc < 0 || Nc > 0xFFFFFFFFll || Dc > 0x7FFFFFFFll) {
            err(root + ": chunk " + std::to_string(ch) + " has a negative or oversized token or document count");
            delete c;
            return nullptr;
        }
        std::vector<int64_t> tl, w2d, wco, dco;
        if (!read_ints(dir + "/TL.txt", tl, (size_t)Nc) || !read_ints(dir + "/word2DocMap.txt", w2d, (size_t)Nc) ||
            !read_ints(dir + "/wordCountOffset.txt", wco, (size_t)(2 * V)) ||
            !read_ints(dir + "/docCountOffset.txt", dco, (size_t)(2 * Dc))) {
            delete c;
            return nullptr;
        }
        // rank of each TL position
        std::vector<int32_t> rank_at(Nc, -1);
        for (int64_t r = 0; r < V; ++r) {
            const int64_t n = wco[2 * r], o = wco[2 * r + 1];
            if (n < 0 || o < 0 || o + n > Nc) {
                err(dir + "/wordCountOffset.txt: range out of bounds at rank " + std::to_string(r));
                delete c;
                return nullptr;
            }
            for (int64_t i = o; i < o + n; ++i) rank_at[i] = (int32_t)r;
        }
        int64_t expect = 0;
        for (int64_t ld = 0; ld < Dc; ++ld) {
            const int64_t n = dco[2 * ld], o = dco[2 * ld + 1];
            if (o != expect || n < 0 || o + n > Nc) {
                err(dir + "/docCountOffset.txt: inconsistent entry " + std::to_string(ld));
                delete c;
                return nullptr;
            }
            expect += n;
            for (int64_t i = o; i < o + n; ++i) {
                const int64_t p = w2d[i];
                if (p < 0 || p >= Nc || rank_at[p] < 0 || tl[p] < 1 || tl[p] > 65535) {
                    err(dir + ": bad token " + std::to_string(i));
                    delete c;
                    return nullptr;
                }
                c->words.push_back(rank_at[p]);
                c->topics.push_back((uint16_t)(tl[p] - 1));
                if (pos) pos->push_back(((uint64_t)ch << 32) | (uint64_t)p);
            }
            c->doc_off.push_back((int64_t)c->words.size());
        }
        if (expect != Nc) {
            err(dir + ": docCountOffset covers " + std::to_string(expect) + " of " + std::to_string(Nc) + " tokens");
            delete c;
            return nullptr;
        }
        c->D += (int32_t)Dc;
    }
    return c;
}

// lengthVec.txt's V, and TLLength.txt / docLength.txt of the folder.
bool read_folder_sizes(const std::string &root, int32_t &num_chunks, int64_t &V, std::vector<int64_t> &tll,
                       std::vector<int64_t> &dl) {
    std::vector<int64_t> lv;
    if (!read_ints(root + "/lengthVec.txt", lv, 8)) return false;
    V = lv[4];
    if (!read_ints(root + "/TLLength.txt", tll, (size_t)-1)) return false;
    if (num_chunks <= 0) num_chunks = (int32_t)tll.size();
    if ((int64_t)tll.size() != num_chunks) {
        err(root + "/TLLength.txt: " + std::to_string(tll.size()) + " chunks, expected " + std::to_string(num_chunks));
        return false;
    }
    if (!read_ints(root + "/docLength.txt", dl, (size_t)num_chunks)) return false;
    if (V < 1 || V > 0x7fffffff) {
        err(root + "/lengthVec.txt: bad vocabulary size");
        return false;
    }
    for (int ch = 0; ch < num_chunks; ++ch)
        if (tll[ch] < 0 || dl[ch] < 0 || tll[ch] > 0xFFFFFFFFll || dl[ch] > 0x7FFFFFFFll) {
            err(root + ": chunk " + std::to_string(ch) + " has a negative or oversized token or document count");
            return false;
        }
    return true;
}

}  // namespace

namespace ezc {

// The chunked folder at root: per doc-major token (the reader's order),
// chunk << 32 | its position in the chunk's TL.txt, and its word (rank);
// tll: tokens per chunk.  false with the error set.
bool chunked_layout(const std::string &root, std::vector<uint64_t> &pos, std::vector<int32_t> &words,
                    std::vector<int64_t> &tll) {
    int32_t chunks = 0;
    int64_t V = 0;
    std::vector<int64_t> dl;
    if (!read_folder_sizes(root, chunks, V, tll, dl)) return false;
    ezlda_corpus *c = read_chunked_text(root, chunks, V, tll, dl, &pos);
    if (!c) return false;
    words = std::move(c->words);
    delete c;
    return true;
}

// Rewrites every chunk's TL.txt with the doc-major topics z (0-based; one per
// entry of pos, as chunked_layout gives it), 1-based as the reference writes
// them, and the folder's binary sidecar with the same topics when it was
// current before.  false with the error set.
bool rewrite_tl(const std::string &root, const std::vector<uint64_t> &pos, const std::vector<int64_t> &tll,
                const uint16_t *z, int K) {
    const int chunks = (int)tll.size();
    int64_t N = 0, D = 0;
    std::vector<int64_t> dl;
    {
        int32_t nc = chunks;
        int64_t V = 0;
        std::vector<int64_t> t2;
        if (!read_folder_sizes(root, nc, V, t2, dl)) return false;
        for (int ch = 0; ch < chunks; ++ch) {
            N += tll[ch];
            D += dl[ch];
        }
        ezlda_corpus *side = read_sidecar(root, chunks, V, D, N);
        std::vector<std::vector<uint16_t>> tl(chunks);
        for (int ch = 0; ch < chunks; ++ch) tl[ch].assign((size_t)tll[ch], 0);
        for (size_t i = 0; i < pos.size(); ++i) {
            const int ch = (int)(pos[i] >> 32);
            const uint32_t p = (uint32_t)pos[i];
            if (z[i] >= K) {
                err("topic out of range at token " + std::to_string(i));
                delete side;
                return false;
            }
            tl[ch][p] = (uint16_t)(z[i] + 1);
        }
        for (int ch = 0; ch < chunks; ++ch) {
            const std::string path = root + "/chunk" + std::to_string(ch) + "/TL.txt", tmp = path + ".tmp";
            std::string buf;
            buf.reserve(tl[ch].size() * 6);
            char num[16];
            for (const uint16_t t : tl[ch]) {
                if (t == 0) {
                    err(path + ": a position without a token");
                    delete side;
                    return false;
                }
                const int n = std::snprintf(num, sizeof num, "%u\n", (unsigned)t);
                buf.append(num, (size_t)n);
            }
            std::FILE *f = std::fopen(tmp.c_str(), "wb");
            const bool ok = f && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            if (f && std::fclose(f) != 0) f = nullptr;
            if (!ok || !f || std::rename(tmp.c_str(), path.c_str()) != 0) {
                std::remove(tmp.c_str());
                err("cannot write " + path);
                delete side;
                return false;
            }
        }
        if (side) {
            std::memcpy(side->topics.data(), z, sizeof(uint16_t) * (size_t)N);
            const bool ok = write_sidecar(root, chunks, *side);
            delete side;
            if (!ok) return false;
        }
    }
    return true;
}

}  // namespace ezc

extern "C" {

ezlda_corpus *ezlda_gen_corpus(const ezlda_gen_params *p) try {
    if (!p || p->V < 1 || p->D < 0 || p->K_true < 1 || p->first_doc < 0) {
        err("bad generator parameters");
        return nullptr;
    }
    const int V = p->V, D = p->D, KT = p->K_true;
    const int T = threads_for(p->threads);
    const int lo = std::max(1, p->len_min), hi = p->len_max > 0 ? p->len_max : 65535;
    auto *c = new ezlda_corpus();
    c->V = V;
    c->D = D;
    // base word distribution
    std::vector<double> base(V);
    for (int w = 0; w < V; ++w) base[w] = p->zipf_s > 0 ? std::pow((double)(w + 1), -p->zipf_s) : 1.0;
    const double bs = std::accumulate(base.begin(), base.end(), 0.0);
    std::vector<double> conc(V);
    for (int w = 0; w < V; ++w) conc[w] = std::max(p->phi_conc * (double)V * base[w] / bs, 1e-300);
    // phi_k ~ Dir(conc): one CDF per true topic
    std::vector<std::vector<double>> cdf(KT);
    pfor(KT, T, [&](int k) {
        Rng r(p->seed, (uint32_t)k, 1u);
        std::vector<double> phi;
        dirichlet(r, conc, phi);
        cdf[k].resize(V);
        double s = 0.0;
        for (int w = 0; w < V; ++w) {
            s += phi[w];
            cdf[k][w] = s;
        }
    });
    // doc lengths
    std::vector<int32_t> len(D);
    pfor(D, T, [&](int d) {
        Rng r(p->seed, (uint32_t)(d + p->first_doc), 2u);
        double x;
        if (p->len_dist == 0) {
            x = r.poisson(p->len_mean);
        } else if (p->len_dist == 1) {
            const double mu = std::log(p->len_mean) - 0.5 * p->len_sigma * p->len_sigma;
            x = std::round(std::exp(mu + p->len_sigma * r.normal()));
        } else {
            x = std::floor(p->pareto_xm / std::pow(r.uniform(), 1.0 / p->pareto_shape));
        }
        if (!(x >= lo)) x = lo;
        if (x > hi) x = hi;
        len[d] = (int32_t)x;
    });
    c->doc_off.assign(D + 1, 0);
    for (int d = 0; d < D; ++d) c->doc_off[d + 1] = c->doc_off[d] + len[d];
    c->words.resize(c->doc_off[D]);
    std::vector<double> tconc(KT, p->theta_conc);
    pfor(D, T, [&](int d) {
        Rng r(p->seed, (uint32_t)(d + p->first_doc), 3u);
        std::vector<double> theta;
        dirichlet(r, tconc, theta);
        for (int k = 1; k < KT; ++k) theta[k] += theta[k - 1];
        for (int64_t i = c->doc_off[d]; i < c->doc_off[d + 1]; ++i) {
            const double u1 = r.uniform() * theta[KT - 1];
            int k = (int)(std::upper_bound(theta.begin(), theta.end(), u1) - theta.begin());
            if (k >= KT) k = KT - 1;
            const std::vector<double> &cd = cdf[k];
            const double u2 = r.uniform() * cd[V - 1];
            int w = (int)(std::upper_bound(cd.begin(), cd.end(), u2) - cd.begin());
            if (w >= V) w = V - 1;
            c->words[i] = w;
        }
    });
    return c;
} catch (...) {
    (void)ezc::exception_status(nullptr);
    return nullptr;
}

int64_t ezlda_corpus_num_tokens(const ezlda_corpus *c) { return c ? c->doc_off.back() : 0; }
int32_t ezlda_corpus_num_docs(const ezlda_corpus *c) { return c ? c->D : 0; }
int32_t ezlda_corpus_num_words(const ezlda_corpus *c) { return c ? c->V : 0; }
const int64_t *ezlda_corpus_doc_off(const ezlda_corpus *c) { return c ? c->doc_off.data() : nullptr; }
const int32_t *ezlda_corpus_words(const ezlda_corpus *c) { return c ? c->words.data() : nullptr; }
const uint16_t *ezlda_corpus_topics(const ezlda_corpus *c) {
    return (c && !c->topics.empty()) ? c->topics.data() : nullptr;
}
void ezlda_corpus_free(ezlda_corpus *c) { delete c; }

ezlda_corpus *ezlda_corpus_from_arrays(int32_t V, int32_t D, const int64_t *doc_off, const int32_t *words,
                                       const uint16_t *topics) {
    if (V < 1 || D < 0 || !doc_off || doc_off[0] != 0) {
        err("bad corpus arrays");
        return nullptr;
    }
    auto *c = new ezlda_corpus();
    c->V = V;
    c->D = D;
    c->doc_off.assign(doc_off, doc_off + D + 1);
    c->words.assign(words, words + doc_off[D]);
    if (topics) c->topics.assign(topics, topics + doc_off[D]);
    return c;
}

// UCI docword: "D W NNZ" then NNZ lines "doc word count", 1-based, grouped
// by doc (the input of preprocess/streamTestBalance.cpp:90-145).
ezlda_corpus *ezlda_read_docword(const char *path) try {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        err(std::string("cannot open ") + path);
        return nullptr;
    }
    int64_t D, W, NNZ;
    if (!(f >> D >> W >> NNZ) || D < 0 || W < 1 || NNZ < 0) {
        err(std::string("bad docword header in ") + path);
        return nullptr;
    }
    auto *c = new ezlda_corpus();
    c->V = (int32_t)W;
    c->D = (int32_t)D;
    std::vector<int64_t> cnt(D, 0);
    std::vector<std::pair<int32_t, int32_t>> trip;
    trip.reserve(NNZ);
    std::vector<int32_t> docs;
    docs.reserve(NNZ);
    for (int64_t i = 0; i < NNZ; ++i) {
        int64_t d, w, n;
        if (!(f >> d >> w >> n) || d < 1 || d > D || w < 1 || w > W || n < 0) {
            err(std::string("bad or truncated docword triple ") + std::to_string(i) + " in " + path);
            delete c;
            return nullptr;
        }
        cnt[d - 1] += n;
        docs.push_back((int32_t)(d - 1));
        trip.emplace_back((int32_t)(w - 1), (int32_t)n);
    }
    c->doc_off.assign(D + 1, 0);
    for (int64_t d = 0; d < D; ++d) c->doc_off[d + 1] = c->doc_off[d] + cnt[d];
    c->words.resize(c->doc_off[D]);
    std::vector<int64_t> cur(c->doc_off.begin(), c->doc_off.end() - 1);
    for (size_t i = 0; i < trip.size(); ++i)
        for (int32_t k = 0; k < trip[i].second; ++k) c->words[cur[docs[i]]++] = trip[i].first;
    return c;
} catch (...) {
    (void)ezc::exception_status(nullptr);
    return nullptr;
}

// Writes the chunked folder of preprocess/streamTestBalance.cpp:356-547.
// Deviations from the reference writer: docs go to chunks by ascending
// length with no 1000-token cap on the length histogram (:117,135), the
// off-by-one docCountVec[j+1] (:145) is not reproduced, and TL.txt holds the
// given initial topics (or Philox draws) instead of unseeded rand() (:399).
int ezlda_write_chunked(const char *prefix, const ezlda_corpus *c, const uint16_t *z0, int32_t K, int32_t num_chunks,
                        uint64_t seed) try {
    if (!c || !prefix || K < 1 || num_chunks < 1) {
        err("bad arguments to ezlda_write_chunked");
        return EZLDA_EINVAL;
    }
    const int V = c->V, D = c->D;
    const int64_t N = c->doc_off[D];
    const std::string root(prefix);
    if (!make_dir(root)) return EZLDA_EIO;
    // word ranks: count desc, stable by id (:281-301)
    std::vector<int64_t> wc(V, 0);
    for (int64_t i = 0; i < N; ++i) wc[c->words[i]]++;
    std::vector<int32_t> ids(V);
    std::iota(ids.begin(), ids.end(), 0);
    std::stable_sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) { return wc[a] > wc[b]; });
    std::vector<int32_t> rank(V);
    for (int r = 0; r < V; ++r) rank[ids[r]] = r;
    int VD = 0;
    while (VD < V && wc[ids[VD]] >= K) ++VD;
    const int VS = V - VD;
    // doc -> chunk: ascending length, cumulative token balance (:148-168)
    std::vector<int32_t> order(D);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return c->doc_off[a + 1] - c->doc_off[a] < c->doc_off[b + 1] - c->doc_off[b];
    });
    std::vector<int32_t> chunk_of(D, 0);
    int64_t acc = 0;
    for (int32_t d : order) {
        int ch = (int)std::min<int64_t>(num_chunks - 1, N > 0 ? acc * num_chunks / N : 0);
        chunk_of[d] = ch;
        acc += c->doc_off[d + 1] - c->doc_off[d];
    }
    std::vector<int64_t> TLLen(num_chunks, 0), DTLen(num_chunks, 0), WTLen(num_chunks, 0), nDocs(num_chunks, 0),
        nHead(num_chunks, 0), nTail(num_chunks, 0);
    std::vector<int64_t> gWT(V, 0);
    int64_t maxTL = 0, maxDocs = 0;
    ezlda_corpus side;  // the corpus as ezlda_read_chunked returns it
    side.V = V;
    side.words.reserve((size_t)N);
    side.topics.reserve((size_t)N);
    for (int ch = 0; ch < num_chunks; ++ch) {
        const std::string dir = root + "/chunk" + std::to_string(ch);
        if (!make_dir(dir)) return EZLDA_EIO;
        std::vector<int32_t> docs;
        for (int d = 0; d < D; ++d)
            if (chunk_of[d] == ch) docs.push_back(d);
        // doc-major token list of the chunk
        struct Tok {
            int32_t doc, rank;
            uint16_t topic;
            int64_t dm;
        };
        std::vector<Tok> toks;
        std::vector<int64_t> dcnt(docs.size()), doff(docs.size());
        for (size_t ld = 0; ld < docs.size(); ++ld) {
            const int d = docs[ld];
            doff[ld] = (int64_t)toks.size();
            for (int64_t i = c->doc_off[d]; i < c->doc_off[d + 1]; ++i) {
                uint16_t t;
                if (z0)
                    t = z0[i];
                else
                    t = (uint16_t)(ez_philox_x0(seed, (uint64_t)i, 0u, EZ_TAG_INIT) % (uint32_t)K);
                toks.push_back({(int32_t)ld + 1, rank[c->words[i]], t, (int64_t)toks.size()});
            }
            dcnt[ld] = (int64_t)toks.size() - doff[ld];
        }
        // the reader's doc-major order of this chunk, for the binary sidecar
        for (const Tok &t : toks) {
            side.words.push_back(t.rank);
            side.topics.push_back(t.topic);
        }
        for (size_t ld = 0; ld < docs.size(); ++ld) side.doc_off.push_back(side.doc_off.back() + dcnt[ld]);
        side.D += (int32_t)docs.size();
        std::vector<Tok> wm(toks);
        std::stable_sort(wm.begin(), wm.end(), [](const Tok &a, const Tok &b) { return a.rank < b.rank; });
        std::vector<int64_t> w2d(toks.size());
        for (size_t i = 0; i < wm.size(); ++i) w2d[wm[i].dm] = (int64_t)i;
        std::vector<int64_t> rc(V, 0);
        for (auto &t : wm) rc[t.rank]++;
        {
            std::ofstream f(dir + "/TL.txt", std::ios::binary), g(dir + "/doc2WordMap.txt", std::ios::binary);
            for (auto &t : wm) {
                f << (int)t.topic + 1 << "\n";
                g << t.doc << "\n";
            }
        }
        {
            std::ofstream f(dir + "/word2DocMap.txt", std::ios::binary);
            for (int64_t x : w2d) f << x << "\n";
        }
        {
            std::ofstream f(dir + "/wordCountOffset.txt", std::ios::binary);
            int64_t o = 0;
            for (int r = 0; r < V; ++r) {
                f << rc[r] << " " << o << "\n";
                o += rc[r];
            }
        }
        {
            std::ofstream f(dir + "/docCountOffset.txt", std::ios::binary), g(dir + "/DTCountOffset.txt", std::ios::binary);
            int64_t o = 0;
            for (size_t ld = 0; ld < docs.size(); ++ld) {
                f << dcnt[ld] << " " << doff[ld] << "\n";
                const int64_t cap = std::min<int64_t>(((dcnt[ld] - 1) / 32 + 1) * 32, K);
                g << cap << " " << o << "\n";
                o += cap;
            }
            DTLen[ch] = o;
        }
        {
            std::ofstream f(dir + "/WTCountOffset.txt", std::ios::binary);
            int64_t o = 0;
            for (int r = VD; r < V; ++r) {
                const int64_t cap = std::min<int64_t>(rc[r], K);
                f << cap << " " << o << "\n";
                o += cap;
            }
            WTLen[ch] = o;
        }
        for (int r = 0; r < V; ++r) {
            (r < VD ? nHead[ch] : nTail[ch]) += rc[r];
            gWT[r] += rc[r];
        }
        TLLen[ch] = (int64_t)toks.size();
        nDocs[ch] = (int64_t)docs.size();
        maxTL = std::max(maxTL, TLLen[ch]);
        maxDocs = std::max(maxDocs, nDocs[ch]);
    }
    int64_t maxWT = 0;
    {
        std::ofstream f(root + "/WTCountOffset.txt", std::ios::binary);
        int64_t o = 0, oD = 0;
        for (int r = 0; r < V; ++r) {
            const int64_t cap = std::min<int64_t>(gWT[r], K);
            if (r == VD) oD = o;
            f << cap << " " << o << "\n";
            o += cap;
        }
        if (VD == V) oD = o;
        maxWT = o - oD;
    }
    auto put = [&](const char *name, const std::vector<int64_t> &a, const std::vector<int64_t> *b) {
        std::ofstream f(root + "/" + name, std::ios::binary);
        for (int ch = 0; ch < num_chunks; ++ch) {
            f << a[ch];
            if (b) f << " " << (*b)[ch];
            f << "\n";
        }
    };
    put("docLength.txt", nDocs, nullptr);
    put("TLLength.txt", TLLen, nullptr);
    put("DTLength.txt", DTLen, nullptr);
    put("WTLength.txt", WTLen, nullptr);
    put("TLSplit.txt", nHead, &nTail);
    {
        std::ofstream f(root + "/lengthVec.txt", std::ios::binary);
        f << maxTL << " " << *std::max_element(DTLen.begin(), DTLen.end()) << " " << maxWT << " " << maxDocs << " " << V
          << " " << *std::max_element(WTLen.begin(), WTLen.end()) << " " << VD << " " << VS;
        if (!f) {
            err("cannot write " + root + "/lengthVec.txt");
            return EZLDA_EIO;
        }
    }
    if (!write_sidecar(root, num_chunks, side)) return EZLDA_EIO;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(nullptr);
}

// Reads the chunked folder back into a doc-major corpus (docs of chunk 0
// first, local doc order inside a chunk; word ids are the ranks; topics from
// TL.txt made 0-based), from its binary sidecar while that is current, else
// from the text files.  num_chunks <= 0 takes the count from TLLength.txt.
ezlda_corpus *ezlda_read_chunked(const char *prefix, int32_t num_chunks) try {
    const std::string root(prefix ? prefix : "");
    int64_t V = 0;
    std::vector<int64_t> tll, dl;
    if (!read_folder_sizes(root, num_chunks, V, tll, dl)) return nullptr;
    int64_t N = 0, D = 0;
    for (int ch = 0; ch < num_chunks; ++ch) {
        N += tll[ch];
        D += dl[ch];
    }
    if (ezlda_corpus *c = read_sidecar(root, num_chunks, V, D, N)) return c;
    return read_chunked_text(root, num_chunks, V, tll, dl, nullptr);
} catch (...) {
    (void)ezc::exception_status(nullptr);
    return nullptr;
}

}  // extern "C"
