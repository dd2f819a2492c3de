// Host runtime behind include/ezlda.h: context, word-major corpus layout,
// device allocation, stage and sweep entry points, hipGraph replay, timing,
// checkpoints.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ezlda.h"
#include "ezlda_device.h"
#include "ezlda_internal.h"
#include "ezlda_philox.h"

namespace {

thread_local std::string g_err;

// Host-side threads for the layout builder (0 = hardware concurrency).
int host_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 1;
    if (n > 16) n = 16;
    return (int)n;
}

template <class F>
void parallel_for(int64_t n, F f) {
    const int T = host_threads();
    if (n < 65536 || T == 1) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int64_t b = n * t / T, e = n * (t + 1) / T;
        th.emplace_back([=] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

// DT rows: 16-bit entries for K <= kCompactK (the oracle's ORC_COMPACT_K),
// tb = max(10, ceil(log2 K)) topic bits, counts in pieces of at most
// 2^(16 - tb) - 1.
constexpr uint32_t kCompactK = 8192;

uint32_t topic_bits(int K) {
    uint32_t b = 10;
    while ((1 << b) < K) ++b;
    return b;
}

uint32_t dt_piece(int K) { return (1u << (16u - topic_bits(K))) - 1u; }

// Entries of a doc's DT row at most: min(n, K), plus one per full piece in
// compact rows.
uint32_t dt_cap(uint32_t n, int K) {
    const uint32_t m = std::min<uint32_t>(n, (uint32_t)K);
    if ((uint32_t)K > kCompactK) return m;
    return std::min<uint32_t>(m + n / dt_piece(K), n);
}

// Device address of the static log-likelihood array on each device, looked
// up once per device (with that device current) by the first ezlda_create
// there; kept for the rest of the process, as the array itself is.
constexpr int kMaxDevices = 64;
struct ScratchSlot {
    bool looked_up = false;
    double *ptr = nullptr;
    std::string err;
};
std::mutex g_scratch_mu;
std::array<ScratchSlot, kMaxDevices> g_scratch;

void scratch_for_current_device(double **ptr, std::string *err) {
    int dev = 0;
    *ptr = nullptr;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *err = "no device index for the static scratch array";
        return;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchSlot &sl = g_scratch[dev];
    if (!sl.looked_up) {
        const hipError_t e = ez::scratch_address(&sl.ptr);
        if (e != hipSuccess || !sl.ptr) {
            sl.ptr = nullptr;
            sl.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        }
        sl.looked_up = true;
    }
    *ptr = sl.ptr;
    *err = sl.err;
}

// Compute units of the current device (256 on MI355X).
int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || n < 1)
        n = 256;
    return n;
}

// LDS bytes per compute unit of the current device (160 KiB on MI355X).
int device_lds_per_cu() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || n < 1)
        n = 160 * 1024;
    return n;
}

// Validates the layout arguments of `a` (those build_layout reads) and
// fills their defaults; returns an error message or "".
std::string check_layout_args(ezlda_args &a) {
    // auto: the smallest workgroup (64, 128, ... threads) with which the
    // sampler workgroups whose LDS fits in a CU hold at least 12 wavefronts
    // (measured best for K = 1024 to 16383: 64 threads up to K ~ 2400, then
    // 128, 256 at K = 10000, 512 at K = 16383)
    if (a.block_dim <= 0) {
        const size_t lds = ez::sample_lds_bytes(a.K, 64);
        const int fit = std::max(1, (int)((size_t)device_lds_per_cu() / std::max<size_t>(lds, 1)));
        int bd = 64;
        while (bd < 1024 && fit * (bd / 64) < 12) bd *= 2;
        a.block_dim = bd;
    }
    if (a.block_dim % 64 != 0 || a.block_dim > 1024) return "block_dim must be a multiple of 64 and <= 1024";
    if (a.token_segment <= 0) a.token_segment = 4096;
    // auto: 16 persistent workgroups per CU at 64 threads, else 8
    if (a.grid_dim < 0) a.grid_dim = (a.block_dim == 64 ? 16 : 8) * device_cus();
    if (a.run_cap == 0) a.run_cap = 4;
    if (a.run_cap < 1 || a.run_cap > 32) return "run_cap must be in [1, 32]";
    return "";
}

}  // namespace

namespace ezc {
void set_error(const std::string &m) { g_err = m; }
}  // namespace ezc

namespace {

int fail(ezlda_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

}  // namespace

int ezc::exception_status(ezlda_ctx *c) {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(c, EZLDA_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(c, EZLDA_EINVAL, std::string("unexpected error: ") + e.what());
    } catch (...) {
        return fail(c, EZLDA_EINVAL, "unexpected error");
    }
}

namespace {

#define EZ_TRY_HIP(ctx, call)                                                                    \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), EZLDA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <class T>
int dalloc(ezlda_ctx *c, T **p, size_t n) {
    *p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return fail(c, EZLDA_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back((void *)*p);
    c->dev_bytes += (int64_t)bytes;
    return EZLDA_OK;
}

template <class T>
int upload(ezlda_ctx *c, T **p, const std::vector<T> &v) {
    int r = dalloc(c, p, v.size());
    if (r) return r;
    if (!v.empty()) EZ_TRY_HIP(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return EZLDA_OK;
}

void free_device(ezlda_ctx *c) {
    // release: nothing to do about a failure here
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    for (void *p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (void *p : c->layout_allocs) (void)hipFree(p);
    c->layout_allocs.clear();
    c->dev_bytes = 0;
    c->layout_bytes = 0;
    c->s = ez::DevState{};
    for (EzChunk &ch : c->chunks)
        if (ch.img) (void)hipHostFree(ch.img);
    c->chunks.clear();
    c->chunk_slot0.clear();
    c->tok_slot_g.clear();
    for (int j = 0; j < 2; ++j) {
        c->cimg[j] = nullptr;
        c->cdt[j] = c->caux[j] = c->cctr[j] = nullptr;
        c->cll[j] = c->cllo[j] = nullptr;
    }
    c->wt_next = c->nk_next = nullptr;
    c->loaded = false;
}

template <class T>
int ldalloc(ezlda_ctx *c, T **p, size_t n) {
    *p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) return fail(c, EZLDA_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->layout_allocs.push_back((void *)*p);
    c->layout_bytes += (int64_t)bytes;
    return EZLDA_OK;
}

template <class T>
int lupload(ezlda_ctx *c, T **p, const std::vector<T> &v) {
    int r = ldalloc(c, p, v.size());
    if (r) return r;
    if (!v.empty()) EZ_TRY_HIP(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return EZLDA_OK;
}

int need_loaded(ezlda_ctx *c) {
    if (!c) return fail(nullptr, EZLDA_EINVAL, "null context");
    if (!c->loaded) return fail(c, EZLDA_ESTATE, "no corpus loaded");
    return EZLDA_OK;
}

// Sampler work items (word, first token, end token).  Head words: the doc
// order (by length) is cut into T tiles of equal head-token work, T = n_xcd * 2^a chosen so one
// tile's DT rows fit in an XCD's L2; XCD x owns tiles [x*T/n_xcd, ...).  A
// head word gets one item per group of tiles (finer for frequent words, so
// an item keeps >= kMinItemTokens tokens); each XCD's items are sorted by
// (tile, word) and the lists are interleaved so that workgroup 8j+x, which
// round-robin dispatch sends to XCD x, is XCD x's j-th item.  Short lists
// are padded with empty items (word 0xFFFFFFFF).  Tail words follow, one
// item per word.  Items longer than token_segment are split.  The order of
// items changes which workgroup draws a token, never the draw.
std::vector<uint4> build_sample_items(const ezlda_args &a, const std::vector<uint32_t> &word_off,
                                      const std::vector<uint32_t> &tok_rank, const std::vector<uint32_t> &soff,
                                      const std::vector<uint32_t> &slot_wrd, int VD, int K, int n_xcd,
                                      int *tiles_used) {
    const uint32_t SEG = (uint32_t)a.token_segment;
    const int V = (int)word_off.size() - 1;
    const int D = (int)soff.size() - 1;
    std::vector<uint4> out;
    auto push_split = [&](std::vector<uint4> &v, uint32_t w, uint32_t b, uint32_t e) {
        for (uint32_t x = b; x < e; x += SEG) v.push_back(make_uint4(w, x, std::min(x + SEG, e), 0u));
    };
    int T = 0;
    uint64_t head_tokens = word_off[VD];
    if (a.doc_tiles >= 0 && VD > 0 && D > 0 && n_xcd > 0 && head_tokens > 0) {
        // DT bytes of the whole corpus -> tiles of ~2 MiB
        uint64_t dt_bytes = 0;
        for (int q = 0; q < D; ++q) dt_bytes += 4ull * std::min<uint32_t>(soff[q + 1] - soff[q], (uint32_t)K);
        if (a.doc_tiles > 0) {
            T = a.doc_tiles;
        } else {
            T = n_xcd;
            while ((uint64_t)T * (2u << 20) < dt_bytes && T < 4096) T *= 2;
        }
        T = std::max(n_xcd, (T / n_xcd) * n_xcd);
    }
    if (T > 0) {
        // tile boundaries: equal head-token work per tile
        std::vector<uint64_t> hw(D + 1, 0);  // head tokens of the first q docs in doc order
        for (int q = 0; q < D; ++q) {
            uint64_t h = 0;
            for (uint32_t i = soff[q]; i < soff[q + 1]; ++i) h += slot_wrd[i] < (uint32_t)VD;
            hw[q + 1] = hw[q] + h;
        }
        std::vector<uint32_t> tile_doc(T + 1, 0);
        for (int t = 0; t <= T; ++t) {
            const uint64_t target = hw[D] * (uint64_t)t / (uint64_t)T;
            tile_doc[t] = (uint32_t)(std::lower_bound(hw.begin(), hw.end(), target) - hw.begin());
        }
        tile_doc[0] = 0;
        tile_doc[T] = (uint32_t)D;
        const int per_xcd = T / n_xcd;
        const uint32_t kMinItemTokens = 512;
        // per XCD: (tile, word, first, end)
        std::vector<std::vector<std::array<uint32_t, 4>>> lists(n_xcd);
        for (int w = 0; w < VD; ++w) {
            const uint32_t b = word_off[w], e = word_off[w + 1];
            const uint32_t n = e - b;
            if (n == 0) continue;
            int g = per_xcd;  // tiles per item: the coarsest that still splits per XCD
            while (g > 1 && (uint64_t)n * (uint64_t)g / (uint64_t)T >= 2ull * kMinItemTokens) g /= 2;
            for (int t = 0; t < T; t += g) {
                const uint32_t d0 = tile_doc[t], d1 = tile_doc[std::min(T, t + g)];
                const uint32_t lo = (uint32_t)(std::lower_bound(tok_rank.begin() + b, tok_rank.begin() + e, d0) - tok_rank.begin());
                const uint32_t hi = (uint32_t)(std::lower_bound(tok_rank.begin() + b, tok_rank.begin() + e, d1) - tok_rank.begin());
                if (hi > lo) lists[t / per_xcd].push_back({(uint32_t)t, (uint32_t)w, lo, hi});
            }
        }
        std::vector<std::vector<uint4>> split(n_xcd);
        size_t longest = 0;
        for (int x = 0; x < n_xcd; ++x) {
            std::sort(lists[x].begin(), lists[x].end());
            for (auto &it : lists[x]) push_split(split[x], it[1], it[2], it[3]);
            longest = std::max(longest, split[x].size());
        }
        for (size_t j = 0; j < longest; ++j)
            for (int x = 0; x < n_xcd; ++x)
                out.push_back(j < split[x].size() ? split[x][j] : make_uint4(0xFFFFFFFFu, 0u, 0u, 0u));
    } else {
        for (int w = 0; w < VD; ++w) push_split(out, (uint32_t)w, word_off[w], word_off[w + 1]);
    }
    for (int w = VD; w < V; ++w) push_split(out, (uint32_t)w, word_off[w], word_off[w + 1]);
    if (tiles_used) *tiles_used = T;
    return out;
}

// Runs: a word's consecutive tokens in one document (tokens of a word are
// numbered in doc order), at most `cap` each, as (first token, first slot,
// doc, tokens); a run's slots are consecutive too (a doc's slots are grouped
// by word, then position).  Inside an item the runs are ordered longest first
// (then by token), so the lanes of a wavefront draw similar numbers of
// tokens; the order does not affect any draw.  Items are rewritten as run
// ranges (w, first run, end run, 0).
std::vector<uint4> build_runs(std::vector<uint4> &items, const std::vector<uint32_t> &tok_doc,
                              const std::vector<uint32_t> &tok_slot, uint32_t cap) {
    const int64_t n = (int64_t)items.size();
    std::vector<uint64_t> cnt(n + 1, 0);
    auto for_runs = [&](const uint4 &it, auto &&emit) {
        uint32_t t = it.y;
        while (t < it.z) {
            uint32_t e = t + 1;
            // same doc implies the next slot (the sampler writes z_doc[slot + j])
            while (e < it.z && e - t < cap && tok_doc[e] == tok_doc[t] && tok_slot[e] == tok_slot[t] + (e - t)) ++e;
            emit(t, e - t);
            t = e;
        }
    };
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            uint64_t k = 0;
            if (items[i].x != 0xFFFFFFFFu) for_runs(items[i], [&](uint32_t, uint32_t) { ++k; });
            cnt[i + 1] = k;
        }
    });
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint4> runs(cnt[n]);
    parallel_for(n, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            uint4 &it = items[i];
            if (it.x == 0xFFFFFFFFu) continue;
            uint64_t k = cnt[i];
            for_runs(it, [&](uint32_t t, uint32_t l) { runs[k++] = make_uint4(t, tok_slot[t], tok_doc[t], l); });
            std::stable_sort(runs.begin() + cnt[i], runs.begin() + cnt[i + 1],
                             [](const uint4 &x, const uint4 &y) { return x.w > y.w; });
            it = make_uint4(it.x, (uint32_t)cnt[i], (uint32_t)cnt[i + 1], 0u);
        }
    });
    return runs;
}

// Dense head words: the first dense_words ranks, else the ranks with at
// least dense_threshold (default K) tokens, and every word with 2^16 or more
// tokens (tail counts are 16-bit).
int head_words(const ezlda_ctx *c, int V, int K) {
    auto count = [&](int r) { return (int64_t)(c->word_off[r + 1] - c->word_off[r]); };
    int VD;
    if (c->a.dense_words >= 0) {
        VD = std::min(c->a.dense_words, V);
    } else {
        const int64_t thr = c->a.dense_threshold > 0 ? c->a.dense_threshold : K;
        VD = 0;
        while (VD < V && count(VD) >= thr) ++VD;
    }
    while (VD < V && count(VD) >= 65536) ++VD;
    return VD;
}

// The part of the device layout that the tuning arguments decide (block_dim,
// grid_dim, token_segment, dense_words / dense_threshold, doc_tiles,
// run_cap): the head/tail split of WT, the WT head segments, the sampler
// items and runs.  Token order, DT rows and z do not depend on it, so it can
// be rebuilt on a loaded corpus; counts are stale afterwards.
int build_layout(ezlda_ctx *c, const std::vector<uint32_t> &tok_doc, const std::vector<uint32_t> &tok_slot,
                 const std::vector<uint32_t> &slot_wrd) {
    ez::DevState &s = c->s;
    const int V = s.V, K = s.K, D = s.D;
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    for (void *p : c->layout_allocs) (void)hipFree(p);
    c->layout_allocs.clear();
    c->layout_bytes = 0;
    // head / tail split (words by rank: count desc)
    auto count = [&](int r) { return (int64_t)(c->word_off[r + 1] - c->word_off[r]); };
    const int VD = head_words(c, V, K);
    std::vector<uint32_t> tail_off(V - VD + 1, 0);
    for (int r = VD; r < V; ++r)
        tail_off[r - VD + 1] = tail_off[r - VD] + std::min<uint32_t>((uint32_t)count(r), (uint32_t)K);
    // work items
    const uint32_t WT_SEG = 32768;
    std::vector<uint4> wt_segs;
    for (int r = 0; r < VD; ++r)
        for (uint32_t b = c->word_off[r]; b < c->word_off[r + 1]; b += WT_SEG)
            wt_segs.push_back(make_uint4((uint32_t)r, b, std::min(b + WT_SEG, c->word_off[r + 1]), 0u));
    (void)D;
    std::vector<uint4> smp_segs =
        build_sample_items(c->a, c->word_off, tok_doc, c->soff, slot_wrd, VD, K, c->n_xcd, &c->doc_tiles);
    std::vector<uint4> runs = build_runs(smp_segs, tok_doc, tok_slot, (uint32_t)c->a.run_cap);
    for (uint4 &rn : runs) {  // doc q -> its DT row offset; the doc's length beside the run's
        rn.w |= (c->soff[rn.z + 1] - c->soff[rn.z]) << 16;
        rn.z = c->row_off[rn.z];
    }
    int r = 0;
    if ((r = ldalloc(c, &s.wt_head, (size_t)VD * K))) return r;
    if ((r = lupload(c, &s.tail_off, tail_off))) return r;
    if ((r = ldalloc(c, &s.wt_tail, tail_off.back()))) return r;
    if ((r = ldalloc(c, &s.tail_nnz, V - VD))) return r;
    if ((r = lupload(c, &s.wt_segs, wt_segs))) return r;
    if ((r = lupload(c, &s.smp_segs, smp_segs))) return r;
    if ((r = lupload(c, &s.runs, runs))) return r;
    if ((r = ldalloc(c, &s.ll_part, smp_segs.size()))) return r;
    s.VD = VD;
    s.n_wt_segs = (int32_t)wt_segs.size();
    s.n_smp_segs = (int32_t)smp_segs.size();
    s.n_runs = (int64_t)runs.size();
    s.n_queues = c->doc_tiles > 0 ? std::min(c->n_xcd, 8) : 1;
    s.block_dim = c->a.block_dim;
    s.grid_dim = c->a.grid_dim;
    s.n_cu = device_cus();
    s.lds_cu = device_lds_per_cu();
    c->tail_cap = tail_off.back();
    c->counts_fresh = false;
    // the sampler keeps z_doc only for head words: bring every slot up to date
    if (s.z_doc) {
        EZ_TRY_HIP(c, ez::launch_z_to_doc(s, c->st));
        EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    }
    return EZLDA_OK;
}

// ------------------------------------------------------------ out-of-core
// num_chunks > 1 (mainKernel.cu:133-310 streams chunks through two CUDA
// streams): every word keeps a dense count row, so counts accumulate over
// chunks with the same kernels; each chunk has its own word-major layout
// (the corpus order restricted to its documents), and its items carry the
// offset from its token ids to the corpus ones, so every draw is keyed as in
// the resident layout and the topics are the same.

bool is_chunked(const ezlda_ctx *c) { return !c->chunks.empty(); }

int refuse_chunked(ezlda_ctx *c, const char *what) {
    return fail(c, EZLDA_ESTATE, std::string(what) + " is not available when the corpus is chunked (num_chunks > 1)");
}

size_t al16(size_t x) { return (x + 15u) & ~(size_t)15u; }

int load_chunked(ezlda_ctx *c, const std::vector<uint32_t> &tok_doc, const std::vector<uint32_t> &tok_slot,
                 const std::vector<uint32_t> &slot_wrd, const std::vector<uint16_t> &z,
                 const std::vector<uint16_t> &z_doc) {
    ez::DevState &s = c->s;
    const int V = s.V, K = s.K, D = s.D;
    const uint32_t N = s.N;
    const int C = std::min<int>(c->a.num_chunks, D);
    // boundaries in doc order: about N / C tokens each, at least one doc each
    std::vector<uint32_t> qb(C + 1, 0);
    for (int k = 1; k < C; ++k) {
        const uint32_t q = (uint32_t)(std::lower_bound(c->soff.begin(), c->soff.end(),
                                                       (uint32_t)((uint64_t)N * (uint64_t)k / (uint64_t)C)) -
                                      c->soff.begin());
        qb[k] = std::min<uint32_t>(std::max<uint32_t>(q, qb[k - 1] + 1), (uint32_t)(D - (C - k)));
    }
    qb[C] = (uint32_t)D;
    std::vector<int32_t> chunk_of_q(D);
    for (int k = 0; k < C; ++k)
        for (uint32_t q = qb[k]; q < qb[k + 1]; ++q) chunk_of_q[q] = k;
    // chunk token ids: the corpus token order restricted to the chunk
    std::vector<uint32_t> loc(N), cnt(C, 0);
    for (uint32_t t = 0; t < N; ++t) loc[t] = cnt[chunk_of_q[tok_doc[t]]]++;
    const int VDi = head_words(c, V, K);  // item tiling as in the resident layout
    const bool compact = s.compact;
    c->chunks.assign(C, EzChunk{});
    c->chunk_slot0.assign(C, 0);
    size_t max_bytes = 0;
    uint64_t max_dt = 0, dt_total = 0;
    uint32_t max_n = 1, max_items = 1;
    int64_t items_total = 0, runs_total = 0;
    for (int k = 0; k < C; ++k) {
        EzChunk &ch = c->chunks[k];
        ch.q0 = qb[k];
        ch.D = qb[k + 1] - qb[k];
        ch.slot0 = c->soff[ch.q0];
        ch.N = c->soff[qb[k + 1]] - ch.slot0;
        c->chunk_slot0[k] = ch.slot0;
        const uint32_t n = ch.N;
        std::vector<uint32_t> woff(V + 1, 0), glob(n), tdoc(n), tslot(n), soffc(ch.D + 1), swrd(n), stok(n);
        std::vector<uint16_t> zc(n), zd(n);
        for (uint32_t pp = 0; pp < n; ++pp) {
            const uint32_t p = ch.slot0 + pp, t = c->slot_tok[p], l = loc[t];
            glob[l] = t;
            tdoc[l] = tok_doc[t] - ch.q0;
            tslot[l] = tok_slot[t] - ch.slot0;
            zc[l] = z[t];
            zd[pp] = z_doc[p];
            swrd[pp] = slot_wrd[p];
            stok[pp] = l;
            woff[slot_wrd[p] + 1]++;
        }
        for (int r = 0; r < V; ++r) woff[r + 1] += woff[r];
        for (uint32_t q = 0; q <= ch.D; ++q) soffc[q] = c->soff[ch.q0 + q] - ch.slot0;
        ch.D_long = 0;
        while ((uint32_t)ch.D_long < ch.D && soffc[ch.D_long + 1] - soffc[ch.D_long] > ez::kDtShort) ++ch.D_long;
        std::vector<uint2> di(ch.D);
        uint64_t dtc = 0;
        for (uint32_t q = 0; q < ch.D; ++q) {
            const uint32_t nq = soffc[q + 1] - soffc[q];
            di[q] = make_uint2((uint32_t)dtc, nq << 16);
            dtc += compact ? (dt_cap(nq, K) + 63u) & ~63u : (dt_cap(nq, K) + 31u) & ~31u;
        }
        if (dtc >= 0xFFFFFFF0ull) return fail(c, EZLDA_EINVAL, "DT capacity of a chunk exceeds 2^32 entries");
        ch.dt_entries = dtc;
        int tiles = 0;
        std::vector<uint4> items = build_sample_items(c->a, woff, tdoc, soffc, swrd, VDi, K, c->n_xcd, &tiles);
        std::vector<uint4> runs = build_runs(items, tdoc, tslot, (uint32_t)c->a.run_cap);
        for (uint4 &rn : runs) {  // doc q -> its DT row offset in the chunk; the doc's length
            rn.w |= (soffc[rn.z + 1] - soffc[rn.z]) << 16;
            rn.z = di[rn.z].x;
        }
        for (uint4 &it : items)
            if (it.x != 0xFFFFFFFFu) it.w = glob[woff[it.x]] - woff[it.x];
        ch.n_queues = tiles > 0 ? (uint32_t)std::min(c->n_xcd, 8) : 1u;
        std::vector<uint4> wsegs;
        for (int r = 0; r < V; ++r)
            for (uint32_t b = woff[r]; b < woff[r + 1]; b += 32768u)
                wsegs.push_back(make_uint4((uint32_t)r, b, std::min(b + 32768u, woff[r + 1]), 0u));
        // image
        size_t o = 4 * (size_t)n;  // z, z_doc
        o = al16(o);
        ch.off_doc_info = o;
        o += sizeof(uint2) * ch.D;
        ch.back_bytes = o;
        o = al16(o);
        ch.off_soff = o;
        o += 4 * ((size_t)ch.D + 1);
        o = al16(o);
        ch.off_slot_wrd = o;
        o += 4 * (size_t)n;
        o = al16(o);
        ch.off_slot_tok = o;
        o += 4 * (size_t)n;
        o = al16(o);
        ch.off_items = o;
        o += sizeof(uint4) * items.size();
        ch.off_runs = o;
        o += sizeof(uint4) * runs.size();
        ch.off_wt_segs = o;
        o += sizeof(uint4) * wsegs.size();
        ch.bytes = al16(o);
        EZ_TRY_HIP(c, hipHostMalloc((void **)&ch.img, ch.bytes, hipHostMallocDefault));
        std::memset(ch.img, 0, ch.bytes);
        auto put = [&](size_t off, const void *src, size_t bytes) {
            if (bytes) std::memcpy(ch.img + off, src, bytes);
        };
        put(0, zc.data(), 2 * (size_t)n);
        put(2 * (size_t)n, zd.data(), 2 * (size_t)n);
        put(ch.off_doc_info, di.data(), sizeof(uint2) * ch.D);
        put(ch.off_soff, soffc.data(), 4 * ((size_t)ch.D + 1));
        put(ch.off_slot_wrd, swrd.data(), 4 * (size_t)n);
        put(ch.off_slot_tok, stok.data(), 4 * (size_t)n);
        put(ch.off_items, items.data(), sizeof(uint4) * items.size());
        put(ch.off_runs, runs.data(), sizeof(uint4) * runs.size());
        put(ch.off_wt_segs, wsegs.data(), sizeof(uint4) * wsegs.size());
        ch.n_items = (uint32_t)items.size();
        ch.n_runs = runs.size();
        ch.n_wt_segs = (uint32_t)wsegs.size();
        max_bytes = std::max(max_bytes, ch.bytes);
        max_dt = std::max(max_dt, dtc);
        max_n = std::max(max_n, n);
        max_items = std::max(max_items, ch.n_items);
        dt_total += dtc;
        items_total += ch.n_items;
        runs_total += (int64_t)ch.n_runs;
    }
    c->tok_slot_g = tok_slot;
    // device: two image buffers, the model and the next sweep's counts
    int r = 0;
    const size_t dt_words = compact ? (max_dt + 256 + 1) / 2 : max_dt + 256;
    for (int j = 0; j < 2; ++j) {
        if ((r = dalloc(c, &c->cimg[j], max_bytes))) return r;
        if ((r = dalloc(c, &c->cdt[j], dt_words))) return r;
        EZ_TRY_HIP(c, hipMemset(c->cdt[j], 0, dt_words * sizeof(uint32_t)));
        if ((r = dalloc(c, &c->caux[j], max_n))) return r;
        if ((r = dalloc(c, &c->cctr[j], ez::kCounterWords))) return r;
        EZ_TRY_HIP(c, hipMemset(c->cctr[j], 0, ez::kCounterWords * sizeof(uint32_t)));
        if ((r = dalloc(c, &c->cll[j], max_items))) return r;
        if ((r = dalloc(c, &c->cllo[j], 1))) return r;
    }
    if ((r = dalloc(c, &s.wt_head, (size_t)V * K))) return r;
    if ((r = dalloc(c, &c->wt_next, (size_t)V * K))) return r;
    if ((r = dalloc(c, &s.nk, K))) return r;
    if ((r = dalloc(c, &c->nk_next, K))) return r;
    if ((r = dalloc(c, &s.inv, K))) return r;
    if ((r = dalloc(c, &s.wtop, V))) return r;
    if ((r = dalloc(c, &s.zero_order, K))) return r;
    if ((r = dalloc(c, &s.tail_off, 1))) return r;
    EZ_TRY_HIP(c, hipMemset(s.tail_off, 0, sizeof(uint32_t)));
    if ((r = dalloc(c, &s.counters, ez::kCounterWords))) return r;
    if ((r = dalloc(c, &s.stats, 2))) return r;
    if ((r = dalloc(c, &s.ll_out, 1))) return r;
    EZ_TRY_HIP(c, hipMemset(s.counters, 0, ez::kCounterWords * sizeof(uint32_t)));
    EZ_TRY_HIP(c, hipMemset(s.stats, 0, 2 * sizeof(unsigned long long)));
    if (!c->st2) EZ_TRY_HIP(c, hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
    s.VD = V;
    s.n_smp_segs = (int32_t)std::min<int64_t>(items_total, INT32_MAX);
    s.n_runs = runs_total;
    s.block_dim = c->a.block_dim;
    s.grid_dim = c->a.grid_dim;
    s.n_cu = device_cus();
    s.lds_cu = device_lds_per_cu();
    c->dt_cap = (int64_t)dt_total;
    c->tail_cap = 0;
    c->doc_tiles = 0;
    EZ_TRY_HIP(c, hipDeviceSynchronize());
    return EZLDA_OK;
}

// Device view of chunk k in image buffer j: the chunk's arrays, the shared
// model (every word dense), the buffer's own sweep counter and item queues.
ez::DevState chunk_state(const ezlda_ctx *c, int k, int j) {
    const EzChunk &ch = c->chunks[k];
    ez::DevState cs = c->s;
    uint8_t *img = c->cimg[j];
    cs.D = (int32_t)ch.D;
    cs.D_long = ch.D_long;
    cs.N = ch.N;
    cs.VD = c->s.V;
    cs.word_off = nullptr;
    cs.z = reinterpret_cast<uint16_t *>(img);
    cs.z_doc = reinterpret_cast<uint16_t *>(img + 2 * (size_t)ch.N);
    cs.doc_info = reinterpret_cast<uint2 *>(img + ch.off_doc_info);
    cs.soff = reinterpret_cast<uint32_t *>(img + ch.off_soff);
    cs.slot_wrd = reinterpret_cast<uint32_t *>(img + ch.off_slot_wrd);
    cs.slot_tok = reinterpret_cast<uint32_t *>(img + ch.off_slot_tok);
    cs.aux_doc = c->caux[j];
    cs.dt = c->cdt[j];
    cs.smp_segs = reinterpret_cast<uint4 *>(img + ch.off_items);
    cs.n_smp_segs = (int32_t)ch.n_items;
    cs.runs = reinterpret_cast<uint4 *>(img + ch.off_runs);
    cs.n_runs = (int64_t)ch.n_runs;
    cs.wt_segs = reinterpret_cast<uint4 *>(img + ch.off_wt_segs);
    cs.n_wt_segs = (int32_t)ch.n_wt_segs;
    cs.n_queues = (int32_t)ch.n_queues;
    cs.counters = c->cctr[j];
    cs.ll_part = c->cll[j];
    cs.ll_out = c->cllo[j];
    return cs;
}

// Counts of the current topics: every chunk's histogram added up.
int chunk_prime(ezlda_ctx *c) {
    EZ_TRY_HIP(c, ez::launch_zero_counts(c->s, c->st));
    for (int k = 0; k < (int)c->chunks.size(); ++k) {
        const EzChunk &ch = c->chunks[k];
        EZ_TRY_HIP(c, hipMemcpyAsync(c->cimg[0], ch.img, ch.bytes, hipMemcpyHostToDevice, c->st));
        EZ_TRY_HIP(c, ez::launch_accumulate_wt(chunk_state(c, k, 0), c->st));
    }
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = true;
    return EZLDA_OK;
}

// n sweeps over the chunks; chunk k goes through image buffer k % 2 on its
// own stream, so one chunk's copies overlap the other's kernels.  Each chunk
// rebuilds its DT rows, samples against the sweep's model, adds its new
// topics to the next counts and copies z, z_doc and its row lengths back.
int chunk_sweeps(ezlda_ctx *c, int n, hipEvent_t ev_begin, hipEvent_t ev_end) {
    if (!c->counts_fresh)
        if (int r = chunk_prime(c)) return r;
    hipStream_t sts[2] = {c->st, c->st2};
    hipEvent_t e_start = nullptr, e_done = nullptr;
    int rc = EZLDA_OK;
    if (hipEventCreateWithFlags(&e_start, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e_done, hipEventDisableTiming) != hipSuccess)
        rc = fail(c, EZLDA_EHIP, "hipEventCreate failed");
    auto chk = [&](hipError_t e, const char *what) {
        if (rc == EZLDA_OK && e != hipSuccess) rc = fail(c, EZLDA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    if (rc == EZLDA_OK && ev_begin) chk(hipEventRecord(ev_begin, c->st), "hipEventRecord");
    const size_t vk = (size_t)c->s.V * c->s.K;
    for (int i = 0; i < n && rc == EZLDA_OK; ++i) {
        chk(ez::launch_topk(c->s, c->st), "top-3");
        ez::DevState nx = c->s;  // zero the next counts
        nx.wt_head = c->wt_next;
        nx.nk = c->nk_next;
        nx.VD = c->s.V;
        chk(ez::launch_zero_counts(nx, c->st), "zero counts");
        chk(hipEventRecord(e_start, c->st), "hipEventRecord");
        chk(hipStreamWaitEvent(c->st2, e_start, 0), "hipStreamWaitEvent");
        for (int k = 0; k < (int)c->chunks.size() && rc == EZLDA_OK; ++k) {
            const int j = k & 1;
            hipStream_t st = sts[j];
            const EzChunk &ch = c->chunks[k];
            const ez::DevState cs = chunk_state(c, k, j);
            chk(hipMemcpyAsync(c->cimg[j], ch.img, ch.bytes, hipMemcpyHostToDevice, st), "chunk upload");
            chk(ez::launch_set_u32(cs.counters, c->sweep, st), "sweep counter");
            chk(ez::launch_build_dt(cs, st), "DT rebuild");
            chk(ez::launch_sample(cs, c->a.shortcut != 0, st), "sample");
            ez::DevState acc = cs;
            acc.wt_head = c->wt_next;
            acc.nk = c->nk_next;
            chk(ez::launch_accumulate_wt(acc, st), "count accumulation");
            chk(hipMemcpyAsync(ch.img, c->cimg[j], ch.back_bytes, hipMemcpyDeviceToHost, st), "chunk download");
        }
        chk(hipEventRecord(e_done, c->st2), "hipEventRecord");
        chk(hipStreamWaitEvent(c->st, e_done, 0), "hipStreamWaitEvent");
        std::swap(c->s.wt_head, c->wt_next);
        std::swap(c->s.nk, c->nk_next);
        c->sweep++;
        (void)vk;
    }
    if (rc == EZLDA_OK && ev_end) chk(hipEventRecord(ev_end, c->st), "hipEventRecord");
    chk(hipStreamSynchronize(c->st), "chunked sweep");
    if (e_start) (void)hipEventDestroy(e_start);
    if (e_done) (void)hipEventDestroy(e_done);
    c->counts_fresh = rc == EZLDA_OK;
    return rc;
}

// Chunk holding corpus slot p.
int chunk_of_slot(const ezlda_ctx *c, uint32_t p) {
    return (int)(std::upper_bound(c->chunk_slot0.begin(), c->chunk_slot0.end(), p) - c->chunk_slot0.begin()) - 1;
}

}  // namespace

int ezc::refresh_counts(ezlda_ctx *c) {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) {
        if (!c->counts_fresh)
            if (int r = chunk_prime(c)) return r;
    } else if (!c->counts_fresh) {
        EZ_TRY_HIP(c, ez::launch_build_wt(c->s, c->st));
    }
    EZ_TRY_HIP(c, ez::launch_topk(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = true;
    return EZLDA_OK;
}

extern "C" {

void ezlda_default_args(ezlda_args *a) {
    std::memset(a, 0, sizeof(*a));
    a->K = 1024;             // Argument.cuh:6
    a->alpha = 50.0f / 1024; // Argument.cuh:9
    a->beta = 0.01f;         // Argument.cuh:8
    a->block_dim = 0;        // auto (Argument.cuh:4 has 256; 64 or 128 measure faster here)
    a->grid_dim = -1;        // auto: 16 or 8 persistent workgroups per CU
    a->num_chunks = 1;
    a->num_streams = 1;
    a->token_segment = 4096;
    a->seed = 2002;
    a->dense_words = -1;
    a->dense_threshold = 0;  // K, as preprocess/streamTestBalance.cpp:291
    a->shortcut = 1;
    a->topic_bits = 16;
    a->run_cap = 4;
}

const char *ezlda_last_error(const ezlda_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

ezlda_ctx *ezlda_create(const ezlda_args *args) try {
    if (!args) {
        fail(nullptr, EZLDA_EINVAL, "null args");
        return nullptr;
    }
    ezlda_args a = *args;
    if (a.K < 1 || a.K > 16383) {
        fail(nullptr, EZLDA_EINVAL, "K must be in [1, 16383]");
        return nullptr;
    }
    if (a.alpha <= 0.f) a.alpha = 50.0f / (float)a.K;
    if (a.beta <= 0.f) a.beta = 0.01f;
    if (a.topic_bits == 0) a.topic_bits = 16;
    if (a.topic_bits != 16) {
        fail(nullptr, EZLDA_EINVAL, "topic_bits must be 16: topics are stored in 16 bits (K <= 16383)");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        fail(nullptr, EZLDA_EHIP, "no HIP device");
        return nullptr;
    }
    {
        const std::string m = check_layout_args(a);
        if (!m.empty()) {
            fail(nullptr, EZLDA_EINVAL, m);
            return nullptr;
        }
    }
    {
        int dev = 0, lmax = 0;
        const size_t lds = ez::max_lds_bytes(a.K, a.block_dim);
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
            fail(nullptr, EZLDA_EHIP, "cannot query the device's LDS size");
            return nullptr;
        }
        if (ez::sampler_static_lds() != 0) {
            fail(nullptr, EZLDA_EHIP, "the sampler kernel has static LDS; its phi table must start at LDS address 0");
            return nullptr;
        }
        if (lds > (size_t)lmax) {
            fail(nullptr, EZLDA_EINVAL,
                 "K=" + std::to_string(a.K) + " needs " + std::to_string(lds) + " bytes of LDS per block; the device "
                 "offers " + std::to_string(lmax));
            return nullptr;
        }
    }
    ezlda_ctx *c = new ezlda_ctx();
    c->a = a;
    {
        int dev = 0, nx = 1;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, dev) == hipSuccess && nx > 0)
            c->n_xcd = nx;
    }
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        fail(nullptr, EZLDA_EHIP, "hipStreamCreate failed");
        delete c;
        return nullptr;
    }
    // a failed lookup only disables ezlda_loglik_noalloc, which reports it
    scratch_for_current_device(&c->ll_scratch, &c->scratch_err);
    return c;
} catch (...) {
    (void)ezc::exception_status(nullptr);
    return nullptr;
}

void ezlda_destroy(ezlda_ctx *c) {
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    free_device(c);
    if (c->st) (void)hipStreamDestroy(c->st);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    delete c;
}

int ezlda_load_docs(ezlda_ctx *c, int32_t V, int32_t D, const int64_t *doc_off, const int32_t *words,
                    const uint16_t *z0) try {
    if (!c) return fail(nullptr, EZLDA_EINVAL, "null context");
    if (V < 1 || D < 0 || !doc_off || (!words && D > 0 && doc_off[D] > 0))
        return fail(c, EZLDA_EINVAL, "bad corpus arguments");
    const int K = c->a.K;
    const int64_t N64 = doc_off[D] - doc_off[0];
    if (doc_off[0] != 0) return fail(c, EZLDA_EINVAL, "doc_off[0] must be 0");
    if (N64 < 0 || N64 >= (int64_t)0xFFFFFFFFll) return fail(c, EZLDA_EINVAL, "token count must be below 2^32");
    for (int d = 0; d < D; ++d) {
        const int64_t n = doc_off[d + 1] - doc_off[d];
        if (n < 0) return fail(c, EZLDA_EINVAL, "doc_off must be non-decreasing");
        if (n > 65535) return fail(c, EZLDA_EINVAL, "a document has more than 65535 tokens (doc " + std::to_string(d) + ")");
    }
    const uint32_t N = (uint32_t)N64;
    // word counts and the rank order (count desc, id asc)
    std::vector<int64_t> cnt(V, 0);
    for (uint32_t i = 0; i < N; ++i) {
        const int32_t w = words[i];
        if (w < 0 || w >= V) return fail(c, EZLDA_EINVAL, "word id out of range at token " + std::to_string(i));
        cnt[w]++;
    }
    if (z0)
        for (uint32_t i = 0; i < N; ++i)
            if (z0[i] >= K) return fail(c, EZLDA_EINVAL, "initial topic out of range at token " + std::to_string(i));
    free_device(c);
    std::vector<int32_t> ids(V);
    for (int w = 0; w < V; ++w) ids[w] = w;
    std::sort(ids.begin(), ids.end(), [&](int32_t x, int32_t y) { return cnt[x] > cnt[y] || (cnt[x] == cnt[y] && x < y); });
    c->rank_of.assign(V, 0);
    for (int r = 0; r < V; ++r) c->rank_of[ids[r]] = r;
    c->word_off.assign(V + 1, 0);
    for (int r = 0; r < V; ++r) c->word_off[r + 1] = c->word_off[r] + (uint32_t)cnt[ids[r]];
    c->doc_off.assign(D + 1, 0);
    for (int d = 0; d <= D; ++d) c->doc_off[d] = (uint32_t)doc_off[d];

    // docs by length desc, id asc: the order tokens of a word are numbered in
    // (neighbouring tokens then have similar DT row lengths)
    std::vector<int32_t> dord(D);
    for (int d = 0; d < D; ++d) dord[d] = d;
    std::sort(dord.begin(), dord.end(), [&](int32_t x, int32_t y) {
        const int64_t lx = doc_off[x + 1] - doc_off[x], ly = doc_off[y + 1] - doc_off[y];
        return lx > ly || (lx == ly && x < y);
    });
    // doc slots: doc q = dord[q] owns [soff[q], soff[q+1])
    c->dord = dord;
    c->soff.assign(D + 1, 0);
    for (int q = 0; q < D; ++q) c->soff[q + 1] = c->soff[q] + (uint32_t)(doc_off[dord[q] + 1] - doc_off[dord[q]]);
    // placement: contiguous ranges of the doc order per thread keep the
    // (doc order, position) order of tokens inside each word
    const int T = host_threads();
    std::vector<int> dsplit(T + 1, 0);
    for (int t = 0; t <= T; ++t)
        dsplit[t] = (int)(std::lower_bound(c->soff.begin(), c->soff.end(), (uint32_t)(N64 * t / T)) - c->soff.begin());
    dsplit[0] = 0;
    dsplit[T] = D;
    std::vector<std::vector<uint32_t>> tcnt(T, std::vector<uint32_t>(V, 0));
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (int q = dsplit[t]; q < dsplit[t + 1]; ++q)
                    for (int64_t i = doc_off[dord[q]]; i < doc_off[dord[q] + 1]; ++i) tcnt[t][c->rank_of[words[i]]]++;
            });
        for (auto &x : th) x.join();
    }
    for (int r = 0; r < V; ++r) {
        uint32_t cur = c->word_off[r];
        for (int t = 0; t < T; ++t) {
            const uint32_t n = tcnt[t][r];
            tcnt[t][r] = cur;
            cur += n;
        }
    }
    // token ids and slots: inside a doc, slots follow (word rank, position);
    // the doc's tokens of one word take consecutive ids in position order
    std::vector<uint32_t> tok_doc(N), tok_slot(N), slot_wrd(N);
    std::vector<uint16_t> z(N);
    c->doc_tok.assign(N, 0);
    c->slot_tok.assign(N, 0);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                std::vector<uint32_t> &cur = tcnt[t];
                std::vector<uint64_t> key;
                for (int q = dsplit[t]; q < dsplit[t + 1]; ++q) {
                    const int d = dord[q];
                    const int64_t b = doc_off[d], n = doc_off[d + 1] - b;
                    key.resize((size_t)n);
                    for (int64_t i = 0; i < n; ++i) key[i] = ((uint64_t)c->rank_of[words[b + i]] << 16) | (uint64_t)i;
                    std::sort(key.begin(), key.end());
                    uint32_t p = c->soff[q];
                    for (int64_t j = 0; j < n; ++j, ++p) {
                        const uint32_t r = (uint32_t)(key[j] >> 16);
                        const int64_t i = b + (int64_t)(key[j] & 0xFFFFu);
                        const uint32_t tk = cur[r]++;
                        tok_doc[tk] = (uint32_t)q;
                        tok_slot[tk] = p;
                        c->slot_tok[p] = tk;
                        slot_wrd[p] = r;
                        c->doc_tok[i] = tk;
                        if (z0) z[tk] = z0[i];
                    }
                }
            });
        for (auto &x : th) x.join();
    }
    tcnt.clear();
    if (!z0) {
        const uint64_t seed = c->a.seed;
        parallel_for(N, [&](int64_t b, int64_t e, int) {
            for (int64_t t = b; t < e; ++t)
                z[t] = (uint16_t)(ez_philox_x0(seed, (uint64_t)t, 0u, EZ_TAG_INIT) % (uint32_t)K);
        });
    }
    std::vector<uint16_t> z_doc(N);
    parallel_for(N, [&](int64_t b, int64_t e, int) {
        for (int64_t p = b; p < e; ++p) z_doc[p] = z[c->slot_tok[p]];
    });
    if (c->a.num_chunks > 1 && D > 1) {  // out of core: chunk images in pinned host memory
        ez::DevState &s = c->s;
        s.compact = K <= (int)kCompactK;
        s.topic_bits = s.compact ? topic_bits(K) : 16u;
        s.K = K; s.V = V; s.D = D; s.N = N;
        s.alpha = c->a.alpha; s.beta = c->a.beta;
        s.vbeta = (float)((double)V * (double)c->a.beta);
        s.seed = c->a.seed;
        c->dt_cap_canon = 0;
        for (int d = 0; d < D; ++d) c->dt_cap_canon += dt_cap(c->doc_off[d + 1] - c->doc_off[d], K);
        if (int r = load_chunked(c, tok_doc, tok_slot, slot_wrd, z, z_doc)) return r;
        c->sweep = 0;
        c->counts_fresh = false;
        c->loaded = true;
        return EZLDA_OK;
    }
    // DT rows in doc order q, each 128-byte aligned (the build zero-fills
    // the rest): K <= 8192 keeps 16-bit entries with counts in pieces
    // (capacity dt_cap(n, K) rounded up to 64 entries), larger K 32-bit
    // entries (min(n_d, K) rounded up to 32)
    const bool compact = K <= (int)kCompactK;
    std::vector<uint2> doc_info(D);
    uint64_t dtc = 0;
    for (int q = 0; q < D; ++q) {
        const uint32_t n = c->soff[q + 1] - c->soff[q];
        doc_info[q] = make_uint2((uint32_t)dtc, n << 16);
        dtc += compact ? (dt_cap(n, K) + 63u) & ~63u : (dt_cap(n, K) + 31u) & ~31u;
    }
    if (dtc >= 0xFFFFFFF0ull) return fail(c, EZLDA_EINVAL, "DT capacity exceeds 2^32 entries");
    c->row_off.resize(D);
    for (int q = 0; q < D; ++q) c->row_off[q] = doc_info[q].x;
    ez::DevState &s = c->s;
    s.compact = compact;
    s.topic_bits = compact ? topic_bits(K) : 16u;
    s.K = K; s.V = V; s.D = D; s.N = N;
    s.alpha = c->a.alpha; s.beta = c->a.beta;
    s.vbeta = (float)((double)V * (double)c->a.beta);
    s.seed = c->a.seed;
    s.D_long = 0;  // docs come longest first
    while (s.D_long < D && c->soff[s.D_long + 1] - c->soff[s.D_long] > ez::kDtShort) ++s.D_long;
    c->dt_cap = (int64_t)dtc;
    c->dt_cap_canon = 0;
    for (int d = 0; d < D; ++d) c->dt_cap_canon += dt_cap(c->doc_off[d + 1] - c->doc_off[d], K);
    int r = 0;
    if ((r = upload(c, &s.word_off, c->word_off))) return r;
    if ((r = upload(c, &s.z, z))) return r;
    if ((r = upload(c, &s.soff, c->soff))) return r;
    if ((r = upload(c, &s.slot_wrd, slot_wrd))) return r;
    if ((r = upload(c, &s.slot_tok, c->slot_tok))) return r;
    if ((r = upload(c, &s.z_doc, z_doc))) return r;
    if ((r = dalloc(c, &s.aux_doc, N))) return r;
    if ((r = upload(c, &s.doc_info, doc_info))) return r;
    // 256 entries of zero padding past the last row
    const size_t dt_words = compact ? (dtc + 256 + 1) / 2 : dtc + 256;
    if ((r = dalloc(c, &s.dt, dt_words))) return r;
    EZ_TRY_HIP(c, hipMemset(s.dt, 0, dt_words * sizeof(uint32_t)));
    if ((r = dalloc(c, &s.nk, K))) return r;
    if ((r = dalloc(c, &s.inv, K))) return r;
    if ((r = dalloc(c, &s.wtop, V))) return r;
    if ((r = dalloc(c, &s.zero_order, K))) return r;
    if ((r = dalloc(c, &s.counters, ez::kCounterWords))) return r;
    if ((r = dalloc(c, &s.stats, 2))) return r;
    if ((r = dalloc(c, &s.ll_out, 1))) return r;
    if ((r = build_layout(c, tok_doc, tok_slot, slot_wrd))) return r;
    EZ_TRY_HIP(c, hipMemsetAsync(s.counters, 0, ez::kCounterWords * sizeof(uint32_t), c->st));
    EZ_TRY_HIP(c, hipMemsetAsync(s.stats, 0, 2 * sizeof(unsigned long long), c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    EZ_TRY_HIP(c, hipDeviceSynchronize());
    c->sweep = 0;
    c->counts_fresh = false;
    c->loaded = true;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_load_chunked(ezlda_ctx *c, const char *prefix, int32_t num_chunks) try {
    if (!c) return fail(nullptr, EZLDA_EINVAL, "null context");
    ezlda_corpus *cp = ezlda_read_chunked(prefix, num_chunks);
    if (!cp) return fail(c, EZLDA_EIO, g_err);
    const int r = ezlda_load_docs(c, ezlda_corpus_num_words(cp), ezlda_corpus_num_docs(cp), ezlda_corpus_doc_off(cp),
                                  ezlda_corpus_words(cp), ezlda_corpus_topics(cp));
    ezlda_corpus_free(cp);
    return r;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_info(ezlda_ctx *c, ezlda_info *info) try {
    if (int r = need_loaded(c)) return r;
    info->num_tokens = c->s.N;
    info->num_words = c->s.V;
    info->num_docs = c->s.D;
    info->num_head_words = c->s.VD;
    info->K = c->s.K;
    info->dt_capacity = c->dt_cap_canon;
    info->wt_tail_capacity = c->tail_cap;
    info->num_sample_items = c->s.n_smp_segs;
    info->sweep = c->sweep;
    info->device_bytes = c->dev_bytes + c->layout_bytes;
    info->doc_tiles = c->doc_tiles;
    info->num_xcd = c->n_xcd;
    info->run_cap = c->a.run_cap;
    info->grid_dim = c->a.grid_dim;
    info->block_dim = c->a.block_dim;
    info->token_segment = c->a.token_segment;
    info->num_runs = c->s.n_runs;
    info->num_chunks = is_chunked(c) ? (int32_t)c->chunks.size() : 1;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_build_wt(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_build_wt");
    EZ_TRY_HIP(c, ez::launch_build_wt(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = true;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_topk(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_topk");
    EZ_TRY_HIP(c, ez::launch_topk(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_build_dt(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_build_dt");
    EZ_TRY_HIP(c, ez::launch_build_dt(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_sample(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_sample");
    EZ_TRY_HIP(c, ez::launch_sample(c->s, c->a.shortcut != 0, c->st));
    EZ_TRY_HIP(c, ez::launch_advance_sweep(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->sweep++;
    c->counts_fresh = false;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

// The reference's launcher-sized stages of the sampler and the counts
// (SamplingKernel.cuh:10-15, WTDenUpdateKernel.cuh:25-26).  They run the
// fused kernels on a part of the work, so any order of shortcut, sample_head
// and sample_tail followed by commit_head gives ezlda_sample's topics.
static int sample_stage(ezlda_ctx *c, uint32_t mode, const char *what) {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, what);
    EZ_TRY_HIP(c, ez::launch_sample(c->s, c->a.shortcut != 0, c->st, mode));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = false;
    return EZLDA_OK;
}

int ezlda_shortcut(ezlda_ctx *c) { return sample_stage(c, ez::kSampleHead | ez::kShortcutOnly, "ezlda_shortcut"); }

int ezlda_sample_head(ezlda_ctx *c) { return sample_stage(c, ez::kSampleHead, "ezlda_sample_head"); }

int ezlda_sample_tail(ezlda_ctx *c) { return sample_stage(c, ez::kSampleTail, "ezlda_sample_tail"); }

int ezlda_commit_head(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_commit_head");
    EZ_TRY_HIP(c, ez::launch_head_rows(c->s, c->st));
    EZ_TRY_HIP(c, ez::launch_advance_sweep(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->sweep++;
    c->counts_fresh = false;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_wt_rowsum(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_wt_rowsum");
    EZ_TRY_HIP(c, ez::launch_topic_hist(c->s, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

// One sweep in stage order; ev (nullable) receives EZLDA_NUM_STAGES + 1
// events that bracket the stages.
static int enqueue_sweep(ezlda_ctx *c, hipEvent_t *ev) {
    if (ev) EZ_TRY_HIP(c, hipEventRecord(ev[0], c->st));
    EZ_TRY_HIP(c, ez::launch_build_wt(c->s, c->st));
    if (ev) EZ_TRY_HIP(c, hipEventRecord(ev[1], c->st));
    EZ_TRY_HIP(c, ez::launch_topk(c->s, c->st));
    if (ev) EZ_TRY_HIP(c, hipEventRecord(ev[2], c->st));
    EZ_TRY_HIP(c, ez::launch_build_dt(c->s, c->st));
    if (ev) EZ_TRY_HIP(c, hipEventRecord(ev[3], c->st));
    EZ_TRY_HIP(c, ez::launch_sample(c->s, c->a.shortcut != 0, c->st));
    if (ev) EZ_TRY_HIP(c, hipEventRecord(ev[4], c->st));
    EZ_TRY_HIP(c, ez::launch_advance_sweep(c->s, c->st));
    return EZLDA_OK;
}

int ezlda_sweep(ezlda_ctx *c, int32_t n) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return chunk_sweeps(c, n, nullptr, nullptr);
    for (int i = 0; i < n; ++i) {
        if (int r = enqueue_sweep(c, nullptr)) return r;
        c->sweep++;
    }
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = false;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

static int ensure_graph(ezlda_ctx *c) {
    if (c->graph_exec) return EZLDA_OK;
    EZ_TRY_HIP(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int r = enqueue_sweep(c, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->st, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(c, EZLDA_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    c->graph = g;
    EZ_TRY_HIP(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    return EZLDA_OK;
}

int ezlda_sweep_graph(ezlda_ctx *c, int32_t n) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_sweep_graph");
    if (int r = ensure_graph(c)) return r;
    for (int i = 0; i < n; ++i) {
        EZ_TRY_HIP(c, hipGraphLaunch(c->graph_exec, c->st));
        c->sweep++;
    }
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->counts_fresh = false;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

// Back-to-back sweeps with no host synchronisation inside the timed region.
// Eager: every stage is bracketed by events, all read after the last sweep.
// Graph: one event pair around n replays (stage times are then -1).
namespace {

// The per-token inputs of the layout, from the host mirrors: doc q and slot
// of every token, word rank of every slot.
void layout_inputs(ezlda_ctx *c, std::vector<uint32_t> &tok_doc, std::vector<uint32_t> &tok_slot,
                   std::vector<uint32_t> &slot_wrd) {
    const uint32_t N = c->s.N;
    const int D = c->s.D, V = c->s.V;
    tok_doc.resize(N);
    tok_slot.resize(N);
    slot_wrd.resize(N);
    parallel_for(D, [&](int64_t b, int64_t e, int) {
        for (int64_t q = b; q < e; ++q)
            for (uint32_t p = c->soff[q]; p < c->soff[q + 1]; ++p) {
                tok_doc[c->slot_tok[p]] = (uint32_t)q;
                tok_slot[c->slot_tok[p]] = p;
            }
    });
    std::vector<uint32_t> tok_wrd(N);
    for (int r = 0; r < V; ++r)
        for (uint32_t t = c->word_off[r]; t < c->word_off[r + 1]; ++t) tok_wrd[t] = (uint32_t)r;
    parallel_for(N, [&](int64_t b, int64_t e, int) {
        for (int64_t p = b; p < e; ++p) slot_wrd[p] = tok_wrd[c->slot_tok[p]];
    });
}

}  // namespace

int ezlda_set_layout(ezlda_ctx *c, const ezlda_args *args) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_set_layout");
    if (!args) return fail(c, EZLDA_EINVAL, "null args");
    ezlda_args a = c->a;
    a.block_dim = args->block_dim;
    a.grid_dim = args->grid_dim;
    a.token_segment = args->token_segment;
    a.dense_words = args->dense_words;
    a.dense_threshold = args->dense_threshold;
    a.doc_tiles = args->doc_tiles;
    a.run_cap = args->run_cap;
    const std::string m = check_layout_args(a);
    if (!m.empty()) return fail(c, EZLDA_EINVAL, m);
    std::vector<uint32_t> tok_doc, tok_slot, slot_wrd;
    layout_inputs(c, tok_doc, tok_slot, slot_wrd);
    c->a = a;
    return build_layout(c, tok_doc, tok_slot, slot_wrd);
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_autotune(ezlda_ctx *c, int32_t sweeps_per_trial, ezlda_args *chosen, double *ms_per_sweep,
                   ezlda_tune_trial *trials, int32_t max_trials, int32_t *num_trials) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_autotune");
    if (sweeps_per_trial < 1) return fail(c, EZLDA_EINVAL, "sweeps_per_trial must be >= 1");
    std::vector<uint32_t> tok_doc, tok_slot, slot_wrd;
    layout_inputs(c, tok_doc, tok_slot, slot_wrd);
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    const int V = c->s.V;
    int32_t nt = 0;
    auto trial = [&](const ezlda_args &a, double *ms) -> int {
        c->a = a;
        if (int r = build_layout(c, tok_doc, tok_slot, slot_wrd)) return r;
        double t = 0.0, st[EZLDA_NUM_STAGES];
        if (int r = ezlda_time_sweeps(c, 1, 0, &t, st)) return r;  // first sweep of a layout: not timed
        if (int r = ezlda_time_sweeps(c, sweeps_per_trial, 0, &t, st)) return r;
        *ms = t / sweeps_per_trial;
        if (trials && nt < max_trials) {
            trials[nt].args = a;
            trials[nt].ms_per_sweep = *ms;
        }
        ++nt;
        return EZLDA_OK;
    };
    ezlda_args best = c->a;
    {
        const std::string m = check_layout_args(best);
        if (!m.empty()) return fail(c, EZLDA_EINVAL, m);
    }
    double best_ms = 0.0;
    if (int r = trial(best, &best_ms)) return r;
    // coordinate descent, one argument at a time, in order of expected effect
    enum { kHead, kRun, kTiles, kBlock, kSeg, kGrid };
    for (int which : {kHead, kRun, kTiles, kBlock, kSeg, kGrid}) {
        std::vector<ezlda_args> cand;
        ezlda_args a = best;
        switch (which) {
        case kHead:  // dense head: a fraction of the vocabulary by rank
            for (double f : {0.0025, 0.005, 0.01, 0.02, 0.04}) {
                a.dense_words = std::max(1, (int)std::ceil(f * V));
                cand.push_back(a);
            }
            break;
        case kRun:
            for (int rc : {1, 2, 4, 8, 16, 32}) { a.run_cap = rc; cand.push_back(a); }
            break;
        case kTiles:
            for (int t : {0, -1}) { a.doc_tiles = t; cand.push_back(a); }
            break;
        case kBlock:
            for (int b : {64, 128, 256, 512}) { a.block_dim = b; cand.push_back(a); }
            break;
        case kSeg:
            for (int g : {1024, 4096, 16384}) { a.token_segment = g; cand.push_back(a); }
            break;
        case kGrid:
            for (int g : {0, 4 * ncu, 8 * ncu, 16 * ncu}) { a.grid_dim = g; cand.push_back(a); }
            break;
        }
        for (const ezlda_args &x : cand) {
            if (std::memcmp(&x, &best, sizeof(x)) == 0) continue;
            double ms = 0.0;
            if (int r = trial(x, &ms)) return r;
            if (ms < best_ms) {
                best_ms = ms;
                best = x;
            }
        }
    }
    c->a = best;
    if (int r = build_layout(c, tok_doc, tok_slot, slot_wrd)) return r;
    if (chosen) *chosen = best;
    if (ms_per_sweep) *ms_per_sweep = best_ms;
    if (num_trials) *num_trials = nt;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_time_sweeps(ezlda_ctx *c, int32_t n, int32_t use_graph, double *total_ms, double *stage_ms) try {
    if (int r = need_loaded(c)) return r;
    if (n < 1) return fail(c, EZLDA_EINVAL, "num_sweeps must be >= 1");
    if (is_chunked(c)) {  // whole sweeps only: stages overlap across the chunk streams
        if (use_graph) return refuse_chunked(c, "a captured sweep");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            return fail(c, EZLDA_EHIP, "hipEventCreate failed");
        }
        int rc = chunk_sweeps(c, n, e0, e1);
        float ms = 0.f;
        if (rc == EZLDA_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(c, EZLDA_EHIP, "hipEventElapsedTime failed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        if (total_ms) *total_ms = ms;
        if (stage_ms)
            for (int k = 0; k < EZLDA_NUM_STAGES; ++k) stage_ms[k] = -1.0;
        return EZLDA_OK;
    }
    double st[EZLDA_NUM_STAGES] = {0, 0, 0, 0};
    double tot = 0.0;
    if (use_graph) {
        if (int r = ensure_graph(c)) return r;
    }
    const int per = use_graph ? 2 : (EZLDA_NUM_STAGES + 1);
    const size_t nev = use_graph ? 2 : (size_t)n * per;
    std::vector<hipEvent_t> ev(nev, nullptr);
    int rc = EZLDA_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(c, EZLDA_EHIP, "hipEventCreate failed");
    if (rc == EZLDA_OK) {
        if (use_graph) {
            if (hipEventRecord(ev[0], c->st) != hipSuccess) rc = fail(c, EZLDA_EHIP, "hipEventRecord failed");
            for (int i = 0; i < n && rc == EZLDA_OK; ++i) {
                if (hipGraphLaunch(c->graph_exec, c->st) != hipSuccess) rc = fail(c, EZLDA_EHIP, "hipGraphLaunch failed");
                c->sweep++;
            }
            if (rc == EZLDA_OK && hipEventRecord(ev[1], c->st) != hipSuccess)
                rc = fail(c, EZLDA_EHIP, "hipEventRecord failed");
        } else {
            for (int i = 0; i < n && rc == EZLDA_OK; ++i) {
                rc = enqueue_sweep(c, &ev[(size_t)i * per]);
                c->sweep++;
            }
        }
    }
    if (rc == EZLDA_OK && hipStreamSynchronize(c->st) != hipSuccess) rc = fail(c, EZLDA_EHIP, "sweep failed");
    if (rc == EZLDA_OK) {
        float ms = 0.f;
        if (use_graph) {
            if (hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) rc = fail(c, EZLDA_EHIP, "hipEventElapsedTime failed");
            tot = ms;
            for (auto &x : st) x = -1.0;
        } else {
            if (hipEventElapsedTime(&ms, ev[0], ev[nev - 1]) != hipSuccess)
                rc = fail(c, EZLDA_EHIP, "hipEventElapsedTime failed");
            tot = ms;
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < EZLDA_NUM_STAGES; ++k) {
                    if (hipEventElapsedTime(&ms, ev[(size_t)i * per + k], ev[(size_t)i * per + k + 1]) != hipSuccess)
                        rc = fail(c, EZLDA_EHIP, "hipEventElapsedTime failed");
                    st[k] += ms;
                }
        }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    c->counts_fresh = false;
    if (rc) return rc;
    if (total_ms) *total_ms = tot;
    if (stage_ms)
        for (int k = 0; k < EZLDA_NUM_STAGES; ++k) stage_ms[k] = st[k];
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_stats(ezlda_ctx *c, uint64_t *stats) try {
    if (int r = need_loaded(c)) return r;
    unsigned long long v[2] = {0, 0};
    EZ_TRY_HIP(c, hipMemcpy(v, c->s.stats, sizeof v, hipMemcpyDeviceToHost));
    stats[0] = v[0];
    stats[1] = v[1];
    if (is_chunked(c)) {  // every row is dense; DT row lengths come back with each chunk
        uint64_t b = 0;
        for (const EzChunk &ch : c->chunks) {
            const uint2 *di = reinterpret_cast<const uint2 *>(ch.img + ch.off_doc_info);
            for (uint32_t q = 0; q < ch.D; ++q) b += di[q].y & 0xFFFFu;
        }
        stats[2] = 0;
        stats[3] = b;
        return EZLDA_OK;
    }
    std::vector<uint32_t> tn(c->s.V - c->s.VD);
    std::vector<uint2> di(c->s.D);
    if (!tn.empty()) EZ_TRY_HIP(c, hipMemcpy(tn.data(), c->s.tail_nnz, tn.size() * 4, hipMemcpyDeviceToHost));
    if (!di.empty()) EZ_TRY_HIP(c, hipMemcpy(di.data(), c->s.doc_info, di.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    uint64_t a = 0, b = 0;
    for (uint32_t x : tn) a += x;
    for (const uint2 &x : di) b += x.y & 0xFFFFu;
    stats[2] = a;
    stats[3] = b;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_reset_stats(ezlda_ctx *c) try {
    if (int r = need_loaded(c)) return r;
    EZ_TRY_HIP(c, hipMemsetAsync(c->s.stats, 0, 2 * sizeof(unsigned long long), c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

namespace {

// The log-likelihood of the current z.  part / out (both null, or both set)
// replace the context's per-item partials and result: the allocation-free
// entry point passes its static array there.  Same kernels, same order.
int loglik_into(ezlda_ctx *c, double *ll, double *part, double *out) {
    if (is_chunked(c)) {  // per chunk, summed in chunk order
        if (int r = ezc::refresh_counts(c)) return r;
        double tot = 0.0;
        for (int k = 0; k < (int)c->chunks.size(); ++k) {
            const EzChunk &ch = c->chunks[k];
            ez::DevState cs = chunk_state(c, k, 0);
            if (part) {
                cs.ll_part = part;
                cs.ll_out = out;
            }
            double v = 0.0;
            EZ_TRY_HIP(c, hipMemcpyAsync(c->cimg[0], ch.img, ch.bytes, hipMemcpyHostToDevice, c->st));
            EZ_TRY_HIP(c, ez::launch_build_dt(cs, c->st));
            EZ_TRY_HIP(c, ez::launch_loglik(cs, c->st));
            EZ_TRY_HIP(c, hipMemcpyAsync(&v, cs.ll_out, sizeof(double), hipMemcpyDeviceToHost, c->st));
            EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
            tot += v;
        }
        *ll = tot;
        return EZLDA_OK;
    }
    ez::DevState s = c->s;
    if (part) {
        s.ll_part = part;
        s.ll_out = out;
    }
    if (!c->counts_fresh) EZ_TRY_HIP(c, ez::launch_build_wt(s, c->st));
    c->counts_fresh = true;
    EZ_TRY_HIP(c, ez::launch_topk(s, c->st));
    EZ_TRY_HIP(c, ez::launch_build_dt(s, c->st));
    EZ_TRY_HIP(c, ez::launch_loglik(s, c->st));
    EZ_TRY_HIP(c, hipMemcpyAsync(ll, s.ll_out, sizeof(double), hipMemcpyDeviceToHost, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    return EZLDA_OK;
}

}  // namespace

int ezlda_loglik(ezlda_ctx *c, double *ll) try {
    if (int r = need_loaded(c)) return r;
    if (!ll) return fail(c, EZLDA_EINVAL, "null result pointer");
    return loglik_into(c, ll, nullptr, nullptr);
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_loglik_noalloc(ezlda_ctx *c, double *ll) try {
    if (int r = need_loaded(c)) return r;
    if (!ll) return fail(c, EZLDA_EINVAL, "null result pointer");
    if (!c->ll_scratch) return fail(c, EZLDA_EHIP, "the static log-likelihood array is unavailable: " + c->scratch_err);
    // one partial per sampler item of the launch (a chunk's items when chunked)
    uint64_t items = (uint64_t)std::max<int32_t>(c->s.n_smp_segs, 0);
    if (is_chunked(c)) {
        items = 0;
        for (const EzChunk &ch : c->chunks) items = std::max<uint64_t>(items, ch.n_items);
    }
    if (items > ez::kLLScratchItems)
        return fail(c, EZLDA_EINVAL, "the corpus has " + std::to_string(items) + " sampler items; "
                    "ezlda_loglik_noalloc holds at most " + std::to_string(ez::kLLScratchItems) +
                    " (EZLDA_LOGLIK_NOALLOC_MAX_ITEMS): use ezlda_loglik");
    // every partial the final sum reads is written by the same launch, so
    // nothing needs clearing
    return loglik_into(c, ll, c->ll_scratch, c->ll_scratch + ez::kLLScratchItems);
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_topics(ezlda_ctx *c, uint16_t *zout) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) {  // from the chunk images (z by doc slot)
        for (size_t i = 0; i < c->doc_tok.size(); ++i) {
            const uint32_t p = c->tok_slot_g[c->doc_tok[i]];
            const EzChunk &ch = c->chunks[chunk_of_slot(c, p)];
            zout[i] = reinterpret_cast<const uint16_t *>(ch.img + 2 * (size_t)ch.N)[p - ch.slot0];
        }
        return EZLDA_OK;
    }
    std::vector<uint16_t> z(c->s.N);
    EZ_TRY_HIP(c, hipMemcpy(z.data(), c->s.z, z.size() * sizeof(uint16_t), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < z.size(); ++i) zout[i] = z[c->doc_tok[i]];
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_set_topics(ezlda_ctx *c, const uint16_t *zin) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) {
        for (size_t i = 0; i < c->doc_tok.size(); ++i)
            if (zin[i] >= c->s.K) return fail(c, EZLDA_EINVAL, "topic out of range");
        for (size_t i = 0; i < c->doc_tok.size(); ++i) {
            const uint32_t p = c->tok_slot_g[c->doc_tok[i]];
            EzChunk &ch = c->chunks[chunk_of_slot(c, p)];
            const uint32_t pp = p - ch.slot0;
            const uint32_t l = reinterpret_cast<const uint32_t *>(ch.img + ch.off_slot_tok)[pp];
            reinterpret_cast<uint16_t *>(ch.img)[l] = zin[i];
            reinterpret_cast<uint16_t *>(ch.img + 2 * (size_t)ch.N)[pp] = zin[i];
        }
        c->counts_fresh = false;
        return EZLDA_OK;
    }
    std::vector<uint16_t> z(c->s.N), z_doc(c->s.N);
    for (size_t i = 0; i < z.size(); ++i) {
        if (zin[i] >= c->s.K) return fail(c, EZLDA_EINVAL, "topic out of range");
        z[c->doc_tok[i]] = zin[i];
    }
    for (size_t p = 0; p < z.size(); ++p) z_doc[p] = z[c->slot_tok[p]];
    EZ_TRY_HIP(c, hipMemcpy(c->s.z, z.data(), z.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    EZ_TRY_HIP(c, hipMemcpy(c->s.z_doc, z_doc.data(), z_doc.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    c->counts_fresh = false;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_set_sweep(ezlda_ctx *c, uint32_t sweep) try {
    if (int r = need_loaded(c)) return r;
    EZ_TRY_HIP(c, ez::launch_set_u32(c->s.counters, sweep, c->st));
    EZ_TRY_HIP(c, hipStreamSynchronize(c->st));
    c->sweep = sweep;
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_nk(ezlda_ctx *c, int32_t *nk) try {
    if (int r = need_loaded(c)) return r;
    EZ_TRY_HIP(c, hipMemcpy(nk, c->s.nk, sizeof(int32_t) * c->s.K, hipMemcpyDeviceToHost));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_inv(ezlda_ctx *c, float *inv) try {
    if (int r = need_loaded(c)) return r;
    EZ_TRY_HIP(c, hipMemcpy(inv, c->s.inv, sizeof(float) * c->s.K, hipMemcpyDeviceToHost));
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_wt_dense(ezlda_ctx *c, int32_t *wt) try {
    if (int r = need_loaded(c)) return r;
    const ez::DevState &s = c->s;
    std::vector<int32_t> head((size_t)s.VD * s.K);
    std::vector<uint32_t> toff(s.V - s.VD + 1), tnnz(s.V - s.VD), tail(c->tail_cap);
    if (!head.empty()) EZ_TRY_HIP(c, hipMemcpy(head.data(), s.wt_head, head.size() * 4, hipMemcpyDeviceToHost));
    EZ_TRY_HIP(c, hipMemcpy(toff.data(), s.tail_off, toff.size() * 4, hipMemcpyDeviceToHost));
    if (!tnnz.empty()) EZ_TRY_HIP(c, hipMemcpy(tnnz.data(), s.tail_nnz, tnnz.size() * 4, hipMemcpyDeviceToHost));
    if (!tail.empty()) EZ_TRY_HIP(c, hipMemcpy(tail.data(), s.wt_tail, tail.size() * 4, hipMemcpyDeviceToHost));
    std::memset(wt, 0, sizeof(int32_t) * (size_t)s.V * s.K);
    for (int w = 0; w < s.V; ++w) {
        const int r = c->rank_of[w];
        int32_t *out = wt + (size_t)w * s.K;
        if (r < s.VD) {
            std::memcpy(out, head.data() + (size_t)r * s.K, sizeof(int32_t) * s.K);
        } else {
            const int tw = r - s.VD;
            for (uint32_t i = 0; i < tnnz[tw]; ++i) {
                const uint32_t e = tail[toff[tw] + i];
                out[e & 0xFFFFu] = (int32_t)(e >> 16);
            }
        }
    }
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_export_model(ezlda_ctx *c, const char *dir) try {
    if (int r = need_loaded(c)) return r;
    if (!dir) return fail(c, EZLDA_EINVAL, "null directory");
    if (int r = ezc::refresh_counts(c)) return r;
    const ez::DevState &s = c->s;
    const std::string d(dir);
    std::vector<int32_t> wt((size_t)s.V * s.K), nk(s.K);
    std::vector<float> inv(s.K);
    if (int r = ezlda_get_wt_dense(c, wt.data())) return r;
    EZ_TRY_HIP(c, hipMemcpy(inv.data(), s.inv, sizeof(float) * s.K, hipMemcpyDeviceToHost));
    {
        std::FILE *f = std::fopen((d + "/phi.f32").c_str(), "wb");
        if (!f) return fail(c, EZLDA_EIO, "cannot write " + d + "/phi.f32");
        std::vector<float> row(s.K);
        bool ok = true;
        for (int w = 0; w < s.V && ok; ++w) {
            for (int k = 0; k < s.K; ++k) row[k] = ((float)wt[(size_t)w * s.K + k] + s.beta) * inv[k];
            ok = std::fwrite(row.data(), sizeof(float), s.K, f) == (size_t)s.K;
        }
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(c, EZLDA_EIO, "short write to " + d + "/phi.f32");
    }
    {  // per-document topic counts of the current topics, ascending topic
        std::vector<uint16_t> z(s.N);
        if (int r = ezlda_get_topics(c, z.data())) return r;
        std::FILE *f = std::fopen((d + "/theta.txt").c_str(), "w");
        if (!f) return fail(c, EZLDA_EIO, "cannot write " + d + "/theta.txt");
        std::vector<uint32_t> cnt(s.K, 0);
        std::vector<uint16_t> seen;
        bool ok = true;
        for (int doc = 0; doc < s.D && ok; ++doc) {  // input doc order
            const uint32_t b = c->doc_off[doc], n = c->doc_off[doc + 1] - b;
            ok = std::fprintf(f, "%u", n) > 0;
            seen.clear();
            for (uint32_t i = 0; i < n; ++i)
                if (cnt[z[b + i]]++ == 0) seen.push_back(z[b + i]);
            std::sort(seen.begin(), seen.end());
            for (const uint16_t k : seen) {
                ok = ok && std::fprintf(f, " %u:%u", (unsigned)k, cnt[k]) > 0;
                cnt[k] = 0;
            }
            ok = ok && std::fputc('\n', f) != EOF;
        }
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(c, EZLDA_EIO, "short write to " + d + "/theta.txt");
    }
    std::FILE *f = std::fopen((d + "/model.txt").c_str(), "w");
    if (!f) return fail(c, EZLDA_EIO, "cannot write " + d + "/model.txt");
    const bool ok = std::fprintf(f, "%d %d %d %.9g %.9g %u\n", s.K, s.V, s.D, (double)s.alpha, (double)s.beta,
                                 c->sweep) > 0;
    if ((std::fclose(f) != 0) || !ok) return fail(c, EZLDA_EIO, "short write to " + d + "/model.txt");
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_dt_raw(ezlda_ctx *c, uint32_t *entries, int32_t *nnz) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_get_dt_raw");
    const ez::DevState &s = c->s;
    std::vector<uint2> di(s.D);
    std::vector<uint32_t> dt(s.compact ? (c->dt_cap + 1) / 2 : c->dt_cap);
    if (s.D) EZ_TRY_HIP(c, hipMemcpy(di.data(), s.doc_info, di.size() * sizeof(uint2), hipMemcpyDeviceToHost));
    if (!dt.empty()) EZ_TRY_HIP(c, hipMemcpy(dt.data(), s.dt, dt.size() * 4, hipMemcpyDeviceToHost));
    const uint16_t *dt16 = reinterpret_cast<const uint16_t *>(dt.data());
    std::vector<int32_t> qof(s.D);
    for (int q = 0; q < s.D; ++q) qof[c->dord[q]] = q;
    uint64_t canon = 0;  // rows packed in input doc order at sum_{e<d} min(n_e, K)
    for (int d = 0; d < s.D; ++d) {
        const uint2 x = di[qof[d]];
        nnz[d] = (int32_t)(x.y & 0xFFFFu);
        for (int32_t i = 0; i < nnz[d]; ++i) {  // device rows hold topic << 2, or 16-bit entries
            if (s.compact) {
                const uint32_t e = dt16[x.x + i];
                entries[canon + i] = (e & ((1u << s.topic_bits) - 1u)) | ((e >> s.topic_bits) << 16);
            } else {
                const uint32_t e = dt[x.x + i];
                entries[canon + i] = ((e & 0xFFFFu) >> 2) | (e & 0xFFFF0000u);
            }
        }
        canon += dt_cap(c->doc_off[d + 1] - c->doc_off[d], s.K);
    }
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_top(ezlda_ctx *c, uint16_t *top) try {
    if (int r = need_loaded(c)) return r;
    std::vector<ushort4> t(c->s.V);
    EZ_TRY_HIP(c, hipMemcpy(t.data(), c->s.wtop, t.size() * sizeof(ushort4), hipMemcpyDeviceToHost));
    for (int w = 0; w < c->s.V; ++w) {
        const ushort4 &x = t[c->rank_of[w]];
        top[3 * (size_t)w] = x.x;
        top[3 * (size_t)w + 1] = x.y;
        top[3 * (size_t)w + 2] = x.z;
    }
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_get_aux(ezlda_ctx *c, uint32_t *aux) try {
    if (int r = need_loaded(c)) return r;
    if (is_chunked(c)) return refuse_chunked(c, "ezlda_get_aux");
    // (n_dk1, n_dk2) of the current topics: the DT rebuild in its other record
    // form, into scratch rows, doc_info and records, so no state changes
    const ez::DevState &s = c->s;
    ez::DevState t = s;
    const size_t dt_words = s.compact ? ((size_t)c->dt_cap + 256 + 1) / 2 : (size_t)c->dt_cap + 256;
    t.aux_doc = nullptr;
    t.dt = nullptr;
    t.doc_info = nullptr;
    int rc = EZLDA_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (rc == EZLDA_OK && e != hipSuccess) rc = fail(c, EZLDA_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(hipMalloc((void **)&t.aux_doc, std::max<size_t>(s.N, 1) * sizeof(uint32_t)), "hipMalloc");
    chk(hipMalloc((void **)&t.dt, dt_words * sizeof(uint32_t)), "hipMalloc");
    chk(hipMalloc((void **)&t.doc_info, std::max<size_t>(s.D, 1) * sizeof(uint2)), "hipMalloc");
    if (rc == EZLDA_OK && s.D)
        chk(hipMemcpyAsync(t.doc_info, s.doc_info, (size_t)s.D * sizeof(uint2), hipMemcpyDeviceToDevice, c->st),
            "hipMemcpyAsync");
    if (rc == EZLDA_OK) chk(ez::launch_build_dt(t, c->st, true), "DT rebuild");
    std::vector<uint32_t> a(s.N), at(s.N);
    if (rc == EZLDA_OK && s.N)
        chk(hipMemcpyAsync(a.data(), t.aux_doc, a.size() * 4, hipMemcpyDeviceToHost, c->st), "hipMemcpyAsync");
    if (rc == EZLDA_OK) chk(hipStreamSynchronize(c->st), "hipStreamSynchronize");
    if (t.aux_doc) (void)hipFree(t.aux_doc);
    if (t.dt) (void)hipFree(t.dt);
    if (t.doc_info) (void)hipFree(t.doc_info);
    if (rc) return rc;
    for (size_t p = 0; p < a.size(); ++p) at[c->slot_tok[p]] = a[p];
    for (size_t i = 0; i < a.size(); ++i) aux[i] = at[c->doc_tok[i]];
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_selftest_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, int32_t n) try {
    if (n < 1 || !ctr || !key || !out) return fail(nullptr, EZLDA_EINVAL, "bad selftest arguments");
    uint32_t *dc = nullptr, *dk = nullptr, *dout = nullptr;
    EZ_TRY_HIP(nullptr, hipMalloc(&dc, (size_t)n * 16));
    EZ_TRY_HIP(nullptr, hipMalloc(&dk, (size_t)n * 8));
    EZ_TRY_HIP(nullptr, hipMalloc(&dout, (size_t)n * 16));
    EZ_TRY_HIP(nullptr, hipMemcpy(dc, ctr, (size_t)n * 16, hipMemcpyHostToDevice));
    EZ_TRY_HIP(nullptr, hipMemcpy(dk, key, (size_t)n * 8, hipMemcpyHostToDevice));
    EZ_TRY_HIP(nullptr, ez::launch_selftest_philox(dc, dk, dout, n, nullptr));
    EZ_TRY_HIP(nullptr, hipMemcpy(out, dout, (size_t)n * 16, hipMemcpyDeviceToHost));
    (void)hipFree(dc);
    (void)hipFree(dk);
    (void)hipFree(dout);
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(nullptr);
}

int ezlda_selftest_copy(int64_t bytes, int32_t reps, double *ms) try {
    if (bytes < 16 || reps < 1) return fail(nullptr, EZLDA_EINVAL, "bad copy calibration arguments");
    bytes &= ~(int64_t)15;
    void *a = nullptr, *b = nullptr;
    EZ_TRY_HIP(nullptr, hipMalloc(&a, (size_t)bytes));
    EZ_TRY_HIP(nullptr, hipMalloc(&b, (size_t)bytes));
    EZ_TRY_HIP(nullptr, hipMemset(a, 1, (size_t)bytes));
    EZ_TRY_HIP(nullptr, hipDeviceSynchronize());
    hipEvent_t e0, e1;
    EZ_TRY_HIP(nullptr, hipEventCreate(&e0));
    EZ_TRY_HIP(nullptr, hipEventCreate(&e1));
    EZ_TRY_HIP(nullptr, ez::launch_copy(a, b, (size_t)bytes, nullptr));
    EZ_TRY_HIP(nullptr, hipEventRecord(e0, nullptr));
    for (int i = 0; i < reps; ++i) EZ_TRY_HIP(nullptr, ez::launch_copy(a, b, (size_t)bytes, nullptr));
    EZ_TRY_HIP(nullptr, hipEventRecord(e1, nullptr));
    EZ_TRY_HIP(nullptr, hipEventSynchronize(e1));
    float t = 0.f;
    const hipError_t et = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (et != hipSuccess) return fail(nullptr, EZLDA_EHIP, "hipEventElapsedTime failed");
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(nullptr);
}

int ezlda_selftest_gather(int64_t bytes, int64_t n_reads, int32_t reps, double *ms) try {
    if (bytes < 64 || n_reads < 1 || reps < 1) return fail(nullptr, EZLDA_EINVAL, "bad gather calibration arguments");
    bytes &= ~(int64_t)63;
    int64_t nblk = 1;
    while (nblk * 2 * 64 <= bytes) nblk *= 2;
    if (n_reads > nblk)
        return fail(nullptr, EZLDA_EINVAL, "n_reads exceeds the buffer's " + std::to_string(nblk) + " distinct blocks");
    void *a = nullptr;
    uint32_t *sink = nullptr;
    EZ_TRY_HIP(nullptr, hipMalloc(&a, (size_t)bytes));
    EZ_TRY_HIP(nullptr, hipMalloc((void **)&sink, 64 * sizeof(uint32_t)));
    EZ_TRY_HIP(nullptr, hipMemset(a, 1, (size_t)bytes));
    EZ_TRY_HIP(nullptr, hipDeviceSynchronize());
    hipEvent_t e0, e1;
    EZ_TRY_HIP(nullptr, hipEventCreate(&e0));
    EZ_TRY_HIP(nullptr, hipEventCreate(&e1));
    EZ_TRY_HIP(nullptr, ez::launch_gather64(a, (size_t)bytes, (uint64_t)n_reads, sink, nullptr));
    EZ_TRY_HIP(nullptr, hipEventRecord(e0, nullptr));
    for (int i = 0; i < reps; ++i)
        EZ_TRY_HIP(nullptr, ez::launch_gather64(a, (size_t)bytes, (uint64_t)n_reads, sink, nullptr));
    EZ_TRY_HIP(nullptr, hipEventRecord(e1, nullptr));
    EZ_TRY_HIP(nullptr, hipEventSynchronize(e1));
    float t = 0.f;
    const hipError_t et = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(sink);
    if (et != hipSuccess) return fail(nullptr, EZLDA_EHIP, "hipEventElapsedTime failed");
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(nullptr);
}

// Checkpoint: "EZLDACK1", then int64 K, V, D, N, uint64 seed, uint32 sweep,
// then z (doc-major, uint16).  Restoring z and the sweep counter restores
// the whole sampler state: counts are rebuilt from z and the draws are keyed
// by (seed, token, sweep).
int ezlda_save_checkpoint(ezlda_ctx *c, const char *path) try {
    if (int r = need_loaded(c)) return r;
    std::vector<uint16_t> z(c->s.N);
    if (int r = ezlda_get_topics(c, z.data())) return r;
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(c, EZLDA_EIO, std::string("cannot write ") + path);
    const int64_t hdr[4] = {c->s.K, c->s.V, c->s.D, (int64_t)c->s.N};
    const uint64_t seed = c->a.seed;
    const uint32_t sweep = c->sweep;
    bool ok = std::fwrite("EZLDACK1", 1, 8, f) == 8 && std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(&seed, sizeof seed, 1, f) == 1 && std::fwrite(&sweep, sizeof sweep, 1, f) == 1 &&
              (z.empty() || std::fwrite(z.data(), 2, z.size(), f) == z.size());
    ok = (std::fclose(f) == 0) && ok;
    return ok ? EZLDA_OK : fail(c, EZLDA_EIO, std::string("short write to ") + path);
} catch (...) {
    return ezc::exception_status(c);
}

// The reference's own checkpoint form: the chunked folder's TL.txt files
// rewritten with the current topics (streamTestBalance.cpp:469 writes the
// initial ones there; DataChunk.cu:85 reads them back) and the Philox state
// in <prefix>/ezlda_state.txt.
int ezlda_save_tl(ezlda_ctx *c, const char *prefix) try {
    if (int r = need_loaded(c)) return r;
    if (!prefix) return fail(c, EZLDA_EINVAL, "null prefix");
    const std::string root(prefix);
    std::vector<uint64_t> pos;
    std::vector<int32_t> words;
    std::vector<int64_t> tll;
    if (!ezc::chunked_layout(root, pos, words, tll)) return fail(c, EZLDA_EIO, ezlda_last_error(nullptr));
    const size_t N = c->doc_tok.size();
    if (pos.size() != N || words.size() != N || (int64_t)c->doc_off.size() - 1 != c->s.D)
        return fail(c, EZLDA_EINVAL, root + " holds " + std::to_string(pos.size()) + " tokens; the context " +
                                         std::to_string(N));
    // the folder must hold the context's corpus: every doc-major token's word
    std::vector<int32_t> id_of_rank(c->rank_of.size());
    for (size_t w = 0; w < c->rank_of.size(); ++w) id_of_rank[c->rank_of[w]] = (int32_t)w;
    std::vector<uint32_t> rank_of_tok(N);
    for (size_t r = 0; r + 1 < c->word_off.size(); ++r)
        for (uint32_t t = c->word_off[r]; t < c->word_off[r + 1]; ++t) rank_of_tok[t] = (uint32_t)r;
    for (size_t i = 0; i < N; ++i)
        if (id_of_rank[rank_of_tok[c->doc_tok[i]]] != words[i])
            return fail(c, EZLDA_EINVAL, root + " does not hold the context's corpus (token " + std::to_string(i) + ")");
    std::vector<uint16_t> z(N);
    if (int r = ezlda_get_topics(c, z.data())) return r;
    if (!ezc::rewrite_tl(root, pos, tll, z.data(), c->s.K)) return fail(c, EZLDA_EIO, ezlda_last_error(nullptr));
    const std::string sp = root + "/ezlda_state.txt";
    FILE *f = std::fopen(sp.c_str(), "w");
    if (!f) return fail(c, EZLDA_EIO, "cannot write " + sp);
    const bool ok = std::fprintf(f, "seed %llu\nsweep %u\nK %d\n", (unsigned long long)c->a.seed, c->sweep,
                                 c->s.K) > 0;
    if (std::fclose(f) != 0 || !ok) return fail(c, EZLDA_EIO, "cannot write " + sp);
    return EZLDA_OK;
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_load_tl(ezlda_ctx *c, const char *prefix) try {
    if (!c) return fail(nullptr, EZLDA_EINVAL, "null context");
    if (!prefix) return fail(c, EZLDA_EINVAL, "null prefix");
    const std::string root(prefix), sp = root + "/ezlda_state.txt";
    unsigned long long seed = c->a.seed;
    unsigned sweep = 0;
    int K = c->a.K;
    bool have_state = false;
    if (FILE *f = std::fopen(sp.c_str(), "r")) {
        const int got = std::fscanf(f, "seed %llu sweep %u K %d", &seed, &sweep, &K);
        std::fclose(f);
        if (got != 3) return fail(c, EZLDA_EIO, sp + ": expected \"seed S sweep N K K\"");
        have_state = true;
    }
    if (have_state && (seed != c->a.seed || K != c->a.K))
        return fail(c, EZLDA_EINVAL, sp + " was written with seed " + std::to_string(seed) + ", K " +
                                         std::to_string(K) + "; the context has seed " + std::to_string(c->a.seed) +
                                         ", K " + std::to_string(c->a.K));
    if (int r = ezlda_load_chunked(c, prefix, 0)) return r;
    return ezlda_set_sweep(c, sweep);
} catch (...) {
    return ezc::exception_status(c);
}

int ezlda_load_checkpoint(ezlda_ctx *c, const char *path) try {
    if (int r = need_loaded(c)) return r;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, EZLDA_EIO, std::string("cannot read ") + path);
    char magic[8];
    int64_t hdr[4];
    uint64_t seed = 0;
    uint32_t sweep = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "EZLDACK1", 8) == 0 &&
              std::fread(hdr, sizeof hdr, 1, f) == 1 && std::fread(&seed, sizeof seed, 1, f) == 1 &&
              std::fread(&sweep, sizeof sweep, 1, f) == 1;
    if (!ok) {
        std::fclose(f);
        return fail(c, EZLDA_EIO, std::string("not an ezlda checkpoint: ") + path);
    }
    if (hdr[0] != c->s.K || hdr[1] != c->s.V || hdr[2] != c->s.D || hdr[3] != (int64_t)c->s.N || seed != c->a.seed) {
        std::fclose(f);
        return fail(c, EZLDA_EIO, "checkpoint does not match the loaded corpus, K or seed");
    }
    std::vector<uint16_t> z(c->s.N);
    ok = z.empty() || std::fread(z.data(), 2, z.size(), f) == z.size();
    std::fclose(f);
    if (!ok) return fail(c, EZLDA_EIO, std::string("truncated checkpoint: ") + path);
    if (int r = ezlda_set_topics(c, z.data())) return r;
    return ezlda_set_sweep(c, sweep);
} catch (...) {
    return ezc::exception_status(c);
}

}  // extern "C"
