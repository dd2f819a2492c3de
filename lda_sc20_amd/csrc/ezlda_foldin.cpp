// Fold-in serving: topic mixtures of incoming documents against a frozen
// model, with continuous batching (no counterpart in the reference, which
// only trains).
//
// The server freezes phi from a trained context (its counts at creation) and
// keeps it on the device with the per-word Q running sums.  Requests carry a
// document; a worker thread admits them into device slots while slots and
// the HBM token pool (the budget) allow, and advances every resident request
// by one fold-in sweep per step, so requests that arrive while others run
// join the next step.  A step is one upload (slot table, finishing table,
// admitted documents' words), the word-major fold-in kernels, and one copy
// back of the finishing documents' counts and topics.  A request leaves
// after `sweeps` sweeps; its pool range and slot are reused.  Draws are keyed
// by (seed, request id, token, sweep), never by slot or position, so a result
// is the same alone and in any mix.  ezlda_foldin_plain, the comparison
// path, runs one document per launch (all of its sweeps in one workgroup) on
// its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/ezlda.h"
#include "ezlda_device.h"
#include "ezlda_internal.h"

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint32_t rid = 0;
    std::vector<uint32_t> words;  // ranks
    double t_submit = 0.0;
};

struct Slot {
    bool active = false;
    uint32_t rid = 0, off = 0, n = 0, sweep = 0;
    double t_submit = 0.0;
};

struct Result {
    std::vector<uint16_t> z;
    std::vector<int32_t> counts;
    double latency_ms = 0.0;
};

// A document finishing with a step: where its topics land in the step's result
// copy (the host counts them).
struct Finishing {
    uint32_t rid = 0, n = 0, fidx = 0, zoff = 0;
    double t_submit = 0.0;
};

// Host side of one step in flight: pinned upload (slot table, finishing
// table, admitted documents), pinned result staging (grown on demand), the
// event after its result copies, and the documents it finishes.
struct StepRec {
    uint8_t *up_h = nullptr;
    uint16_t *z_h = nullptr;
    size_t z_cap = 0;
    hipEvent_t ev = nullptr;
    std::vector<Finishing> fins;
    bool busy = false;
};

}  // namespace

struct ezlda_server {
    int K = 0, V = 0, NS = 0, sweeps = 20, max_slots = 0;
    float alpha = 0.f;
    uint64_t seed = 0;
    std::vector<int32_t> rank_of;
    // device: frozen model, pools (in tokens; ranges are multiples of 8),
    // step buffers, upload buffer
    float *phi = nullptr, *qs = nullptr, *qacc = nullptr;
    uint32_t *wpool = nullptr;
    uint16_t *zpool = nullptr;
    ez::FoldinBatch fb;
    uint8_t *up_d = nullptr;
    size_t up_bytes = 0;
    uint32_t budget_tokens = 0, pool_tokens = 0, up_cap = 0;
    StepRec rec[2];     // two steps in flight: host buffers and finishing documents
    int64_t retiring = 0;  // documents done on the device, results not yet published
    // plain path: one document's words (ranks) and topics
    uint32_t *pw = nullptr;
    uint16_t *pz = nullptr;
    hipStream_t st = nullptr, pst = nullptr;
    std::vector<void *> allocs;
    // token pool allocator: free ranges offset -> length
    std::map<uint32_t, uint32_t> free_ranges;
    std::vector<Slot> slots;
    std::vector<int> free_slots, active;
    std::deque<Request> pending;
    std::unordered_map<uint32_t, Result> done;
    std::unordered_set<uint32_t> inflight;  // queued or resident request ids
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    std::string err;
    int worker_rc = EZLDA_OK;
    // statistics
    std::vector<double> lat;
    int64_t submitted = 0, completed = 0, rejected = 0, steps = 0;
    int max_active = 0;
    double t_first = -1.0, t_last = 0.0;
    std::mutex plain_mu;
    int device = 0;
};

namespace {

int sfail(ezlda_server *s, int code, const std::string &m) {
    if (s) s->err = m;
    ezc::set_error(m);
    return code;
}

}  // namespace

int ezc::server_exception_status(ezlda_server *s) {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return sfail(s, EZLDA_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return sfail(s, EZLDA_EINVAL, std::string("unexpected error: ") + e.what());
    } catch (...) {
        return sfail(s, EZLDA_EINVAL, "unexpected error");
    }
}

namespace {

#define SV_HIP(srv, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return sfail((srv), EZLDA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int salloc(ezlda_server *s, T **p, size_t n) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return sfail(s, EZLDA_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocs.push_back((void *)*p);
    return EZLDA_OK;
}

// First fit in the token pool; returns false when no range is long enough.
bool pool_alloc(ezlda_server *s, uint32_t n, uint32_t *off) {
    if (n == 0) {
        *off = 0;
        return true;
    }
    for (auto it = s->free_ranges.begin(); it != s->free_ranges.end(); ++it) {
        if (it->second >= n) {
            *off = it->first;
            const uint32_t rest = it->second - n, o = it->first + n;
            s->free_ranges.erase(it);
            if (rest) s->free_ranges[o] = rest;
            return true;
        }
    }
    return false;
}

void pool_free(ezlda_server *s, uint32_t off, uint32_t n) {
    if (n == 0) return;
    auto it = s->free_ranges.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != s->free_ranges.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        s->free_ranges.erase(nx);
    }
    if (it != s->free_ranges.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            s->free_ranges.erase(it);
        }
    }
}

uint32_t round8(uint32_t n) { return (n + 7u) & ~7u; }
size_t align16(size_t n) { return (n + 15u) & ~(size_t)15u; }

// Grows a pinned host buffer to hold `need` elements (contents not kept).
template <class T>
bool ensure_pinned(T **p, size_t *cap, size_t need) {
    if (need <= *cap) return true;
    const size_t n = std::max(need, *cap * 2);
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipHostMalloc((void **)p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) {
        *p = nullptr;
        return false;
    }
    *cap = n;
    return true;
}

// Upload buffer layout for ns slots and na admitted documents:
// [slots: ns uint4][fin: ns uint2][pad to 16][adm: na uint4][words]
struct UpLayout {
    size_t fin, adm, words, end;
};
UpLayout up_layout(size_t ns, size_t na, size_t nw) {
    UpLayout l;
    l.fin = ns * sizeof(uint4);
    l.adm = align16(l.fin + ns * sizeof(uint2));
    l.words = l.adm + na * sizeof(uint4);
    l.end = l.words + nw * sizeof(uint32_t);
    return l;
}

// Waits for a step's result copies and publishes its finishing documents.
bool publish(ezlda_server *s, StepRec &R) {
    const hipError_t e = hipEventSynchronize(R.ev);
    R.busy = false;
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_rc = EZLDA_EHIP;
        s->err = std::string("fold-in step failed: ") + hipGetErrorString(e);
        s->cv_done.notify_all();
        return false;
    }
    std::vector<Result> res(R.fins.size());
    for (size_t f = 0; f < R.fins.size(); ++f) {
        const Finishing &x = R.fins[f];
        res[f].z.assign(R.z_h + x.zoff, R.z_h + x.zoff + x.n);
        res[f].counts.assign(s->K, 0);
        for (const uint16_t t : res[f].z) res[f].counts[t]++;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = now_ms();
    for (size_t f = 0; f < R.fins.size(); ++f) {
        Result &r = res[f];
        r.latency_ms = t - R.fins[f].t_submit;
        s->lat.push_back(r.latency_ms);
        s->completed++;
        s->t_last = t;
        s->done[R.fins[f].rid] = std::move(r);
        s->inflight.erase(R.fins[f].rid);
    }
    s->retiring -= (int64_t)R.fins.size();
    if (!R.fins.empty()) s->cv_done.notify_all();
    R.fins.clear();
    return true;
}

// Worker: admit, build and enqueue a step, publish the previous one, until
// stopped.  Two steps are in flight: the host builds step s + 1 while the
// device runs step s.  Documents finishing with a step give back their slot
// and pool range when the step is built: the stream runs the step (and its
// result copies) before any later step's upload or admission touches them.
void worker_main(ezlda_server *s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_rc = EZLDA_EHIP;
        s->err = "fold-in worker: hipSetDevice failed";
        s->cv_done.notify_all();
        return;
    }
    std::vector<Request> adm;
    std::vector<uint4> adm_rec;
    int cur = 0;
    while (true) {
        StepRec &R = s->rec[cur], &P = s->rec[cur ^ 1];
        adm.clear();
        adm_rec.clear();
        uint32_t up_words = 0, ns = 0, nfin = 0, zsum = 0, ntok = 0;
        UpLayout L{};
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // admission: a free slot, pool room and upload room, in arrival
            // order, at most 2 * ceil(slots / sweeps) documents per step: a
            // burst then enters as a steady flow instead of a cohort that
            // finishes (and is replaced) in one step
            const size_t pace = 2 * (((size_t)s->max_slots + (size_t)s->sweeps - 1) / (size_t)s->sweeps);
            while (!s->pending.empty() && !s->free_slots.empty() && adm.size() < pace) {
                const uint32_t n = (uint32_t)s->pending.front().words.size();
                if (up_words + n > s->up_cap) break;
                uint32_t off = 0;
                if (!pool_alloc(s, round8(n), &off)) break;
                adm.push_back(std::move(s->pending.front()));
                s->pending.pop_front();
                const int j = s->free_slots.back();
                s->free_slots.pop_back();
                Slot &x = s->slots[j];
                x.active = true;
                x.rid = adm.back().rid;
                x.off = off;
                x.n = n;
                x.sweep = 0;
                x.t_submit = adm.back().t_submit;
                s->active.push_back(j);
                adm_rec.push_back(make_uint4(up_words, off, n, 0u));
                up_words += n;
            }
            s->max_active = std::max(s->max_active, (int)s->active.size());
            if (s->active.empty() && !P.busy) {
                if (s->stop) return;
                s->cv_work.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
            if (!s->active.empty()) {
                // the slot table, the finishing table and the bookkeeping of
                // this step: sweeps advance, finishing documents retire
                ns = (uint32_t)s->active.size();
                L = up_layout(ns, adm_rec.size(), up_words);
                uint4 *hs = (uint4 *)R.up_h;
                uint2 *hf = (uint2 *)(R.up_h + L.fin);
                R.fins.clear();
                for (uint32_t j = 0; j < ns; ++j) {
                    const int si = s->active[j];
                    Slot &x = s->slots[si];
                    hs[j] = make_uint4(x.rid, x.off, x.n, x.sweep);
                    ntok += x.n;
                    if (++x.sweep == (uint32_t)s->sweeps) {
                        Finishing f;
                        f.rid = x.rid;
                        f.n = x.n;
                        f.fidx = nfin++;
                        f.zoff = zsum;
                        f.t_submit = x.t_submit;
                        R.fins.push_back(f);
                        hf[j] = make_uint2(f.fidx, f.zoff);
                        zsum += x.n;
                        pool_free(s, x.off, round8(x.n));
                        x.active = false;
                        s->free_slots.push_back(si);
                    } else {
                        hf[j] = make_uint2(0xFFFFFFFFu, 0u);
                    }
                }
                if (nfin) {
                    s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                                   [&](int i) { return !s->slots[i].active; }),
                                    s->active.end());
                    s->retiring += nfin;
                }
            }
        }
        if (ns) {
            if (!adm_rec.empty()) std::memcpy(R.up_h + L.adm, adm_rec.data(), adm_rec.size() * sizeof(uint4));
            for (size_t k = 0; k < adm.size(); ++k)
                if (!adm[k].words.empty())
                    std::memcpy(R.up_h + L.words + (size_t)adm_rec[k].x * sizeof(uint32_t), adm[k].words.data(),
                                adm[k].words.size() * sizeof(uint32_t));
            hipError_t e = hipSuccess;
            if (!ensure_pinned(&R.z_h, &R.z_cap, (size_t)zsum + 1))
                e = hipErrorOutOfMemory;
            if (e == hipSuccess) e = hipMemcpyAsync(s->up_d, R.up_h, L.end, hipMemcpyHostToDevice, s->st);
            if (e == hipSuccess)
                e = ez::launch_foldin_admit((const uint4 *)(s->up_d + L.adm), (int)adm_rec.size(),
                                            (const uint32_t *)(s->up_d + L.words), s->wpool, s->st);
            if (e == hipSuccess)
                e = ez::launch_foldin_step(s->fb, (const uint4 *)s->up_d, (const uint2 *)(s->up_d + L.fin), (int)ns,
                                           ntok, s->V, s->K, s->alpha, s->phi, s->qs, s->qacc, s->wpool, s->zpool,
                                           s->seed, s->st);
            if (e == hipSuccess && zsum)
                e = hipMemcpyAsync(R.z_h, s->fb.zout, (size_t)zsum * sizeof(uint16_t), hipMemcpyDeviceToHost, s->st);
            if (e == hipSuccess) e = hipEventRecord(R.ev, s->st);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->steps++;
                if (e != hipSuccess) {
                    s->worker_rc = EZLDA_EHIP;
                    s->err = std::string("fold-in step failed: ") + hipGetErrorString(e);
                    s->cv_done.notify_all();
                    return;
                }
            }
            R.busy = true;
        }
        if (P.busy && !publish(s, P)) return;
        if (ns) cur ^= 1;
    }
}

// Device buffers of one fold-in batch: `tokens` pool tokens, `nslots` slots.
bool alloc_batch(ezlda_server *s, ez::FoldinBatch *b, uint32_t tokens, int nslots, int grid) {
    const size_t nb = ((size_t)s->V + 1023) / 1024;
    b->grid = grid;
    return salloc(s, &b->rows, tokens) == EZLDA_OK && salloc(s, &b->row_nnz, nslots) == EZLDA_OK &&
           salloc(s, &b->wcount, s->V) == EZLDA_OK && salloc(s, &b->wcursor, s->V) == EZLDA_OK &&
           salloc(s, &b->bsum, 3 * nb) == EZLDA_OK && salloc(s, &b->tokw, tokens) == EZLDA_OK &&
           salloc(s, &b->items, ez::foldin_max_items(s->V, tokens)) == EZLDA_OK &&
           salloc(s, &b->ctr, 4) == EZLDA_OK && salloc(s, &b->zout, tokens) == EZLDA_OK;
}

}  // namespace

extern "C" {

ezlda_server *ezlda_server_create(ezlda_ctx *ctx, const ezlda_server_args *a) try {
    if (!ctx || !a || a->max_slots < 1 || a->max_slots > 65535 || a->sweeps < 1 ||
        a->hbm_budget_bytes < EZLDA_FOLDIN_TOKEN_BYTES) {
        ezc::set_error("bad fold-in server arguments (max_slots 1..65535, sweeps >= 1, budget >= one token)");
        return nullptr;
    }
    if (ezc::refresh_counts(ctx) != EZLDA_OK) return nullptr;
    auto *s = new ezlda_server();
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || cus < 1)
        cus = 1;
    s->K = ctx->s.K;
    s->V = ctx->s.V;
    s->NS = (s->K + 63) / 64;
    s->alpha = ctx->s.alpha;
    s->seed = a->seed;
    s->sweeps = a->sweeps;
    s->max_slots = a->max_slots;
    s->rank_of = ctx->rank_of;
    // the pool holds the budget's tokens plus 8 per slot, the most that
    // rounding every resident document up to a multiple of 8 can add
    const uint64_t budget = std::min<uint64_t>((uint64_t)a->hbm_budget_bytes / EZLDA_FOLDIN_TOKEN_BYTES, 0xF0000000ull);
    s->budget_tokens = (uint32_t)budget;
    s->pool_tokens = round8(s->budget_tokens) + 8u * (uint32_t)s->max_slots;
    s->up_cap = std::min<uint32_t>(s->pool_tokens, 1u << 22);
    s->up_bytes = up_layout((size_t)s->max_slots, (size_t)s->max_slots, s->up_cap).end;
    bool ok = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->pst, hipStreamNonBlocking) == hipSuccess;
    ok = ok && salloc(s, &s->phi, (size_t)s->V * s->K) == EZLDA_OK &&
         salloc(s, &s->qs, (size_t)s->V * s->K) == EZLDA_OK && salloc(s, &s->qacc, (size_t)s->V * s->NS) == EZLDA_OK &&
         salloc(s, &s->wpool, s->pool_tokens) == EZLDA_OK && salloc(s, &s->zpool, s->pool_tokens) == EZLDA_OK &&
         salloc(s, &s->up_d, s->up_bytes) == EZLDA_OK &&
         alloc_batch(s, &s->fb, s->pool_tokens, s->max_slots, 8 * cus) &&
         salloc(s, &s->pw, 65536) == EZLDA_OK && salloc(s, &s->pz, 65536) == EZLDA_OK;
    for (StepRec &R : s->rec)
        ok = ok && hipHostMalloc((void **)&R.up_h, s->up_bytes, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&R.ev, hipEventDisableTiming) == hipSuccess;
    ok = ok && ez::launch_foldin_tables(ctx->s, s->phi, s->qs, s->qacc, s->st) == hipSuccess &&
         hipStreamSynchronize(s->st) == hipSuccess;
    if (!ok) {
        ezc::set_error(s->err.empty() ? "fold-in server set-up failed" : s->err);
        ezlda_server_destroy(s);
        return nullptr;
    }
    s->slots.resize(s->max_slots);
    for (int j = s->max_slots - 1; j >= 0; --j) s->free_slots.push_back(j);
    s->free_ranges[0] = s->pool_tokens;
    s->worker = std::thread(worker_main, s);
    return s;
} catch (...) {
    (void)ezc::exception_status(ctx);
    return nullptr;
}

void ezlda_server_destroy(ezlda_server *s) {
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->pending.clear();
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    // release: nothing to do about a failure here
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->pst) (void)hipStreamSynchronize(s->pst);
    for (void *p : s->allocs) (void)hipFree(p);
    for (StepRec &R : s->rec) {
        if (R.up_h) (void)hipHostFree(R.up_h);
        if (R.z_h) (void)hipHostFree(R.z_h);
        if (R.ev) (void)hipEventDestroy(R.ev);
    }
    if (s->st) (void)hipStreamDestroy(s->st);
    if (s->pst) (void)hipStreamDestroy(s->pst);
    delete s;
}

const char *ezlda_server_last_error(const ezlda_server *s) { return s ? s->err.c_str() : ezlda_last_error(nullptr); }

int ezlda_server_submit(ezlda_server *s, uint32_t request_id, const int32_t *words, int32_t n) try {
    if (!s || n < 0 || (n > 0 && !words)) return sfail(s, EZLDA_EINVAL, "bad fold-in request");
    if (n > 65535) return sfail(s, EZLDA_EINVAL, "a fold-in document has more than 65535 tokens");
    Request rq;
    rq.rid = request_id;
    rq.words.resize(n);
    for (int32_t i = 0; i < n; ++i) {
        if (words[i] < 0 || words[i] >= s->V) return sfail(s, EZLDA_EINVAL, "word id out of range in fold-in request");
        rq.words[i] = (uint32_t)s->rank_of[words[i]];
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if ((uint32_t)n > s->budget_tokens) {
        s->rejected++;
        return sfail(s, EZLDA_ENOMEM, "document larger than the server's HBM token budget");
    }
    if (s->worker_rc != EZLDA_OK) return sfail(s, s->worker_rc, s->err);
    if (s->done.count(request_id)) return sfail(s, EZLDA_EINVAL, "request id already has an unclaimed result");
    if (s->inflight.count(request_id)) return sfail(s, EZLDA_EINVAL, "request id already in flight");
    s->inflight.insert(request_id);
    rq.t_submit = now_ms();
    if (s->t_first < 0) s->t_first = rq.t_submit;
    s->submitted++;
    s->pending.push_back(std::move(rq));
    s->cv_work.notify_one();
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_server_submit_many(ezlda_server *s, int32_t n_docs, const uint32_t *request_ids, const int64_t *doc_off,
                             const int32_t *words) try {
    if (!s || n_docs < 0 || (n_docs > 0 && (!request_ids || !doc_off)) || (n_docs > 0 && doc_off[0] != 0))
        return sfail(s, EZLDA_EINVAL, "bad fold-in batch");
    std::vector<Request> rq((size_t)n_docs);
    for (int32_t d = 0; d < n_docs; ++d) {  // validate and map everything before queuing anything
        const int64_t n = doc_off[d + 1] - doc_off[d];
        if (n < 0) return sfail(s, EZLDA_EINVAL, "doc_off must be non-decreasing");
        if (n > 65535) return sfail(s, EZLDA_EINVAL, "a fold-in document has more than 65535 tokens");
        if (n > 0 && !words) return sfail(s, EZLDA_EINVAL, "bad fold-in batch");
        rq[d].rid = request_ids[d];
        rq[d].words.resize((size_t)n);
        const int32_t *w = words + doc_off[d];
        for (int64_t i = 0; i < n; ++i) {
            if (w[i] < 0 || w[i] >= s->V) return sfail(s, EZLDA_EINVAL, "word id out of range in fold-in request");
            rq[d].words[i] = (uint32_t)s->rank_of[w[i]];
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->worker_rc != EZLDA_OK) return sfail(s, s->worker_rc, s->err);
    std::unordered_set<uint32_t> ids;
    for (const Request &r : rq) {
        if (r.words.size() > s->budget_tokens) {
            s->rejected++;
            return sfail(s, EZLDA_ENOMEM, "document larger than the server's HBM token budget");
        }
        if (s->done.count(r.rid)) return sfail(s, EZLDA_EINVAL, "request id already has an unclaimed result");
        if (s->inflight.count(r.rid) || !ids.insert(r.rid).second)
            return sfail(s, EZLDA_EINVAL, "request id already in flight");
    }
    const double t = now_ms();
    if (s->t_first < 0 && n_docs > 0) s->t_first = t;
    for (Request &r : rq) {
        s->inflight.insert(r.rid);
        r.t_submit = t;
        s->pending.push_back(std::move(r));
    }
    s->submitted += n_docs;
    s->cv_work.notify_one();
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_server_wait(ezlda_server *s, uint32_t request_id, uint16_t *z_out, int32_t *counts_out, double *latency_ms,
                      int32_t timeout_ms) try {
    if (!s) return sfail(nullptr, EZLDA_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->done.count(request_id) && !s->inflight.count(request_id))
        return sfail(s, EZLDA_EINVAL, "unknown fold-in request " + std::to_string(request_id));
    auto ready = [&] { return s->done.count(request_id) || s->worker_rc != EZLDA_OK; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return sfail(s, EZLDA_ESTATE, "timed out waiting for fold-in request " + std::to_string(request_id));
    auto it = s->done.find(request_id);
    if (it == s->done.end()) return sfail(s, s->worker_rc, s->err);
    if (z_out && !it->second.z.empty()) std::memcpy(z_out, it->second.z.data(), it->second.z.size() * sizeof(uint16_t));
    if (counts_out) std::memcpy(counts_out, it->second.counts.data(), it->second.counts.size() * sizeof(int32_t));
    if (latency_ms) *latency_ms = it->second.latency_ms;
    s->done.erase(it);
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_server_drain(ezlda_server *s) try {
    if (!s) return sfail(nullptr, EZLDA_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (s->worker_rc != EZLDA_OK) return true;
        return s->pending.empty() && s->active.empty() && s->retiring == 0;
    });
    return s->worker_rc;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_server_get_stats(ezlda_server *s, ezlda_server_stats *st) try {
    if (!s || !st) return sfail(s, EZLDA_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->steps = s->steps;
    st->max_active = s->max_active;
    int act = 0;
    act = (int)s->active.size();
    st->active = act;
    st->pending = (int32_t)s->pending.size();
    if (!s->lat.empty()) {
        std::vector<double> v = s->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5))]; };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = v.back();
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_ms = sum / (double)v.size();
        if (s->t_last > s->t_first) st->docs_per_s = 1e3 * (double)s->completed / (s->t_last - s->t_first);
    }
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_server_reset_stats(ezlda_server *s) try {
    if (!s) return sfail(nullptr, EZLDA_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->submitted = s->completed = s->rejected = s->steps = 0;
    s->max_active = 0;
    s->t_first = -1.0;
    s->t_last = 0.0;
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

int ezlda_foldin_plain(ezlda_server *s, uint32_t request_id, const int32_t *words, int32_t n, uint16_t *z_out,
                       int32_t *counts_out) try {
    if (!s || n < 0 || (n > 0 && !words)) return sfail(s, EZLDA_EINVAL, "bad fold-in request");
    if (n > 65535) return sfail(s, EZLDA_EINVAL, "a fold-in document has more than 65535 tokens");
    std::vector<uint32_t> r(n);
    for (int32_t i = 0; i < n; ++i) {
        if (words[i] < 0 || words[i] >= s->V) return sfail(s, EZLDA_EINVAL, "word id out of range in fold-in request");
        r[i] = (uint32_t)s->rank_of[words[i]];
    }
    std::lock_guard<std::mutex> lk(s->plain_mu);
    // one document, one launch: every sweep runs inside k_fi_doc
    if (n) SV_HIP(s, hipMemcpyAsync(s->pw, r.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s->pst));
    SV_HIP(s, ez::launch_foldin_doc(s->pw, (uint32_t)n, request_id, s->sweeps, s->K, s->alpha, s->phi, s->qs, s->qacc,
                                    s->seed, s->pz, s->pst));
    std::vector<uint16_t> z((size_t)n);
    if (n) SV_HIP(s, hipMemcpyAsync(z.data(), s->pz, n * sizeof(uint16_t), hipMemcpyDeviceToHost, s->pst));
    SV_HIP(s, hipStreamSynchronize(s->pst));
    if (n && z_out) std::memcpy(z_out, z.data(), n * sizeof(uint16_t));
    if (counts_out) {
        std::memset(counts_out, 0, sizeof(int32_t) * (size_t)s->K);
        for (const uint16_t t : z) counts_out[t]++;
    }
    return EZLDA_OK;
} catch (...) {
    return ezc::server_exception_status(s);
}

}  // extern "C"
