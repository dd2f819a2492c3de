// Fold-in serving: topic mixtures of incoming documents against a frozen
// model, with continuous batching (no counterpart in the reference, which
// only trains).
//
// The server freezes phi from a trained context (its counts at creation) and
// keeps it on the device with the per-word Q running sums.  Requests carry a
// document; a worker thread admits them into device slots while slots and
// the HBM token pool (the budget) allow, and advances every resident request
// by one fold-in sweep per launch (k_foldin_step), so requests that arrive
// while others run join the next launch.  A request leaves after `sweeps`
// sweeps; its pool range and slot are reused.  Draws are keyed by (seed,
// request id, token, sweep), never by slot or position, so a result is the
// same alone and in any mix.  ezlda_foldin_plain runs one document per launch
// on its own stream, the comparison path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
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

}  // namespace

struct ezlda_server {
    int K = 0, V = 0, NS = 0, sweeps = 20, max_slots = 0;
    float alpha = 0.f;
    uint64_t seed = 0;
    std::vector<int32_t> rank_of;
    // device: frozen model, pools, slot table, per-slot counts
    float *phi = nullptr, *qs = nullptr, *qacc = nullptr;
    uint32_t *wpool = nullptr;
    uint16_t *zpool = nullptr;
    int32_t *counts = nullptr;
    uint4 *dslots = nullptr;
    uint4 *hslots = nullptr;  // pinned
    uint32_t pool_tokens = 0;
    // plain path buffers
    uint32_t *pw = nullptr;
    uint16_t *pz = nullptr;
    int32_t *pc = nullptr;
    uint4 *pslot = nullptr;
    hipStream_t st = nullptr, pst = nullptr;
    std::vector<void *> allocs;
    // token pool allocator: free ranges offset -> length
    std::map<uint32_t, uint32_t> free_ranges;
    std::vector<Slot> slots;
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

// Worker: admit, step, retire, until stopped.
void worker_main(ezlda_server *s) {
    if (hipSetDevice(s->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_rc = EZLDA_EHIP;
        s->err = "fold-in worker: hipSetDevice failed";
        s->cv_done.notify_all();
        return;
    }
    std::vector<int> act;
    std::vector<uint16_t> zbuf;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            // admission: free slot and pool room, in arrival order
            while (!s->pending.empty()) {
                auto fs = std::find_if(s->slots.begin(), s->slots.end(), [](const Slot &x) { return !x.active; });
                if (fs == s->slots.end()) break;
                uint32_t off = 0;
                const uint32_t n = (uint32_t)s->pending.front().words.size();
                if (!pool_alloc(s, n, &off)) break;
                Request rq = std::move(s->pending.front());
                s->pending.pop_front();
                fs->active = true;
                fs->rid = rq.rid;
                fs->off = off;
                fs->n = n;
                fs->sweep = 0;
                fs->t_submit = rq.t_submit;
                // the pageable source must stay valid until the copy is done
                if (n && (hipMemcpyAsync(s->wpool + off, rq.words.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                         s->st) != hipSuccess ||
                          hipStreamSynchronize(s->st) != hipSuccess)) {
                    s->worker_rc = EZLDA_EHIP;
                    s->err = "fold-in upload failed";
                }
            }
            act.clear();
            for (int i = 0; i < s->max_slots; ++i)
                if (s->slots[i].active) act.push_back(i);
            s->max_active = std::max(s->max_active, (int)act.size());
            if (act.empty()) {
                if (s->stop) return;
                s->cv_work.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
        }
        for (size_t j = 0; j < act.size(); ++j) {
            const Slot &x = s->slots[act[j]];
            s->hslots[j] = make_uint4(x.rid, x.off, x.n, x.sweep);
        }
        const int ns = (int)act.size();
        hipError_t e = hipMemcpyAsync(s->dslots, s->hslots, ns * sizeof(uint4), hipMemcpyHostToDevice, s->st);
        if (e == hipSuccess)
            e = ez::launch_foldin_step(s->dslots, ns, s->K, s->sweeps, s->alpha, s->phi, s->qs, s->qacc, s->wpool,
                                       s->zpool, s->counts, s->seed, s->st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->st);
        std::unique_lock<std::mutex> lk(s->mu);
        s->steps++;
        if (e != hipSuccess) {
            s->worker_rc = EZLDA_EHIP;
            s->err = std::string("fold-in step failed: ") + hipGetErrorString(e);
            for (int i : act) s->slots[i].active = false;
            s->cv_done.notify_all();
            return;
        }
        bool any = false;
        for (size_t j = 0; j < act.size(); ++j) {
            Slot &x = s->slots[act[j]];
            if (++x.sweep < (uint32_t)s->sweeps) continue;
            Result r;
            r.z.resize(x.n);
            r.counts.resize(s->K);
            lk.unlock();
            hipError_t ce = hipSuccess;
            if (x.n) ce = hipMemcpy(r.z.data(), s->zpool + x.off, x.n * sizeof(uint16_t), hipMemcpyDeviceToHost);
            if (ce == hipSuccess)
                ce = hipMemcpy(r.counts.data(), s->counts + j * (size_t)s->K, s->K * sizeof(int32_t),
                               hipMemcpyDeviceToHost);
            lk.lock();
            if (ce != hipSuccess) {
                s->worker_rc = EZLDA_EHIP;
                s->err = std::string("fold-in result copy failed: ") + hipGetErrorString(ce);
            }
            const double t = now_ms();
            r.latency_ms = t - x.t_submit;
            s->lat.push_back(r.latency_ms);
            s->completed++;
            s->t_last = t;
            s->done[x.rid] = std::move(r);
            s->inflight.erase(x.rid);
            pool_free(s, x.off, x.n);
            x.active = false;
            any = true;
        }
        if (any) s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

ezlda_server *ezlda_server_create(ezlda_ctx *ctx, const ezlda_server_args *a) {
    if (!ctx || !a || a->max_slots < 1 || a->sweeps < 1 || a->hbm_budget_bytes < 6) {
        ezc::set_error("bad fold-in server arguments");
        return nullptr;
    }
    if (ezc::refresh_counts(ctx) != EZLDA_OK) return nullptr;
    auto *s = new ezlda_server();
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    s->K = ctx->s.K;
    s->V = ctx->s.V;
    s->NS = (s->K + 63) / 64;
    s->alpha = ctx->s.alpha;
    s->seed = a->seed;
    s->sweeps = a->sweeps;
    s->max_slots = a->max_slots;
    s->rank_of = ctx->rank_of;
    const uint64_t tok = std::min<uint64_t>((uint64_t)a->hbm_budget_bytes / 6u, 0xFFFFFFF0ull);
    s->pool_tokens = (uint32_t)tok;
    bool ok = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->pst, hipStreamNonBlocking) == hipSuccess;
    ok = ok && salloc(s, &s->phi, (size_t)s->V * s->K) == EZLDA_OK &&
         salloc(s, &s->qs, (size_t)s->V * s->K) == EZLDA_OK && salloc(s, &s->qacc, (size_t)s->V * s->NS) == EZLDA_OK &&
         salloc(s, &s->wpool, s->pool_tokens) == EZLDA_OK && salloc(s, &s->zpool, s->pool_tokens) == EZLDA_OK &&
         salloc(s, &s->counts, (size_t)s->max_slots * s->K) == EZLDA_OK &&
         salloc(s, &s->dslots, s->max_slots) == EZLDA_OK && salloc(s, &s->pw, 65536) == EZLDA_OK &&
         salloc(s, &s->pz, 65536) == EZLDA_OK && salloc(s, &s->pc, s->K) == EZLDA_OK &&
         salloc(s, &s->pslot, 1) == EZLDA_OK;
    ok = ok && hipHostMalloc((void **)&s->hslots, s->max_slots * sizeof(uint4), hipHostMallocDefault) == hipSuccess;
    ok = ok && ez::launch_foldin_tables(ctx->s, s->phi, s->qs, s->qacc, s->st) == hipSuccess &&
         hipStreamSynchronize(s->st) == hipSuccess;
    if (!ok) {
        ezc::set_error(s->err.empty() ? "fold-in server set-up failed" : s->err);
        ezlda_server_destroy(s);
        return nullptr;
    }
    s->slots.resize(s->max_slots);
    s->free_ranges[0] = s->pool_tokens;
    s->worker = std::thread(worker_main, s);
    return s;
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
    if (s->hslots) (void)hipHostFree(s->hslots);
    if (s->st) (void)hipStreamDestroy(s->st);
    if (s->pst) (void)hipStreamDestroy(s->pst);
    delete s;
}

const char *ezlda_server_last_error(const ezlda_server *s) { return s ? s->err.c_str() : ezlda_last_error(nullptr); }

int ezlda_server_submit(ezlda_server *s, uint32_t request_id, const int32_t *words, int32_t n) {
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
    if ((uint32_t)n > s->pool_tokens) {
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
}

int ezlda_server_wait(ezlda_server *s, uint32_t request_id, uint16_t *z_out, int32_t *counts_out, double *latency_ms,
                      int32_t timeout_ms) {
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
}

int ezlda_server_drain(ezlda_server *s) {
    if (!s) return sfail(nullptr, EZLDA_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        if (s->worker_rc != EZLDA_OK) return true;
        if (!s->pending.empty()) return false;
        for (const Slot &x : s->slots)
            if (x.active) return false;
        return true;
    });
    return s->worker_rc;
}

int ezlda_server_get_stats(ezlda_server *s, ezlda_server_stats *st) {
    if (!s || !st) return sfail(s, EZLDA_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->steps = s->steps;
    st->max_active = s->max_active;
    int act = 0;
    for (const Slot &x : s->slots) act += x.active;
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
}

int ezlda_server_reset_stats(ezlda_server *s) {
    if (!s) return sfail(nullptr, EZLDA_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->submitted = s->completed = s->rejected = s->steps = 0;
    s->max_active = 0;
    s->t_first = -1.0;
    s->t_last = 0.0;
    return EZLDA_OK;
}

int ezlda_foldin_plain(ezlda_server *s, uint32_t request_id, const int32_t *words, int32_t n, uint16_t *z_out,
                       int32_t *counts_out) {
    if (!s || n < 0 || (n > 0 && !words)) return sfail(s, EZLDA_EINVAL, "bad fold-in request");
    if (n > 65535) return sfail(s, EZLDA_EINVAL, "a fold-in document has more than 65535 tokens");
    std::vector<uint32_t> r(n);
    for (int32_t i = 0; i < n; ++i) {
        if (words[i] < 0 || words[i] >= s->V) return sfail(s, EZLDA_EINVAL, "word id out of range in fold-in request");
        r[i] = (uint32_t)s->rank_of[words[i]];
    }
    std::lock_guard<std::mutex> lk(s->plain_mu);
    if (n) SV_HIP(s, hipMemcpyAsync(s->pw, r.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s->pst));
    for (int sw = 0; sw < s->sweeps; ++sw) {
        const uint4 sl = make_uint4(request_id, 0u, (uint32_t)n, (uint32_t)sw);
        SV_HIP(s, hipMemcpyAsync(s->pslot, &sl, sizeof(uint4), hipMemcpyHostToDevice, s->pst));
        SV_HIP(s, ez::launch_foldin_step(s->pslot, 1, s->K, s->sweeps, s->alpha, s->phi, s->qs, s->qacc, s->pw, s->pz,
                                         s->pc, s->seed, s->pst));
        SV_HIP(s, hipStreamSynchronize(s->pst));
    }
    if (n && z_out) SV_HIP(s, hipMemcpy(z_out, s->pz, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    if (counts_out) SV_HIP(s, hipMemcpy(counts_out, s->pc, s->K * sizeof(int32_t), hipMemcpyDeviceToHost));
    return EZLDA_OK;
}

}  // extern "C"
