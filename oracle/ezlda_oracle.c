/*
 * ezlda_oracle.c -- CPU restatement of the EZLDA Gibbs sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU baseline.  The product (lda_sc20_amd/libezlda.so) never links,
 * loads or calls it.
 *
 * Built twice from this one source (oracle/Makefile):
 *   liboracle.so      single-threaded: the parity oracle;
 *   liboracle_omp.so  the same code with its OpenMP pragmas enabled: the
 *                     reported CPU baseline.  Every parallel loop writes
 *                     disjoint outputs and sums integers only, so both builds
 *                     return bit-identical results.
 * Float rules: -ffp-contract=off, IEEE division, no -ffast-math; every float
 * expression below is evaluated exactly as written, in f32.
 *
 * The sweep it defines (canonical semantics; the GPU path must match it bit
 * for bit):
 *   1. counts from z: n_wk, n_k, n_dk  (reference: EZLDA/utility.cu:10-76 WT,
 *      :1067-1126 dense head + n_k, :79-284 DT).  Every token of the sweep is
 *      sampled against these counts, as the reference does: its WT is only
 *      rebuilt between sweeps (EZLDA/mainKernel.cu:163-185, :355-356) and its
 *      DT for a chunk is built before the chunk is sampled (:228).
 *   2. inv_k = 1/(n_k + V*beta); phi_wk = (n_wk + beta) * inv_k
 *      (reference computes the same quotient, utility.cu:336,712,719,1259,1267).
 *   3. per word: top-3 topics by phi, ties -> lower topic
 *      (MaxTopic*_Update_Kernel, utility.cu:305-678, :679-1062).
 *   4. per token: n_dk1, n_dk2 of its doc at the word's k1, k2
 *      (DT_Update_Kernel re-pack, utility.cu:125-161).
 *   5. per token: three buckets in this order (utility.cu:1863-1992):
 *        P1 = (n_dk1 + alpha) phi_k1
 *        S  = sum_{k != k1} n_dk phi_wk   over the doc's sorted DT row
 *        Q  = alpha sum_{k != k1} phi_wk  over all topics
 *      u = Philox4x32-10(seed; token, sweep, 0).x >> 8 scaled by 2^-24
 *      (replaces curand_uniform, utility.cu:1353).  target = u * total.
 *      S is accumulated left to right over the DT row with fused
 *      multiply-adds (fmaf: C99, correctly rounded, as the device FMA); Q as
 *      running sums inside each 64-topic slab plus a running sum of the slab
 *      totals.  Running sums of non-negative terms never decrease, so the
 *      draw is the first element whose running sum exceeds the target (it
 *      always has positive mass); when rounding leaves no running sum above
 *      the target, the last element of positive mass wins.
 *   6. max-topic shortcut (UpdateProbKernelTrainD1, utility.cu:2680-3312;
 *      inline tail test utility.cu:1341-1373): accept k1 when
 *      u * (P1 + maxS + Q) < P1 * (1 - 2^-9), maxS = n_dk2 phi_k2 +
 *      (n_d - n_dk1 - n_dk2) phi_k3 >= S.  The full path's target carries at
 *      most (nnz + 4) roundings of 2^-24 (nnz <= K <= 16383), below the
 *      2^-9 margin, so the shortcut only ever accepts tokens the full path
 *      also sends to k1: shortcut on and off give identical z.  The same
 *      holds for any bound >= S; shortcut mode 2 uses the GPU path's
 *      maxS = (n_d - n_dk1) phi_k2 (every topic but k1 has phi <= phi_k2),
 *      which needs no n_dk2: same z, more full draws.
 *   7. log-likelihood in FP64: sum_t log(sum_k theta_dk phi_wk), the per-token
 *      term of profilePerp/utility.cu:2398,2810 (log((S+Q)/(n_d+K alpha))),
 *      summed over all tokens (profilePerp/utility.cu:2858-2928).
 *
 * Token ids: tokens are numbered in word-major order -- words by descending
 * count, ties by ascending word id (preprocess/streamTestBalance.cpp:281-301
 * ranks words the same way); within a word, docs by descending length (ties
 * by ascending doc id), then position in the doc.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MARGIN 0.998046875f /* 1 - 2^-9 */
#define ORC_NONE 0xFFFFu
/* Compact DT rows (K <= ORC_COMPACT_K): 16-bit entries of tb topic bits,
 * tb = max(10, ceil(log2 K)), and 16 - tb count bits; a topic's count is
 * stored in pieces of at most piece = 2^(16 - tb) - 1, one entry per piece,
 * in order, and every S sum and running sum walks the pieces.  Larger K
 * keeps one entry per topic. */
#define ORC_COMPACT_K 8192

static int topic_bits(int K) {
    int b = 10;
    while ((1 << b) < K) ++b;
    return b;
}
static uint32_t dt_piece(int K) { return K <= ORC_COMPACT_K ? (1u << (16 - topic_bits(K))) - 1u : 0xFFFFFFFFu; }

/* Entries of a doc's DT row at most: min(n, K), plus one per full piece in
 * compact rows. */
static int64_t dt_cap(int64_t n, int K) {
    const int64_t m = n < K ? n : K;
    if (K > ORC_COMPACT_K) return m;
    const int64_t c = m + n / (int64_t)dt_piece(K);
    return c < n ? c : n;
}

/* ---------------------------------------------------------------- Philox */
/* Salmon et al., SC'11, Philox4x32 with 10 rounds (Random123 constants). */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static uint32_t philox_x0(uint64_t seed, uint64_t tok, uint32_t sweep, uint32_t tag) {
    uint32_t c[4] = {(uint32_t)tok, (uint32_t)(tok >> 32), sweep, tag};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox(c, k, o);
    return o[0];
}

static float u24(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

/* ------------------------------------------------------------ the state */
typedef struct {
    int K, V, D;
    int64_t N;
    float alpha, beta, vbeta;
    uint64_t seed;
    int shortcut;       /* 0 off, 1 the reference's bound, 2 the GPU path's looser bound */
    int64_t n_full;     /* tokens that took the full draw, all sweeps */
    uint32_t sweep;
    int64_t *word_off;  /* V+1, word-major token ranges (rank order) */
    int32_t *tok_doc;   /* N */
    int64_t *doc_off;   /* D+1, doc-major ranges */
    int64_t *doc_tok;   /* N: doc-major position -> token id */
    int32_t *doc_wrd;   /* N: doc-major position -> word rank */
    int32_t *rank_of;   /* V: input word id -> rank */
    uint16_t *z;        /* N */
    /* counts */
    int32_t *nwk;       /* V*K */
    int32_t *nk;        /* K */
    int64_t *dt_off;    /* D+1, capacity min(n_d,K) */
    uint32_t *dt;       /* sorted (topic | count<<16) */
    int32_t *dt_nnz;    /* D */
    float *inv;         /* K */
    uint16_t *top;      /* V*3 */
    uint32_t *aux;      /* N: n_dk1 | n_dk2<<16 */
    int counts_fresh;
} orc_t;

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s ? s : 1); }

/* Word-major layout of a doc-major corpus (the token numbering above). */
void *orc_create(int K, float alpha, float beta, uint64_t seed, int V, int D,
                 const int64_t *doc_off, const int32_t *words, const uint16_t *z0,
                 int shortcut) {
    orc_t *o = (orc_t *)xcalloc(1, sizeof(orc_t));
    o->K = K; o->V = V; o->D = D;
    o->alpha = alpha; o->beta = beta;
    o->vbeta = (float)((double)V * (double)beta);
    o->seed = seed; o->shortcut = shortcut;
    int64_t N = doc_off[D];
    o->N = N;
    int64_t *cnt = (int64_t *)xcalloc(V, sizeof(int64_t));
    for (int64_t i = 0; i < N; ++i) cnt[words[i]]++;
    /* rank: count desc, id asc (stable selection via counting is overkill; a
     * simple insertion-free approach: sort ids with qsort on (count, id)). */
    int32_t *ids = (int32_t *)xcalloc(V, sizeof(int32_t));
    for (int w = 0; w < V; ++w) ids[w] = w;
    /* shell sort by (count desc, id asc) */
    for (int gap = V / 2; gap > 0; gap /= 2)
        for (int i = gap; i < V; ++i) {
            int32_t x = ids[i];
            int j = i;
            while (j >= gap) {
                int32_t y = ids[j - gap];
                int before = cnt[y] > cnt[x] || (cnt[y] == cnt[x] && y < x);
                if (before) break;
                ids[j] = y;
                j -= gap;
            }
            ids[j] = x;
        }
    o->rank_of = (int32_t *)xcalloc(V, sizeof(int32_t));
    for (int r = 0; r < V; ++r) o->rank_of[ids[r]] = r;
    o->word_off = (int64_t *)xcalloc(V + 1, sizeof(int64_t));
    for (int r = 0; r < V; ++r) o->word_off[r + 1] = o->word_off[r] + cnt[ids[r]];
    o->doc_off = (int64_t *)xcalloc(D + 1, sizeof(int64_t));
    memcpy(o->doc_off, doc_off, sizeof(int64_t) * (D + 1));
    o->tok_doc = (int32_t *)xcalloc(N, sizeof(int32_t));
    o->doc_tok = (int64_t *)xcalloc(N, sizeof(int64_t));
    o->doc_wrd = (int32_t *)xcalloc(N, sizeof(int32_t));
    o->z = (uint16_t *)xcalloc(N, sizeof(uint16_t));
    /* docs by length desc, id asc: tokens of a word come doc by doc in this
     * order (so neighbouring tokens of a word have similar DT row lengths) */
    int32_t *dord = (int32_t *)xcalloc(D, sizeof(int32_t));
    for (int d = 0; d < D; ++d) dord[d] = d;
    for (int gap = D / 2; gap > 0; gap /= 2)
        for (int i = gap; i < D; ++i) {
            int32_t x = dord[i];
            int64_t lx = doc_off[x + 1] - doc_off[x];
            int j = i;
            while (j >= gap) {
                int32_t y = dord[j - gap];
                int64_t ly = doc_off[y + 1] - doc_off[y];
                if (ly > lx || (ly == lx && y < x)) break;
                dord[j] = y;
                j -= gap;
            }
            dord[j] = x;
        }
    int64_t *cur = (int64_t *)xcalloc(V, sizeof(int64_t));
    for (int r = 0; r < V; ++r) cur[r] = o->word_off[r];
    for (int q = 0; q < D; ++q)
        for (int64_t i = doc_off[dord[q]]; i < doc_off[dord[q] + 1]; ++i) {
            const int d = dord[q];
            int r = o->rank_of[words[i]];
            int64_t t = cur[r]++;
            o->tok_doc[t] = d;
            o->doc_tok[i] = t;
            o->doc_wrd[i] = r;
            if (z0) o->z[t] = z0[i];
        }
    if (!z0)
        for (int64_t t = 0; t < N; ++t)
            o->z[t] = (uint16_t)(philox_x0(seed, (uint64_t)t, 0, 1u) % (uint32_t)K);
    free(cur); free(ids); free(cnt); free(dord);
    o->nwk = (int32_t *)xcalloc((size_t)V * K, sizeof(int32_t));
    o->nk = (int32_t *)xcalloc(K, sizeof(int32_t));
    o->dt_off = (int64_t *)xcalloc(D + 1, sizeof(int64_t));
    for (int d = 0; d < D; ++d) {
        int64_t n = doc_off[d + 1] - doc_off[d];
        o->dt_off[d + 1] = o->dt_off[d] + dt_cap(n, K);
    }
    o->dt = (uint32_t *)xcalloc(o->dt_off[D], sizeof(uint32_t));
    o->dt_nnz = (int32_t *)xcalloc(D, sizeof(int32_t));
    o->inv = (float *)xcalloc(K, sizeof(float));
    o->top = (uint16_t *)xcalloc((size_t)V * 3, sizeof(uint16_t));
    o->aux = (uint32_t *)xcalloc(N, sizeof(uint32_t));
    return o;
}

void orc_destroy(void *h) {
    orc_t *o = (orc_t *)h;
    if (!o) return;
    free(o->word_off); free(o->tok_doc); free(o->doc_off); free(o->doc_tok);
    free(o->doc_wrd); free(o->rank_of); free(o->z); free(o->nwk); free(o->nk);
    free(o->dt_off); free(o->dt); free(o->dt_nnz); free(o->inv); free(o->top);
    free(o->aux); free(o);
}

int64_t orc_num_tokens(void *h) { return ((orc_t *)h)->N; }
uint32_t orc_get_sweep(void *h) { return ((orc_t *)h)->sweep; }
int64_t orc_full_draws(void *h) { return ((orc_t *)h)->n_full; }
void orc_set_sweep(void *h, uint32_t s) { ((orc_t *)h)->sweep = s; }

/* --------------------------------------------- 1. WT and n_k from z */
/* utility.cu:10-76 (tail WT), :1067-1102 (head WT), :1104-1126 and
 * :1132-1204 (n_k): here one dense histogram per word. */
void orc_build_wt(void *h) {
    orc_t *o = (orc_t *)h;
    const int K = o->K, V = o->V;
    memset(o->nk, 0, sizeof(int32_t) * K);
#pragma omp parallel for schedule(dynamic, 64)
    for (int w = 0; w < V; ++w) {
        int32_t *row = o->nwk + (size_t)w * K;
        memset(row, 0, sizeof(int32_t) * K);
        for (int64_t t = o->word_off[w]; t < o->word_off[w + 1]; ++t) row[o->z[t]]++;
    }
    for (int w = 0; w < V; ++w) {
        const int32_t *row = o->nwk + (size_t)w * K;
        for (int k = 0; k < K; ++k) o->nk[k] += row[k];
    }
    o->counts_fresh = 1;
}

static int better(float va, int ka, float vb, int kb) {
    return va > vb || (va == vb && ka < kb);
}

static void phi_row(const orc_t *o, int w, float *phi) {
    const int32_t *row = o->nwk + (size_t)w * o->K;
    for (int k = 0; k < o->K; ++k) phi[k] = ((float)row[k] + o->beta) * o->inv[k];
}

/* ------------------------------------- 2-3. inv_k and per-word top-3 */
/* MaxTopicDense/Sparse_Update_Kernel, utility.cu:305-678, :679-1062: the
 * reference keeps the three largest phi with ties to the lower topic
 * (:375-641); topics beyond K (K < 3) are ORC_NONE. */
void orc_topk(void *h) {
    orc_t *o = (orc_t *)h;
    const int K = o->K;
    for (int k = 0; k < K; ++k) o->inv[k] = 1.0f / ((float)o->nk[k] + o->vbeta);
#pragma omp parallel
    {
        float *phi = (float *)xcalloc(K, sizeof(float));
#pragma omp for schedule(dynamic, 64)
        for (int w = 0; w < o->V; ++w) {
            phi_row(o, w, phi);
            float v[3] = {-1.f, -1.f, -1.f};
            int kk[3] = {ORC_NONE, ORC_NONE, ORC_NONE};
            for (int k = 0; k < K; ++k) {
                float x = phi[k];
                if (kk[2] != ORC_NONE && !better(x, k, v[2], kk[2])) continue;
                int pos = 2;
                while (pos > 0 && (kk[pos - 1] == ORC_NONE || better(x, k, v[pos - 1], kk[pos - 1]))) {
                    v[pos] = v[pos - 1];
                    kk[pos] = kk[pos - 1];
                    --pos;
                }
                v[pos] = x;
                kk[pos] = k;
            }
            for (int i = 0; i < 3; ++i) o->top[(size_t)w * 3 + i] = (uint16_t)kk[i];
        }
        free(phi);
    }
}

/* ---------------------------------------- 4. DT rows and per-token aux */
/* DT_Update_Kernel, utility.cu:79-284: histogram of the doc's topics,
 * per-token (n_dk1, n_dk2) re-pack (:125-161), ascending-topic compaction
 * (:183-238); for K <= ORC_COMPACT_K a count above dt_piece(K) becomes
 * pieces of dt_piece(K) and a remainder (the reference packs the count in 16
 * bits and never splits). */
void orc_build_dt(void *h) {
    orc_t *o = (orc_t *)h;
    const int K = o->K;
    const uint32_t piece = dt_piece(K);
#pragma omp parallel
    {
        int32_t *hist = (int32_t *)xcalloc(K, sizeof(int32_t));
#pragma omp for schedule(dynamic, 256)
        for (int d = 0; d < o->D; ++d) {
            for (int64_t i = o->doc_off[d]; i < o->doc_off[d + 1]; ++i) hist[o->z[o->doc_tok[i]]]++;
            for (int64_t i = o->doc_off[d]; i < o->doc_off[d + 1]; ++i) {
                const uint16_t *tp = o->top + (size_t)o->doc_wrd[i] * 3;
                uint32_t n1 = tp[0] != ORC_NONE ? (uint32_t)hist[tp[0]] : 0u;
                uint32_t n2 = tp[1] != ORC_NONE ? (uint32_t)hist[tp[1]] : 0u;
                o->aux[o->doc_tok[i]] = n1 | (n2 << 16);
            }
            int nz = 0;
            uint32_t *row = o->dt + o->dt_off[d];
            for (int k = 0; k < K; ++k)
                if (hist[k]) {
                    uint32_t c = (uint32_t)hist[k];
                    if (K <= ORC_COMPACT_K)
                        for (; c > piece; c -= piece) row[nz++] = (uint32_t)k | (piece << 16);
                    row[nz++] = (uint32_t)k | (c << 16);
                    hist[k] = 0;
                }
            o->dt_nnz[d] = nz;
        }
        free(hist);
    }
}

/* ------------------------------------------------------- 5-6. sampler */
/* Full three-bucket draw for one token (LDAKernelTrain(D), utility.cu:
 * 1217-1611, :1625-2103; bucket search :1878-1992).  phi has phi[k1] = 0.
 * S: the DT row (ascending topic) accumulated left to right with fused
 * multiply-adds, run = fma(n_dk, phi_k, run), one rounding per entry; the
 * draw takes the first entry whose running sum exceeds the target.
 * Q: qs[k] is the running sum of alpha*phi inside each 64-topic slab,
 * qacc[j] the running sum of the slab totals; the draw takes the first slab,
 * then the first topic in it, whose running sum exceeds the target.  All
 * running sums of non-negative terms are non-decreasing, so "first" is
 * well defined and can be found by bisection. */
static int draw_full(const orc_t *o, const float *phi, const float *qs, const float *qt, const float *qacc,
                     int nsq, int k1, float P1, float Q, float u, int d) {
    const uint32_t *row = o->dt + o->dt_off[d];
    const int nnz = o->dt_nnz[d];
    float S = 0.0f;
    for (int i = 0; i < nnz; ++i) S = fmaf((float)(row[i] >> 16), phi[row[i] & 0xFFFFu], S);
    const float A = P1 + S;
    const float total = A + Q;
    const float target = u * total;
    if (target < P1) return k1;
    if (target < A) {
        const float t2 = target - P1;
        float run = 0.0f;
        int last = -1;
        for (int i = 0; i < nnz; ++i) {
            const float c = (float)(row[i] >> 16), f = phi[row[i] & 0xFFFFu];
            run = fmaf(c, f, run);
            if (run > t2) return (int)(row[i] & 0xFFFFu);
            if (c > 0.0f && f > 0.0f) last = i;
        }
        return last >= 0 ? (int)(row[last] & 0xFFFFu) : k1;
    }
    const float t2 = target - A;
    int js = -1;
    for (int j = 0; j < nsq; ++j)
        if (qacc[j] > t2) { js = j; break; }
    if (js < 0)
        for (int j = nsq - 1; j >= 0; --j)
            if (qt[j] > 0.0f) { js = j; break; }
    if (js < 0) return k1;
    const float t3 = t2 - (js > 0 ? qacc[js - 1] : 0.0f);
    const int k0 = 64 * js, k_end = (k0 + 64 < o->K) ? k0 + 64 : o->K;
    for (int k = k0; k < k_end; ++k)
        if (qs[k] > t3) return k;
    for (int k = k_end - 1; k >= k0; --k)
        if (o->alpha * phi[k] > 0.0f) return k;
    return k1;
}

void orc_sample(void *h) {
    orc_t *o = (orc_t *)h;
    const int K = o->K;
    const int nsq = (K + 63) / 64;
    const float alpha = o->alpha;
#pragma omp parallel
    {
        float *phi = (float *)xcalloc(K, sizeof(float));
        float *qs = (float *)xcalloc(K, sizeof(float));
        float *qt = (float *)xcalloc(nsq, sizeof(float));
        float *qacc = (float *)xcalloc(nsq, sizeof(float));
        int64_t n_full = 0;
#pragma omp for schedule(dynamic, 16)
        for (int w = 0; w < o->V; ++w) {
            if (o->word_off[w + 1] == o->word_off[w]) continue;
            phi_row(o, w, phi);
            const uint16_t *tp = o->top + (size_t)w * 3;
            const int k1 = tp[0];
            const float f1 = phi[k1];
            const float f2 = tp[1] != ORC_NONE ? phi[tp[1]] : 0.0f;
            const float f3 = tp[2] != ORC_NONE ? phi[tp[2]] : 0.0f;
            phi[k1] = 0.0f;
            float acc = 0.0f;
            for (int j = 0; j < nsq; ++j) {
                float run = 0.0f;
                for (int k = 64 * j; k < 64 * j + 64 && k < K; ++k) {
                    run = run + alpha * phi[k];
                    qs[k] = run;
                }
                qt[j] = run;
                acc = acc + run;
                qacc[j] = acc;
            }
            const float Q = qacc[nsq - 1];
            for (int64_t t = o->word_off[w]; t < o->word_off[w + 1]; ++t) {
                const int d = o->tok_doc[t];
                const uint32_t a = o->aux[t];
                const uint32_t n1 = a & 0xFFFFu, n2 = a >> 16;
                const uint32_t nd = (uint32_t)(o->doc_off[d + 1] - o->doc_off[d]);
                const float u = u24(philox_x0(o->seed, (uint64_t)t, o->sweep, 0u));
                const float P1 = ((float)n1 + alpha) * f1;
                if (o->shortcut) {
                    const float maxS = o->shortcut == 2 ? (float)(nd - n1) * f2
                                                        : (float)n2 * f2 + (float)(nd - n1 - n2) * f3;
                    const float B = (P1 + maxS) + Q;
                    if (u * B < P1 * ORC_MARGIN) {
                        o->z[t] = (uint16_t)k1;
                        continue;
                    }
                }
                ++n_full;
                o->z[t] = (uint16_t)draw_full(o, phi, qs, qt, qacc, nsq, k1, P1, Q, u, d);
            }
        }
        free(phi); free(qs); free(qt); free(qacc);
#pragma omp atomic
        o->n_full += n_full;
    }
    o->sweep++;
    o->counts_fresh = 0;
}

/* One sweep: mainKernel.cu:158-380 in order (WT rebuild, top-k, DT, sample). */
void orc_sweep(void *h, int n) {
    orc_t *o = (orc_t *)h;
    for (int s = 0; s < n; ++s) {
        if (!o->counts_fresh) orc_build_wt(o);
        orc_topk(o);
        orc_build_dt(o);
        orc_sample(o);
    }
}

/* --------------------------------------------- 7. FP64 log-likelihood */
/* Per token log(sum_k (n_dk+alpha)/(n_d+K alpha) * (n_wk+beta)/(n_k+V beta))
 * (profilePerp/utility.cu:2398,2810), summed over every token of the current
 * z.  Returns LL; perplexity = exp(-LL/N). */
double orc_loglik(void *h) {
    orc_t *o = (orc_t *)h;
    if (!o->counts_fresh) orc_build_wt(o);
    orc_topk(o);
    orc_build_dt(o);
    const int K = o->K;
    const double a = (double)o->alpha, b = (double)o->beta;
    const double vb = (double)o->V * b;
    double *den = (double *)xcalloc(K, sizeof(double));
    for (int k = 0; k < K; ++k) den[k] = (double)o->nk[k] + vb;
    double ll = 0.0;
#pragma omp parallel
    {
        double *phi = (double *)xcalloc(K, sizeof(double));
#pragma omp for schedule(dynamic, 16) reduction(+ : ll)
        for (int w = 0; w < o->V; ++w) {
            if (o->word_off[w + 1] == o->word_off[w]) continue;
            const int32_t *row = o->nwk + (size_t)w * K;
            double q = 0.0;
            for (int k = 0; k < K; ++k) {
                phi[k] = ((double)row[k] + b) / den[k];
                q += phi[k];
            }
            q *= a;
            for (int64_t t = o->word_off[w]; t < o->word_off[w + 1]; ++t) {
                const int d = o->tok_doc[t];
                const uint32_t *dr = o->dt + o->dt_off[d];
                double s = 0.0;
                for (int i = 0; i < o->dt_nnz[d]; ++i) s += (double)(dr[i] >> 16) * phi[dr[i] & 0xFFFFu];
                const double nd = (double)(o->doc_off[d + 1] - o->doc_off[d]);
                ll += log((s + q) / (nd + (double)K * a));
            }
        }
        free(phi);
    }
    free(den);
    return ll;
}

/* ------------------------------------------------------------ fold-in */
/* Fold-in of new documents against the frozen model of the current counts
 * (the reference has no counterpart; the conditional is its sampler's with
 * WT frozen: p(k) ~ (n_dk + alpha) phi_wk).  Per request r (id rid) with
 * tokens w_0..w_{n-1} (input word ids):
 *   z_i = Philox(seed; rid<<32 | i, 0, 4).x mod K;
 *   each of `sweeps` Jacobi sweeps rebuilds the doc's DT row from z and
 *   draws every token from two buckets in this order:
 *     S = doc row (ascending topic) accumulated with fmaf(n_dk, phi_wk, S),
 *     Q = alpha * phi_w. running sums inside 64-topic slabs + running sum
 *         of the slab totals (no topic excluded),
 *   u = Philox(seed; rid<<32 | i, sweep, 3).x >> 8 scaled by 2^-24,
 *   target = u * (S + Q); S bucket if target < S (first entry whose running
 *   sum exceeds target), else Q bucket with target - S; the same fallbacks
 *   as training.
 * Outputs the final z and the doc's topic counts (K per doc). */
void orc_foldin(void *h, int nd, const int64_t *doc_off, const int32_t *words, const uint32_t *rids,
                int sweeps, uint64_t seed, uint16_t *z_out, int32_t *counts_out) {
    orc_t *o = (orc_t *)h;
    const int K = o->K, V = o->V;
    const int nsq = (K + 63) / 64;
    const float alpha = o->alpha;
    if (!o->counts_fresh) orc_build_wt(o);
    for (int k = 0; k < K; ++k) o->inv[k] = 1.0f / ((float)o->nk[k] + o->vbeta);
    float *phi = (float *)xcalloc((size_t)V * K, sizeof(float));
    float *qs = (float *)xcalloc((size_t)V * K, sizeof(float));
    float *qacc = (float *)xcalloc((size_t)V * nsq, sizeof(float));
    for (int w = 0; w < V; ++w) {
        float *ph = phi + (size_t)w * K;
        phi_row(o, w, ph);
        float acc = 0.0f;
        for (int j = 0; j < nsq; ++j) {
            float run = 0.0f;
            for (int k = 64 * j; k < 64 * j + 64 && k < K; ++k) {
                run = run + alpha * ph[k];
                qs[(size_t)w * K + k] = run;
            }
            acc = acc + run;
            qacc[(size_t)w * nsq + j] = acc;
        }
    }
    int32_t *hist = (int32_t *)xcalloc(K, sizeof(int32_t));
    uint32_t *row = (uint32_t *)xcalloc(K, sizeof(uint32_t));
    for (int r = 0; r < nd; ++r) {
        const int64_t b = doc_off[r], n = doc_off[r + 1] - doc_off[r];
        const uint64_t key = (uint64_t)rids[r] << 32;
        uint16_t *z = z_out + b;
        for (int64_t i = 0; i < n; ++i) z[i] = (uint16_t)(philox_x0(seed, key | (uint64_t)i, 0u, 4u) % (uint32_t)K);
        for (int s = 0; s < sweeps; ++s) {
            memset(hist, 0, sizeof(int32_t) * K);
            for (int64_t i = 0; i < n; ++i) hist[z[i]]++;
            int nnz = 0;
            for (int k = 0; k < K; ++k)
                if (hist[k]) row[nnz++] = (uint32_t)k | ((uint32_t)hist[k] << 16);
            for (int64_t i = 0; i < n; ++i) {
                const int w = o->rank_of[words[b + i]];
                const float *ph = phi + (size_t)w * K;
                const float *qw = qs + (size_t)w * K;
                const float *qa = qacc + (size_t)w * nsq;
                const float u = u24(philox_x0(seed, key | (uint64_t)i, (uint32_t)s, 3u));
                float S = 0.0f;
                for (int j = 0; j < nnz; ++j) S = fmaf((float)(row[j] >> 16), ph[row[j] & 0xFFFFu], S);
                const float target = u * (S + qa[nsq - 1]);
                int zz = -1;
                if (target < S) {
                    float run = 0.0f;
                    int last = -1;
                    for (int j = 0; j < nnz && zz < 0; ++j) {
                        const float c = (float)(row[j] >> 16), f = ph[row[j] & 0xFFFFu];
                        run = fmaf(c, f, run);
                        if (run > target) zz = (int)(row[j] & 0xFFFFu);
                        else if (c > 0.0f && f > 0.0f) last = (int)(row[j] & 0xFFFFu);
                    }
                    if (zz < 0) zz = last;
                } else {
                    const float t2 = target - S;
                    int js = -1;
                    for (int j = 0; j < nsq; ++j)
                        if (qa[j] > t2) { js = j; break; }
                    if (js < 0)
                        for (int j = nsq - 1; j >= 0; --j)
                            if (qa[j] > (j > 0 ? qa[j - 1] : 0.0f)) { js = j; break; }
                    if (js >= 0) {
                        const float t3 = t2 - (js > 0 ? qa[js - 1] : 0.0f);
                        const int k0 = 64 * js, ke = (k0 + 64 < K) ? k0 + 64 : K;
                        for (int k = k0; k < ke && zz < 0; ++k)
                            if (qw[k] > t3) zz = k;
                        for (int k = ke - 1; k >= k0 && zz < 0; --k)
                            if (alpha * ph[k] > 0.0f) zz = k;
                    }
                }
                /* Jacobi: the row was built before any token of this sweep moved */
                z_out[b + i] = (uint16_t)(zz < 0 ? 0 : zz);
            }
        }
        int32_t *cnt = counts_out + (size_t)r * K;
        memset(cnt, 0, sizeof(int32_t) * K);
        for (int64_t i = 0; i < n; ++i) cnt[z[i]]++;
    }
    free(hist); free(row); free(phi); free(qs); free(qacc);
}

/* ------------------------------------------------------------ getters */
void orc_get_topics(void *h, uint16_t *z_docmajor) {
    orc_t *o = (orc_t *)h;
    for (int64_t i = 0; i < o->N; ++i) z_docmajor[i] = o->z[o->doc_tok[i]];
}
void orc_get_topics_wordmajor(void *h, uint16_t *z) {
    orc_t *o = (orc_t *)h;
    memcpy(z, o->z, sizeof(uint16_t) * o->N);
}
void orc_set_topics(void *h, const uint16_t *z_docmajor) {
    orc_t *o = (orc_t *)h;
    for (int64_t i = 0; i < o->N; ++i) o->z[o->doc_tok[i]] = z_docmajor[i];
    o->counts_fresh = 0;
}
void orc_get_nk(void *h, int32_t *nk) {
    orc_t *o = (orc_t *)h;
    memcpy(nk, o->nk, sizeof(int32_t) * o->K);
}
/* Dense n_wk indexed by INPUT word id (not rank). */
void orc_get_wt_dense(void *h, int32_t *out) {
    orc_t *o = (orc_t *)h;
    for (int w = 0; w < o->V; ++w)
        memcpy(out + (size_t)w * o->K, o->nwk + (size_t)o->rank_of[w] * o->K, sizeof(int32_t) * o->K);
}
void orc_get_dt_dense(void *h, int32_t *out) {
    orc_t *o = (orc_t *)h;
    memset(out, 0, sizeof(int32_t) * (size_t)o->D * o->K);
    for (int d = 0; d < o->D; ++d)
        for (int i = 0; i < o->dt_nnz[d]; ++i) {
            uint32_t e = o->dt[o->dt_off[d] + i];
            out[(size_t)d * o->K + (e & 0xFFFFu)] += (int32_t)(e >> 16);
        }
}
/* Packed DT rows exactly as stored: out has dt_off[D] entries; nnz per doc. */
void orc_get_dt_raw(void *h, uint32_t *entries, int32_t *nnz) {
    orc_t *o = (orc_t *)h;
    for (int d = 0; d < o->D; ++d) {
        nnz[d] = o->dt_nnz[d];
        memcpy(entries + o->dt_off[d], o->dt + o->dt_off[d], sizeof(uint32_t) * o->dt_nnz[d]);
    }
}
/* Top-3 per word, indexed by INPUT word id. */
void orc_get_top(void *h, uint16_t *out) {
    orc_t *o = (orc_t *)h;
    for (int w = 0; w < o->V; ++w)
        memcpy(out + (size_t)w * 3, o->top + (size_t)o->rank_of[w] * 3, sizeof(uint16_t) * 3);
}
/* Per-token aux in doc-major order. */
void orc_get_aux(void *h, uint32_t *out) {
    orc_t *o = (orc_t *)h;
    for (int64_t i = 0; i < o->N; ++i) out[i] = o->aux[o->doc_tok[i]];
}
void orc_get_inv(void *h, float *out) {
    orc_t *o = (orc_t *)h;
    memcpy(out, o->inv, sizeof(float) * o->K);
}
/* Word-major layout (for checking the product's layout builder). */
void orc_get_layout(void *h, int64_t *word_off, int32_t *tok_doc, int64_t *doc_tok, int32_t *rank_of) {
    orc_t *o = (orc_t *)h;
    memcpy(word_off, o->word_off, sizeof(int64_t) * (o->V + 1));
    memcpy(tok_doc, o->tok_doc, sizeof(int32_t) * o->N);
    memcpy(doc_tok, o->doc_tok, sizeof(int64_t) * o->N);
    memcpy(rank_of, o->rank_of, sizeof(int32_t) * o->V);
}

/* Uniform drawn for a token in a sweep (exposed for the statistical tests). */
float orc_uniform(uint64_t seed, uint64_t tok, uint32_t sweep) {
    return u24(philox_x0(seed, tok, sweep, 0u));
}

int orc_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
