/*
 * ezlda.h -- C ABI of the MI355X-native EZLDA Gibbs-sampling sweep.
 *
 * Plain C: no HIP, CUDA or torch types.  Every function returns an int
 * status (EZLDA_OK = 0) unless stated; the message of the last failure is
 * ezlda_last_error(ctx) (ctx may be NULL for failures before a context
 * exists).  A context owns all of its device memory; arrays passed in or out
 * are host arrays owned by the caller.  One context per host thread; calls
 * on one context must not overlap.
 *
 * The reference exposes its hot path as ten host launchers that take its
 * container classes and a cudaStream_t& (EZLDA/SamplingKernel.cuh:10-15,
 * DTUpdateKernel.cuh:13, WTUpdateKernel.cuh:23, WTDenUpdateKernel.cuh:25-26,
 * WTAddKernel.cuh:24).  Each stage entry point below names the launchers it
 * replaces.  The corpus stays resident on the device, so the reference's
 * per-chunk staging methods (Doc.cu:266-319, WT.cu:212-239, DT.cu:135-142)
 * have no counterpart here.
 */
#ifndef EZLDA_H
#define EZLDA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EZLDA_OK 0
#define EZLDA_EINVAL 1  /* bad argument or corpus */
#define EZLDA_EHIP 2    /* HIP runtime error */
#define EZLDA_EIO 3     /* file missing, truncated or inconsistent */
#define EZLDA_ESTATE 4  /* call out of order (e.g. no corpus loaded) */
#define EZLDA_ENOMEM 5  /* over the context's memory budget */

typedef struct ezlda_ctx ezlda_ctx;

/* Run-time form of the reference's compile-time Argument surface
 * (EZLDA/Argument.cuh:3-9) plus the constants main() hard-codes
 * (mainKernel.cu:50-51) and Doc.cuh:30 tokenSegment. */
typedef struct ezlda_args {
    int32_t K;               /* topics (Argument.cuh:6); 1 <= K <= 16383 */
    float alpha;             /* Argument.cuh:9; <= 0 means 50/K */
    float beta;              /* Argument.cuh:8; <= 0 means 0.01 */
    int32_t block_dim;       /* Argument.cuh:4 BlockDim: threads per workgroup
                                of the sampler; multiple of 64, <= 1024
                                (<= 0 = auto: the smallest of 64, 128, ...
                                with which the workgroups whose LDS fits in
                                a CU at this K hold >= 12 wavefronts: 64 up
                                to K ~ 2400, 256 at K = 10000) */
    int32_t grid_dim;        /* Argument.cuh:5 GridDim: > 0 = that many
                                persistent sampler workgroups pulling items
                                from per-XCD queues (then from the others');
                                0 = one workgroup per item; < 0 = 16 per CU
                                at 64 threads, else 8 per CU.
                                Topics do not depend on it */
    int32_t num_chunks;      /* mainKernel.cu:50: > 1 = out-of-core sweeps:
                                the documents are cut into that many chunks
                                (about equal token counts) whose token state
                                and layout stay in pinned host memory and pass
                                through two device buffers on two streams per
                                sweep; topics equal the resident ones, and
                                the counts (ezlda_get_nk, _get_wt_dense) are
                                those of the current topics.  The per-stage
                                calls, graphs, layout tuning and DT/aux
                                getters then return EZLDA_ESTATE; model
                                export works in both modes.
                                <= 1 = the corpus is resident */
    int32_t num_streams;     /* Argument.cuh:7; recorded (chunked sweeps use 2) */
    int32_t token_segment;   /* Doc.cuh:30: max tokens of one word per sampler
                                work item (0 = 4096) */
    uint64_t seed;           /* Philox key */
    int32_t dense_words;     /* head size V_D (words by rank); < 0 = use
                                dense_threshold */
    int32_t dense_threshold; /* count >= threshold makes a head word
                                (preprocess/streamTestBalance.cpp:291 uses K);
                                <= 0 = K */
    int32_t shortcut;        /* 1 = max-topic shortcut, 0 = plain sampler;
                                both give identical topics */
    int32_t topic_bits;      /* width of a stored topic (the CLI's
                                --topic-bits): 16 (0 = 16); K <= 16383 keeps
                                every topic in 16 bits, other widths are
                                refused.  Compact DT rows pack a topic in
                                max(10, ceil(log2 K)) bits internally */
    int32_t doc_tiles;       /* head-word sampler items cut by doc tiles for
                                L2 locality, dealt to XCDs: 0 = auto,
                                > 0 = that many tiles, < 0 = off; topics do
                                not depend on it */
    int32_t run_cap;         /* the sampler shares one DT row scan among the
                                repeats of a word in a document, at most
                                run_cap tokens (1..32, 0 = 4); 1 = one scan
                                per token.  Topics do not depend on it */
} ezlda_args;

typedef struct ezlda_info {
    int64_t num_tokens;      /* N */
    int32_t num_words;       /* V */
    int32_t num_docs;        /* D */
    int32_t num_head_words;  /* V_D */
    int32_t K;
    int64_t dt_capacity;     /* entries of ezlda_get_dt_raw: sum_d of min(n_d, K),
                                plus n_d / p when K <= 8192 (rows then hold a
                                count in pieces of at most p = 2^(16 - tb) - 1,
                                tb = max(10, ceil(log2 K)): 63 up to K = 1024,
                                31, 15, 7 up to 2048, 4096, 8192) */
    int64_t wt_tail_capacity;/* sum_tail min(n_w, K) */
    int32_t num_sample_items;/* sampler work items (word segments) */
    uint32_t sweep;          /* sweeps done (the Philox sweep counter) */
    int64_t device_bytes;    /* device memory held by the context */
    int32_t doc_tiles;       /* doc tiles used for head-word items */
    int32_t num_xcd;         /* XCDs of the device */
    int32_t run_cap;         /* tokens per run in effect */
    int32_t grid_dim;        /* sampler workgroups in effect (0 = one per item) */
    int32_t block_dim;       /* sampler workgroup size in effect */
    int32_t token_segment;   /* in effect */
    int64_t num_runs;        /* sampler runs (a word's tokens in one doc) */
    int32_t num_chunks;      /* chunks of the out-of-core path (1 = resident) */
} ezlda_info;

void ezlda_default_args(ezlda_args *args);
const char *ezlda_last_error(const ezlda_ctx *ctx);

/* Creates a context on the current HIP device.  NULL on failure. */
ezlda_ctx *ezlda_create(const ezlda_args *args);
void ezlda_destroy(ezlda_ctx *ctx);

/* Loads a doc-major corpus: doc d holds words[doc_off[d] .. doc_off[d+1]),
 * word ids in [0, V).  z0 (nullable) gives the initial topic of every token
 * in the same order, 0-based; NULL draws it from Philox.  Replaces
 * Document::loadDocument + GPUMemAllocate (Doc.cu:45-108, :179-246),
 * DTChunk::loadDocDTLength/loadDTCountOffset (DT.cu:4-100), WTAll/WTD
 * set-up (WT.cu:3-193, WTDense.cu:3-43). */
int ezlda_load_docs(ezlda_ctx *ctx, int32_t V, int32_t D, const int64_t *doc_off,
                    const int32_t *words, const uint16_t *z0);

/* Loads the chunked text corpus written by preprocess/streamTestBalance.cpp
 * (lengthVec.txt, TLLength.txt, docLength.txt, chunk<c>/TL.txt,
 * word2DocMap.txt, wordCountOffset.txt, docCountOffset.txt ...); topics in
 * TL.txt are 1-based.  Reference readers: DataChunk.cu:50-100, Doc.cu:45-108. */
int ezlda_load_chunked(ezlda_ctx *ctx, const char *prefix, int32_t num_chunks);

int ezlda_get_info(ezlda_ctx *ctx, ezlda_info *info);

/* ---- stages (for parity tests and custom drivers; ezlda_sweep runs all) */

/* WT (dense head rows + sorted sparse tail rows) and n_k from the current z.
 * Replaces UpdateWTDenKernel (WTDenUpdateKernel.cuh:25), UpdateWTKernel
 * (WTUpdateKernel.cuh:23), WTAdditionKernel (WTAddKernel.cuh:24) and
 * UpdateWTDenRowSumKernel (WTDenUpdateKernel.cuh:26). */
int ezlda_build_wt(ezlda_ctx *ctx);
/* 1/(n_k + V beta) and every word's top-3 topics by phi.
 * Replaces MaxTopicKernel (SamplingKernel.cuh:12). */
int ezlda_topk(ezlda_ctx *ctx);
/* Sorted DT rows and each token's record for the sampler (n_dk1 and its
 * doc's row length; ezlda_get_aux gives the reference's (n_dk1, n_dk2)).
 * Replaces UpdateDTKernel (DTUpdateKernel.cuh:13). */
int ezlda_build_dt(ezlda_ctx *ctx);
/* Draws a new topic for every token and advances the sweep counter.
 * Replaces UpdateProbKernelD (SamplingKernel.cuh:15), SampleKernelD (:10)
 * and SampleKernel (:11). */
int ezlda_sample(ezlda_ctx *ctx);
/* The sampler and count stages at the reference's launcher granularity.  A
 * token's draw depends only on the counts built at the start of the sweep,
 * so ezlda_shortcut, ezlda_sample_head and ezlda_sample_tail (in any order)
 * followed by ezlda_commit_head give exactly the topics of ezlda_sample.
 * Max-topic shortcut of the head words' tokens: each accepted token gets its
 * word's top topic, the others keep theirs for the head sampler.  A no-op
 * when args.shortcut is 0.  Replaces UpdateProbKernelD (SamplingKernel.cuh:15). */
int ezlda_shortcut(ezlda_ctx *ctx);
/* Draws the head words' tokens (the shortcut test first; tokens it accepts
 * get the top topic, as after ezlda_shortcut).  Replaces SampleKernelD
 * (SamplingKernel.cuh:10). */
int ezlda_sample_head(ezlda_ctx *ctx);
/* Draws the tail words' tokens, with the same inline shortcut.  Replaces
 * SampleKernel (SamplingKernel.cuh:11). */
int ezlda_sample_tail(ezlda_ctx *ctx);
/* Ends the sweep: the dense head rows are rebuilt from the new z (the
 * reference copies its accumulated head counts back, mainKernel.cu:355-356)
 * and the sweep counter advances.  Tail rows and n_k are rebuilt by the next
 * ezlda_build_wt. */
int ezlda_commit_head(ezlda_ctx *ctx);
/* n_k = column sums of WT, from the current z.  Replaces
 * UpdateWTDenRowSumKernel (WTDenUpdateKernel.cuh:26). */
int ezlda_wt_rowsum(ezlda_ctx *ctx);
/* FP64 log-likelihood sum_t log p(w_t | d_t) of the current z (rebuilds the
 * counts it needs).  Replaces PerplexityKernel (SamplingKernel.cuh:13);
 * perplexity = exp(-ll / N). */
int ezlda_loglik(ezlda_ctx *ctx, double *ll);
/* The same log-likelihood, bit for bit, from a call that allocates nothing:
 * no reference counterpart.  It is for callers that cannot allocate device
 * memory at this point (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its per-item partials and result live
 * in a static device array of EZLDA_LOGLIK_NOALLOC_MAX_ITEMS + 1 doubles
 * (8 MiB) that each GPU holds until the process ends; a corpus with more
 * sampler items (ezlda_info.num_sample_items; a chunk's items when chunked)
 * is rejected with EZLDA_EINVAL.  Every call on one GPU shares the array:
 * calls that overlap on one GPU (from different contexts or threads) race. */
#define EZLDA_LOGLIK_NOALLOC_MAX_ITEMS 1048576
int ezlda_loglik_noalloc(ezlda_ctx *ctx, double *ll);

/* ---- whole sweeps (mainKernel.cu:158-380) */
int ezlda_sweep(ezlda_ctx *ctx, int32_t num_sweeps);
/* The same sweeps replayed from one captured hipGraph. */
int ezlda_sweep_graph(ezlda_ctx *ctx, int32_t num_sweeps);
/* Timed sweeps: total wall time (HIP events on the context's stream) and
 * the summed time of each stage (EZLDA_STAGE_*), all in ms. */
#define EZLDA_STAGE_WT 0
#define EZLDA_STAGE_TOPK 1
#define EZLDA_STAGE_DT 2
#define EZLDA_STAGE_SAMPLE 3
#define EZLDA_NUM_STAGES 4
int ezlda_time_sweeps(ezlda_ctx *ctx, int32_t num_sweeps, int32_t use_graph,
                      double *total_ms, double *stage_ms);
/* Sampler counters summed since the last reset: stats[0] = tokens that took
 * the full three-bucket draw (the rest took the max-topic shortcut),
 * stats[1] = DT entries those draws read.  The reference counts its
 * full-sampled tokens the same way (utility.cu:1570,2004).  Also, of the
 * current counts: stats[2] = nonzeros of the sparse WT tail, stats[3] =
 * nonzeros of DT.  stats must hold 4 values. */
int ezlda_get_stats(ezlda_ctx *ctx, uint64_t *stats);
int ezlda_reset_stats(ezlda_ctx *ctx);

/* ---- state in / out (host arrays, caller-owned) */
/* z of every token, in the doc-major order it was loaded in. */
int ezlda_get_topics(ezlda_ctx *ctx, uint16_t *z);
int ezlda_set_topics(ezlda_ctx *ctx, const uint16_t *z);
int ezlda_set_sweep(ezlda_ctx *ctx, uint32_t sweep);

/* ---- layout tuning (no reference counterpart: the reference's BlockDim,
 * GridDim, tokenSegment and head threshold are compile-time constants,
 * Argument.cuh:3-9, preprocess/streamTestBalance.cpp:291).  Topics never
 * depend on these arguments, so they can change between sweeps. */
/* Rebuilds the layout with args' block_dim, grid_dim, token_segment,
 * dense_words, dense_threshold, doc_tiles and run_cap (others ignored). */
int ezlda_set_layout(ezlda_ctx *ctx, const ezlda_args *args);

typedef struct ezlda_tune_trial {
    ezlda_args args;
    double ms_per_sweep;
} ezlda_tune_trial;
/* Coordinate descent over the layout arguments (head size, run cap, doc
 * tiles, block size, token segment, persistent grid), each trial one untimed
 * and sweeps_per_trial timed sweeps; the sweeps are ordinary sweeps of the
 * run.  Leaves the context on the fastest layout and reports it; trials may
 * be NULL. */
int ezlda_autotune(ezlda_ctx *ctx, int32_t sweeps_per_trial, ezlda_args *chosen, double *ms_per_sweep,
                   ezlda_tune_trial *trials, int32_t max_trials, int32_t *num_trials);
int ezlda_get_nk(ezlda_ctx *ctx, int32_t *nk);
/* Dense V x K word-topic counts, rows by input word id. */
int ezlda_get_wt_dense(ezlda_ctx *ctx, int32_t *wt);
/* Model export of the current state (the reference keeps only debug dumps,
 * WT.cu:242-280, DT.cu:154-168): dir/model.txt ("K V D alpha beta sweep"),
 * dir/phi.f32 (V x K little-endian float32, rows by input word id, phi =
 * (n_wk + beta) / (n_k + V beta) as the sampler computes it) and
 * dir/theta.txt (per input doc: "n_d topic:count ..." in ascending topic).
 * The directory must exist. */
int ezlda_export_model(ezlda_ctx *ctx, const char *dir);
/* DT rows as the sampler walks them: entries[dt_offset(d) + i] = topic |
 * count << 16 for i < nnz[d], ascending topic; for K <= 8192 a count above
 * p (see dt_capacity) appears as pieces of p and a remainder.
 * dt_offset(d) = sum_{e<d} cap(n_e), cap(n) = min(n, K) (+ n / p, at most
 * n, when K <= 8192). */
int ezlda_get_dt_raw(ezlda_ctx *ctx, uint32_t *entries, int32_t *nnz);
/* Top-3 topics per word (3 x uint16 per input word id; 0xFFFF = none). */
int ezlda_get_top(ezlda_ctx *ctx, uint16_t *top);
/* Per-token n_dk1 | n_dk2 << 16, doc-major order, of the current topics and
 * counts (DT_Update_Kernel's per-token pair, EZLDA/utility.cu:125-161).  The
 * sampler keeps n_dk1 with the row length instead, so this runs the DT
 * rebuild in its n_dk2 form into scratch device buffers (it allocates; no
 * state changes). */
int ezlda_get_aux(ezlda_ctx *ctx, uint32_t *aux);
int ezlda_get_inv(ezlda_ctx *ctx, float *inv);

/* ---- device self-tests (no reference counterpart; they pin the RNG the
 * oracle restates to what the device computes, and time a plain copy) */
/* out[4i..4i+3] = Philox4x32-10(ctr[4i..4i+3], key[2i..2i+1]). */
int ezlda_selftest_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out, int32_t n);
/* Streams `bytes` from one device buffer to another `reps` times with a
 * plain copy kernel; *ms = mean time per copy.  The known-size workload
 * that scales hardware traffic counters. */
int ezlda_selftest_copy(int64_t bytes, int32_t reps, double *ms);
/* n_reads scattered 64-byte block reads (four 16-byte loads per lane, the
 * sampler's DT row-block pattern) over a `bytes` buffer, each of them a
 * distinct block (n_reads must not exceed the largest power of two of 64-byte
 * blocks the buffer holds), `reps` times; *ms = mean time per launch.  The
 * known-size workload that scales the traffic counters for that request mix
 * (64 * n_reads bytes per launch). */
int ezlda_selftest_gather(int64_t bytes, int64_t n_reads, int32_t reps, double *ms);

/* ---- checkpoint (the state is z plus the sweep counter) */
/* A private binary file: header, seed, sweep, z in doc-major order. */
int ezlda_save_checkpoint(ezlda_ctx *ctx, const char *path);
int ezlda_load_checkpoint(ezlda_ctx *ctx, const char *path);
/* The reference's own form (no reference counterpart writes it; its
 * preprocess writes the initial topics there, preprocess/streamTestBalance.cpp:469,
 * and its loader reads them, DataChunk.cu:85): every chunk<c>/TL.txt of the
 * chunked folder `prefix` rewritten with the current topics (1-based, in the
 * file's own token order) and <prefix>/ezlda_state.txt = "seed S\nsweep N\nK K\n".
 * The folder must hold the context's corpus (the one ezlda_load_chunked read
 * from it; checked token by token).  A current binary sidecar is rewritten
 * with the same topics. */
int ezlda_save_tl(ezlda_ctx *ctx, const char *prefix);
/* ezlda_load_chunked(prefix) plus the sweep counter from ezlda_state.txt
 * (when present; its seed and K must be the context's), so the next sweeps
 * are those the saving context would have run. */
int ezlda_load_tl(ezlda_ctx *ctx, const char *prefix);

/* ---- fold-in serving (no reference counterpart: EZLDA only trains)
 * A server freezes phi from a trained context's current counts and infers
 * the topics of incoming documents: `sweeps` Jacobi fold-in sweeps with
 * p(k) ~ (n_dk + alpha) phi_wk per document, drawn with Philox keyed by
 * (seed, request id, token, sweep).  Requests are admitted into device slots
 * while slots and the HBM token budget allow, and every resident request is
 * advanced by one sweep per launch (continuous batching), so a result does
 * not depend on what else is in flight.  Thread-safe: submit and wait may be
 * called from any thread; one internal worker thread drives the device. */
typedef struct ezlda_server ezlda_server;
/* Device bytes per resident token: words, topics, row pool, word-major
 * token list, result staging. */
#define EZLDA_FOLDIN_TOKEN_BYTES 16
typedef struct ezlda_server_args {
    int32_t max_slots;        /* documents resident at once */
    int32_t sweeps;           /* fold-in sweeps per document (e.g. 20) */
    int64_t hbm_budget_bytes; /* device bytes for resident request state
                                 (EZLDA_FOLDIN_TOKEN_BYTES per token; slots
                                 and per-word step buffers come on top);
                                 admission waits for room, a document larger
                                 than all of it is rejected */
    uint64_t seed;
} ezlda_server_args;

typedef struct ezlda_server_stats {
    int64_t submitted, completed, rejected, steps;
    int32_t active, pending, max_active;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> result ready */
    double docs_per_s;                      /* completed / (last result - first submit) */
} ezlda_server_stats;

ezlda_server *ezlda_server_create(ezlda_ctx *model, const ezlda_server_args *args);
void ezlda_server_destroy(ezlda_server *srv);
const char *ezlda_server_last_error(const ezlda_server *srv);
/* Queues one document (input word ids). */
int ezlda_server_submit(ezlda_server *srv, uint32_t request_id, const int32_t *words, int32_t n);
/* Queues n_docs documents at once: document d has id request_ids[d] and the
 * words [doc_off[d], doc_off[d+1]) (doc_off[0] = 0).  Every document is
 * checked first; if any is rejected (as ezlda_server_submit would reject it,
 * or an id repeats within the batch), none is queued. */
int ezlda_server_submit_many(ezlda_server *srv, int32_t n_docs, const uint32_t *request_ids, const int64_t *doc_off,
                             const int32_t *words);
/* Blocks until the request is done (timeout_ms < 0: no limit) and copies its
 * final topics (as many values as the request was submitted with: z_out
 * must hold them) and topic counts (K values); either may be NULL.
 * The result is then released.  An id that is neither queued, resident nor
 * done is EZLDA_EINVAL; so is submitting an id that is in flight or has an
 * unclaimed result. */
int ezlda_server_wait(ezlda_server *srv, uint32_t request_id, uint16_t *z_out, int32_t *counts_out,
                      double *latency_ms, int32_t timeout_ms);
/* Blocks until no request is queued or resident. */
int ezlda_server_drain(ezlda_server *srv);
int ezlda_server_get_stats(ezlda_server *srv, ezlda_server_stats *stats);
int ezlda_server_reset_stats(ezlda_server *srv);
/* The plain path: one document per kernel launch (one workgroup runs all of
 * its sweeps, its row in LDS), synchronous, on its own stream; same results
 * as the server.  Calls on one server are serialized. */
int ezlda_foldin_plain(ezlda_server *srv, uint32_t request_id, const int32_t *words, int32_t n, uint16_t *z_out,
                       int32_t *counts_out);

/* ---- synthetic corpora (LDA generative process, Philox-keyed) */
typedef struct ezlda_gen_params {
    int32_t V, D, K_true;
    int32_t len_dist;     /* 0 Poisson(len_mean); 1 log-normal(len_mean,
                             len_sigma); 2 Pareto(pareto_shape, pareto_xm) */
    double len_mean, len_sigma, pareto_shape, pareto_xm;
    int32_t len_min, len_max;
    double zipf_s;        /* base word distribution ~ rank^-s; 0 = uniform */
    double phi_conc;      /* phi_k ~ Dir(phi_conc * V * base) */
    double theta_conc;    /* theta_d ~ Dir(theta_conc) over K_true topics */
    uint64_t seed;
    int32_t threads;      /* host threads (0 = all) */
    int32_t first_doc;    /* documents are numbered from here: the same seed
                             with first_doc = D of a corpus draws further
                             documents of the same model (0: a fresh corpus) */
} ezlda_gen_params;

typedef struct ezlda_corpus ezlda_corpus;
ezlda_corpus *ezlda_gen_corpus(const ezlda_gen_params *p);
int64_t ezlda_corpus_num_tokens(const ezlda_corpus *c);
int32_t ezlda_corpus_num_docs(const ezlda_corpus *c);
int32_t ezlda_corpus_num_words(const ezlda_corpus *c);
const int64_t *ezlda_corpus_doc_off(const ezlda_corpus *c);
const int32_t *ezlda_corpus_words(const ezlda_corpus *c);
/* Initial topics (0-based) when the source carried them, else NULL. */
const uint16_t *ezlda_corpus_topics(const ezlda_corpus *c);
void ezlda_corpus_free(ezlda_corpus *c);
/* Copies caller arrays into a corpus handle (topics nullable). */
ezlda_corpus *ezlda_corpus_from_arrays(int32_t V, int32_t D, const int64_t *doc_off, const int32_t *words,
                                       const uint16_t *topics);
/* UCI docword bag of words ("D W NNZ", then "doc word count"), the input
 * preprocess/streamTestBalance.cpp:90-145 reads. */
ezlda_corpus *ezlda_read_docword(const char *path);
/* The chunked text folder of preprocess/streamTestBalance.cpp:356-547
 * (Appendix format: lengthVec.txt, TLLength.txt, docLength.txt, DTLength.txt,
 * WTLength.txt, TLSplit.txt, WTCountOffset.txt, chunk<c>/{TL, word2DocMap,
 * doc2WordMap, wordCountOffset, docCountOffset, DTCountOffset,
 * WTCountOffset}.txt).  Topics are written 1-based.  z0 nullable (Philox
 * draws keyed by seed).  Returns a status. */
int ezlda_write_chunked(const char *prefix, const ezlda_corpus *c, const uint16_t *z0, int32_t K,
                        int32_t num_chunks, uint64_t seed);
/* Reads that folder (DataChunk.cu:50-100, Doc.cu:45-108): docs of chunk 0
 * first; word ids are the file's ranks; topics 0-based.  num_chunks <= 0
 * reads the count from TLLength.txt.  NULL on failure.
 * The binary sidecar: ezlda_write_chunked also writes <prefix>/corpus.ezb,
 * the same corpus in binary with the size and modification time of every
 * chunk's TL.txt, word2DocMap.txt and docCountOffset.txt; this reader takes
 * it (no text parsing) while all of them still match, else the text. */
ezlda_corpus *ezlda_read_chunked(const char *prefix, int32_t num_chunks);

#ifdef __cplusplus
}
#endif
#endif /* EZLDA_H */
