/*
 * bm25mi.h - C ABI of the MI355X (gfx950) BM25 full-Search library.
 *
 * One shared library, libbm25mi.so, keeps a BM25 index resident in HBM and
 * answers Search(query term IDs, K) -> (doc IDs, FP32 scores) under strict
 * ranking: score descending, then doc ID ascending. Every signature uses plain
 * pointers and sizes; no framework types cross this boundary.
 *
 * Reference interfaces replaced (paths relative to the reference checkout,
 * FR = feature_retrieval/src/ascendfaiss):
 *   bm25mi_create   <- faiss::ascend::AscendIndexBm25::Init
 *                      (FR/ascendhost/include/index/AscendIndexBm25.h:47,
 *                       FR/ascendhost/src/bm25/AscendIndexBm25.cpp:33-48) and the
 *                      bm25large::Session constructor it runs
 *                      (FR/ascendhost/src/bm25/large_search.cpp:107-188)
 *   bm25mi_search   <- AscendIndexBm25::Search (AscendIndexBm25.h:49,
 *                      AscendIndexBm25.cpp:50-65) -> Session::Search
 *                      (large_search.cpp:190-212); the two aclnn operators it
 *                      replays, aclnnBm25FusedScoreTopK and aclnnBm25TopKMerge
 *                      (FR/ascendhost/src/bm25/fused_runtime.h:28-39)
 *   bm25mi_destroy  <- AscendIndexBm25::Close (AscendIndexBm25.h:51,
 *                      AscendIndexBm25.cpp:67-78)
 *   bm25mi_last_error <- AscendIndexBm25::GetLastError (AscendIndexBm25.h:52;
 *                      ErrorInfo in FR/common/bm25/error_info.h:8-20)
 *
 * Error convention: every call returns a status that mirrors the four classes
 * of FR/ascendhost/include/impl/Bm25ErrorAdapter.h:11-45. Input errors leave a
 * session usable; a device error makes the session fail for good and later
 * searches return BM25MI_ERR_ILLEGAL_OPERATION (large_search.cpp:193-195,208-210).
 */
#ifndef BM25MI_H
#define BM25MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BM25MI_OK = 0,
    BM25MI_ERR_INVALID_PARAM = 1,     /* InvalidArgument -> APP_ERR_INVALID_PARAM */
    BM25MI_ERR_ILLEGAL_OPERATION = 2, /* InvalidState    -> APP_ERR_ILLEGAL_OPERATION */
    BM25MI_ERR_INNER = 3              /* DeviceError/other -> APP_ERR_INNER_ERROR */
};

/* Hard limits of the library. */
#define BM25MI_MAX_K 1024            /* reference: 32 (FR/common/bm25/large_search_plan.h:34) */
#define BM25MI_MAX_TERMS_PER_QUERY 64 /* FR/common/bm25/fixed_query_packet.h:32 */
#define BM25MI_MAX_QUERIES 10000      /* fixed_query_packet.h:28 */

typedef struct bm25mi_session bm25mi_session;

/* Bm25Config (AscendIndexBm25.h:17-25) plus the index-time constants the
 * reference hardcodes (k1 = 1.2 in large_corpus.py:174 and 2.2F = k1 + 1 in
 * cpu_index_search.h:185). */
typedef struct {
    int32_t device_id;       /* deviceId */
    int32_t max_queries;     /* queryCount: largest batch one Search accepts */
    int64_t max_words;       /* maxQueryWords: total term IDs per batch */
    int32_t max_k;           /* topK upper bound; each Search names its own K <= max_k */
    int32_t block_docs;      /* docs per scoring task (one wave): 2048, 4096 or 8192; 0 = 2048 */
    double k1;               /* BM25 k1; impact = idf*(tf*(k1+1)) / (tf+norm) in FP32 */
    int32_t use_graph;       /* 1: replay one captured hipGraph per Search; 0 (default): the same copy
                                and kernels launched directly, measured about 10 us faster per
                                Search on ROCm 7.2 (10M docs, Q8 x 3 terms, K10) */
    int32_t grid_blocks;     /* persistent scoring waves; 0 = as many as fit the CUs */
    int32_t task_cost;       /* estimated fixed cost of one scoring task in postings, used to give
                                every wave equal work; 0 = library default */
    int32_t slots;           /* in-flight Search contexts (buffers, stream, graph) for the server;
                                0 or 1 = one */
    int64_t pool_entries;    /* candidate-pool entries per slot; 0 = max_queries queries of K = max_k
                                (a batch whose queries need more is rejected / split by the server) */
    int32_t live_max_k;      /* largest K a threshold query is scored with on the live path (larger
                                ones take the plain path); 0 = 64 */
    int32_t keep_cap;        /* candidates a scoring task may append before it cuts to its top
                                max(K, keep_cap) (at most block_docs; sizes the candidate pools);
                                0 = max(2K, 64). A launch setting like block_docs: results are
                                the same for every value (the reference's candidate capacity
                                variants, tools/capacity_sources.py) */
} bm25mi_config;

/* Per-Search counters, filled by bm25mi_last_stats. */
typedef struct {
    double device_ms;          /* HIP-event time of the device work of the last timed Search
                                  (bm25mi_search_timed); -1 after an untimed one */
    double score_kernel_ms;    /* HIP-event time of the scoring kernel alone (timed Search only, else -1) */
    double host_ms;            /* wall time of the whole bm25mi_search* call */
    int64_t postings_scanned;  /* postings streamed from HBM by the scoring kernel */
    int64_t postings_total;    /* postings of the query terms (plain-path work) */
    int64_t blocks_scored;     /* (query, block) tasks scored */
    int64_t blocks_skipped;    /* (query, block) tasks skipped (no postings, or below a threshold) */
    int64_t candidates;        /* entries written to the candidate pools */
    int32_t fallbacks;         /* queries rerun on the plain path */
    int32_t queries;
    double final_ms;           /* HIP-event time of the final selection kernel (timed Search only, else -1) */
    int64_t postings_in_range; /* streamed postings inside their task's doc range */
    int64_t emitting_blocks;   /* scoring tasks that appended candidates */
    int64_t exact_blocks;      /* tasks that cut their candidates to the per-block cap */
    int32_t score_grid;        /* scoring waves launched (persistent grid) */
    int32_t block_docs;        /* docs per scoring task */
    int64_t phase_cycles[4];   /* scoring-kernel shader cycles summed over waves per phase: locate postings,
                                  stream and add, select, append; zero unless built with -DBM25MI_PHASES */
    int64_t wave_busy_max;     /* -DBM25MI_PHASES builds: longest busy time of a scoring wave (100 MHz ticks), */
    int64_t wave_span;         /* first wave start to last wave end, */
    int64_t wave_start_spread; /* and first to last wave start */
} bm25mi_stats;

void bm25mi_config_default(bm25mi_config *cfg);

/* Copies, validates (FR/common/bm25/index_data.h:13-41) and uploads the index,
 * precomputes FP32 impacts, builds the block directory and block-max table,
 * and captures the Search graph. Arrays follow the reference index files
 * (standalone/resident_index.py:12-18): term_offsets int64[vocab+1],
 * doc_ids int32[postings], tf float[postings], doc_norm float[docs],
 * idf float[vocab]. */
int bm25mi_create(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids, const float *tf,
                  int64_t postings, const float *doc_norm, int64_t docs, const float *idf,
                  const bm25mi_config *cfg, bm25mi_session **out);

/* Plain full Search. query_offsets int64[queries+1] (offsets[0] == 0), query_terms
 * int32[offsets[queries]] in query order (duplicates kept, IDs >= vocab skipped,
 * negative IDs rejected). Writes out_scores float[queries*k], out_ids
 * int32[queries*k], row-major, strict ranking. */
int bm25mi_search(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms, int32_t queries,
                  int32_t k, float *out_scores, int32_t *out_ids);

/* Uneven batch: per-query K (k_per_query[q] in [1, max_k]). Results of query q
 * start at row offset sum_{p<q} k_per_query[p]. */
int bm25mi_search_batch(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                        int32_t queries, const int32_t *k_per_query, float *out_scores, int32_t *out_ids);

/* Plain Search that allocates nothing (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, serving loops and harnesses
 * that forbid hipMalloc). Same arguments and results as bm25mi_search_batch, but the Search's
 * device scratch (request packet, per-query counters, score histograms and bounds, per-wave
 * counter rows, candidate pools) lives in one static device array of
 * BM25MI_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the process ends;
 * the call makes no hipMalloc/hipFree and no host allocation. A batch whose candidate pools
 * need more than the array holds (each query needs min(blocks x max(2K, 64), docs) 8-byte
 * entries; e.g. 8 queries of K = 100 on 10M docs with 2048-doc tiles) is rejected with
 * BM25MI_ERR_INVALID_PARAM. Every call on one GPU uses that GPU's array: calls that overlap
 * on one GPU (two sessions on one device searched from two threads) race; calls on one
 * session are serialised by its lock. */
#define BM25MI_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int bm25mi_search_batch_noalloc(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                int32_t queries, const int32_t *k_per_query, float *out_scores, int32_t *out_ids);

/* Threshold-pruned Search (no reference counterpart: the learned path).
 * thresholds[q] is a predicted K-th score (0 = plain). 256-doc granules whose
 * upper bound (the FP32 sum of the query words' granule maxima) is below it are
 * skipped; only docs scoring at least the threshold are kept. A query with fewer
 * than K survivors, or whose candidate pool fills, is rerun on the plain path,
 * so results always equal bm25mi_search_batch; fallback_out[q] (optional) is
 * set to 1 for each rerun query. A threshold below 0.3 of the query's score
 * upper bound (the FP32 sum of its terms' largest impacts) would leave most
 * granules live, and a query with K above 64 gains nothing from the live path:
 * such queries are scored on the plain path directly. */
int bm25mi_search_pruned(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                         int32_t queries, const int32_t *k_per_query, const float *thresholds,
                         float *out_scores, int32_t *out_ids, int32_t *fallback_out);

/* The resident K-th-score model of learned Searches (no reference counterpart). model holds
 * the MLP's parameters (bm25mi_mlp_param_count(BM25MI_FEATURES) floats in the layout of the
 * bm25mi_mlp_* kernels) followed by the features' mean[BM25MI_FEATURES] and
 * scale[BM25MI_FEATURES]; count must be their total. For a query with features f
 * (bm25mi_query_features) and score upper bound u (bm25mi_query_bounds), the model's output z
 * on (f - mean) / scale gives the threshold clamp(safety * u * exp(z * y_scale + y_mean), lower,
 * u). The model is copied to device memory (allocated by the first call on a session). */
int bm25mi_set_threshold_model(bm25mi_session *s, const float *model, int64_t count, float y_mean, float y_scale,
                               float safety);

/* The thresholds a learned Search of this batch would use, predicted on the device by the
 * session's model (no reference counterpart): out_thresholds[q] = 0 for a query routed to the
 * plain path. */
int bm25mi_learned_thresholds(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                              int32_t queries, const int32_t *k_per_query, float *out_thresholds);

/* Learned Search (no reference counterpart): bm25mi_search_pruned whose thresholds the
 * session's resident model predicts on the device, inside the Search (one kernel before the
 * scoring kernels). A query is routed to the live path when its threshold is at least 0.3 of its
 * score upper bound and its K at most cfg.live_max_k, else to the plain path; a query that keeps
 * fewer than K docs is rerun on the plain path (fallback_out[q] = 1), so results always equal
 * bm25mi_search_batch. timed != 0: direct launches with HIP events, as bm25mi_search_timed
 * (score_kernel_ms then spans threshold prediction through plain and live scoring). Requires
 * bm25mi_set_threshold_model. */
int bm25mi_search_learned(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                          int32_t queries, const int32_t *k_per_query, int32_t timed, float *out_scores,
                          int32_t *out_ids, int32_t *fallback_out);

/* Same as bm25mi_search_batch (thresholds == NULL) or bm25mi_search_pruned, but
 * launches the kernels directly instead of replaying the graph, with HIP events
 * around the scoring kernel (bm25mi_stats.score_kernel_ms). For measurement. */
int bm25mi_search_timed(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                        int32_t queries, const int32_t *k_per_query, const float *thresholds, float *out_scores,
                        int32_t *out_ids, int32_t *fallback_out);

/* Query features for the learned threshold (no reference counterpart): writes
 * float[queries * BM25MI_FEATURES], per query: log K, the number of known terms, log of the score
 * upper bound, log of the largest and second-largest K-th term impacts (the largest is the
 * provable lower bound of the K-th score), the three largest term maxima, the sum of the K-th
 * term impacts, of the term mean impacts and of the idf, log1p of the sum, maximum and minimum
 * df, the density sum of term maximum * df / docs, and the sum of log(df / docs) (at least -40). */
#define BM25MI_FEATURES 16
int bm25mi_query_features(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                          int32_t queries, const int32_t *k_per_query, float *out_features);

/* Per query (no reference counterpart): out_lower[q] = the K-th largest impact of any of its
 * terms, a lower bound of its K-th score (every doc scores at least each of its impacts);
 * out_upper[q] = the FP32 sum of its terms' largest impacts in query order, an upper bound of
 * every doc's score. Either output may be NULL. */
int bm25mi_query_bounds(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                        int32_t queries, const int32_t *k_per_query, float *out_lower, float *out_upper);

/* Changes the scoring launch of a session without rebuilding it (for the autotuner):
 * grid_waves persistent scoring waves (at most twice the waves that fit the GPU at once, or
 * the configured grid_blocks) and the schedule's per-task cost; 0 keeps the current value.
 * Results do not depend on either. */
int bm25mi_set_launch(bm25mi_session *s, int32_t grid_waves, int32_t task_cost);

int bm25mi_last_stats(const bm25mi_session *s, bm25mi_stats *out);

/* Close: idempotent; frees device memory. The handle is freed on success. */
int bm25mi_destroy(bm25mi_session *s);

/* Failure injection for tests of the error contract (no counterpart in the reference's public
 * API: its unit tests inject the same failures through mock ACL functions,
 * FR/ut/cases/bm25/sdk_facade_tests.cpp:53-62, runtime_tests.cpp:19-23). The next library call
 * that reaches `stage` fails once with BM25MI_ERR_INNER, device code device_code and the given
 * message (kept whole here; bm25mi_last_error returns at most 1023 characters of it):
 *   "create"  bm25mi_create, before any device work;
 *   "search"  a Search on a session, after its request was accepted and before its device
 *             work: the session fails for good and the outputs are left untouched;
 *   "close"   bm25mi_destroy, before anything is released: the session stays usable and
 *             Close may be retried;
 *   "result"  the next Search's first query rows are corrupted on the host after the device
 *             wrote them (its first two rows exchanged, or its first doc ID set to the document
 *             count when K = 1), before the result checks of Session::Receive
 *             (large_search.cpp:64-85) run: they reject it as a device error (code 0, device_code
 *             and message unused), which fails the session.
 * A NULL stage clears every pending injection. Injections are process-wide. */
int bm25mi_inject_failure(const char *stage, int64_t device_code, const char *message);

/* Copies the last error message of s (or of the calling thread when s is NULL)
 * into buf (truncated to size-1 chars, NUL-terminated) and returns the raw
 * HIP error code of a device failure, else 0 (ErrorInfo::aclCode). */
int64_t bm25mi_last_error(const bm25mi_session *s, char *buf, size_t size);

/* Device properties of the session's GPU: compute units, LDS bytes per block. */
int bm25mi_device_info(int32_t device_id, int32_t *cus, int32_t *lds_per_block, int32_t *lds_per_cu);

/* ---------------------------------------------------------------------------
 * Serving loop with continuous batching (no reference counterpart). A server takes
 * requests from any thread and keeps up to cfg.slots batches in flight on the session:
 * whenever a slot is free it admits the oldest waiting requests, up to max_batch queries,
 * the word capacity and the slot's candidate-pool memory (pool_entries), and launches them
 * without waiting, so requests that arrive while others run are batched and overlap on the
 * GPU. Results equal a Search of the same query alone. While a server runs, the session's
 * own Search entry points and bm25mi_destroy return BM25MI_ERR_ILLEGAL_OPERATION.
 * ------------------------------------------------------------------------- */
typedef struct bm25mi_server bm25mi_server;
int bm25mi_server_start(bm25mi_session *s, int32_t max_batch, bm25mi_server **out);
/* Queues one query (n term IDs, top-K); *ticket identifies it for bm25mi_server_wait. */
int bm25mi_server_submit(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k, uint64_t *ticket);
/* Same with a pruning threshold (a predicted K-th score, as bm25mi_search_pruned): a query
 * that keeps fewer than K docs at or above it is rerun as a plain query, so the result is
 * the same. A request with K above 64 is queued as a plain one. */
int bm25mi_server_submit_pruned(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k, float threshold,
                                uint64_t *ticket);
/* Same as a learned Search of one query: its threshold is predicted on the device by the
 * session's model (bm25mi_set_threshold_model) in the batch it is admitted to; a request that
 * keeps fewer than K docs is rerun as a plain one. Batches hold learned and plain requests, or
 * requests with thresholds of their own and plain ones. */
int bm25mi_server_submit_learned(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k, uint64_t *ticket);
/* Blocks until the query is done; writes its K scores and doc IDs (either may be NULL) and
 * the submit-to-completion latency in microseconds (may be NULL). */
int bm25mi_server_wait(bm25mi_server *srv, uint64_t ticket, float *scores, int32_t *ids, double *latency_us);
/* Open-loop load: query i (terms offsets[i]..offsets[i+1], K ks[i], threshold thresholds[i]
 * when thresholds is not NULL, learned when learned != 0) is submitted at the i-th arrival of a
 * seeded Poisson process of rate qps; latency_us[i] gets its latency, scores / ids (optional) its
 * rows at offset sum_{j<i} ks[j], *elapsed_s the wall time of the run. */
int bm25mi_server_run_poisson(bm25mi_server *srv, const int64_t *offsets, const int32_t *terms, int32_t n,
                              const int32_t *ks, const float *thresholds, int32_t learned, double qps,
                              uint64_t seed, double *latency_us, float *scores, int32_t *ids, double *elapsed_s);
/* Batches launched, queries served (reruns included), the largest number of batches in flight
 * and the pruned queries rerun as plain ones, so far. Any output may be NULL. */
int bm25mi_server_counters(bm25mi_server *srv, int64_t *batches, int64_t *queries, int32_t *max_inflight,
                           int64_t *fallbacks);
/* Finishes the submitted queries and stops; NULL is a no-op. */
int bm25mi_server_stop(bm25mi_server *srv);

/* ---------------------------------------------------------------------------
 * Index builder (host C++, no GPU). The reference's Python builder is
 * standalone/large_corpus.py:95-297; the synthetic generator here follows the
 * seeded Zipf/lognormal corpus spec of this project instead of the
 * reference's uniform vocab-1024 stream.
 * ------------------------------------------------------------------------- */
typedef struct bm25mi_corpus bm25mi_corpus;

typedef struct {
    int64_t num_docs;
    int64_t vocab;
    double zipf_s;     /* token rank r in [1, vocab] drawn with P(r) ~ r^-s; term ID = r-1 */
    double mean_len;   /* lognormal doc length mean (tokens) */
    double sigma;      /* lognormal sigma */
    uint64_t seed;
    int32_t threads;   /* 0 = hardware concurrency; output never depends on it */
} bm25mi_corpus_spec;

int bm25mi_corpus_generate(const bm25mi_corpus_spec *spec, bm25mi_corpus **out);
/* Borrowed views, valid until bm25mi_corpus_free: term_offsets int64[vocab+1],
 * doc_ids int32[postings], tf float[postings], doc_len int32[num_docs]. */
int bm25mi_corpus_view(const bm25mi_corpus *c, int64_t *num_docs, int64_t *vocab, int64_t *postings,
                       const int64_t **term_offsets, const int32_t **doc_ids, const float **tf,
                       const int32_t **doc_len);
void bm25mi_corpus_free(bm25mi_corpus *c);

/* idf = f32(log1p((N-df+0.5)/(df+0.5))) and norm = f32(k1*(1-b+b*len/avgdl)), both
 * evaluated in double (standalone/large_corpus.py:148,174). */
int bm25mi_bm25_weights(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_len, int64_t docs,
                        double k1, double b, float *out_idf, float *out_norm, int32_t threads);

/* impact[p] = f32(f32(idf[t] * f32(tf[p] * f32(k1+1))) / f32(tf[p] + norm[doc[p]]))
 * (FR/test/ascendhost/bm25/reference/cpu_index_search.h:185-186). */
int bm25mi_impacts(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids, const float *tf,
                   const float *doc_norm, const float *idf, double k1, float *out_impact, int32_t threads);

/* Index validation of FR/common/bm25/index_data.h:13-41; returns
 * BM25MI_ERR_INVALID_PARAM with a message on the first violation. */
int bm25mi_validate_index(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids, const float *tf,
                          int64_t postings, const float *doc_norm, int64_t docs, const float *idf);

/* ---------------------------------------------------------------------------
 * Learned K-th-score model kernels (no reference counterpart). A 3-layer MLP,
 * in -> 64 -> 64 -> 1, ReLU, FP32. All pointers are device pointers;
 * params is the flat layout W1[64*in] b1[64] W2[64*64] b2[64] W3[64] b3[1].
 * ------------------------------------------------------------------------- */
#define BM25MI_MLP_HIDDEN 64
int64_t bm25mi_mlp_param_count(int32_t in_features);
/* y[n] = MLP(x[n, in]) */
int bm25mi_mlp_forward(const float *params, const float *x, int32_t n, int32_t in_features, float *y,
                       void *stream);
/* Mean-squared-error loss and gradient: grad (same layout as params) +=
 * d/dparams of (1/n) * sum (y - target)^2 ; loss_out[0] = that mean. Asynchronous on stream:
 * calls share one buffer of workgroup partials, and each call's stream waits on the previous
 * call's work (from any stream) before reusing it; the host waits only when the buffer grows. */
int bm25mi_mlp_backward(const float *params, const float *x, const float *target, int32_t n, int32_t in_features,
                        float *grad, float *loss_out, void *stream);
/* Adam step over count parameters; state m, v are device arrays of count floats. */
int bm25mi_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t count, float lr, float beta1,
                    float beta2, float eps, int32_t step, void *stream);

/* Host-pointer forms (the library owns the device buffers). */
/* He-uniform weights, zero biases, from a seeded stream; params_out float[param_count]. */
int bm25mi_mlp_init(int32_t in_features, uint64_t seed, float *params_out);
/* One GPU backward pass: grad_out = d/dparams (1/n) sum (y - target)^2, loss_out = that mean. */
int bm25mi_mlp_gradient(const float *params, const float *x, const float *target, int32_t n, int32_t in_features,
                        float *grad_out, float *loss_out);
/* y_out[n] = MLP(x[n, in]) on the GPU. */
int bm25mi_mlp_predict(const float *params, const float *x, int32_t n, int32_t in_features, float *y_out);
/* Minibatch Adam (beta 0.9/0.999, eps 1e-8) on the GPU for `epochs` seeded shuffles of the data;
 * params in/out; epoch_loss float[epochs] (optional) gets each epoch's mean minibatch loss. */
int bm25mi_mlp_train(float *params, const float *x, const float *target, int32_t n, int32_t in_features,
                     int32_t epochs, int32_t batch, float lr, uint64_t seed, float *epoch_loss);

#ifdef __cplusplus
}
#endif

#endif /* BM25MI_H */
