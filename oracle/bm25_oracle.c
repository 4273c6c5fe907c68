/*
 * bm25_oracle.c - CPU oracle for BM25 full Search. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline; the
 * product path (libbm25mi.so) never calls it.
 *
 * Plain C11 + OpenMP restatement of the reference's CPU Search
 * (paths relative to the reference checkout, FR = feature_retrieval/src/ascendfaiss):
 *   oracle_search       FR/test/ascendhost/bm25/reference/cpu_index_search.h:154-211
 *                       (document shards, per-shard zero fill, lower_bound into the
 *                       shard range, FP32 contribution (idf*(tf*kp1))/(tf+norm)
 *                       added in query-term order) and
 *                       FR/common/bm25/topk_document_shards.h:13-47 (per-shard
 *                       bounded heaps over all docs including zero scores, then a
 *                       per-query heap merge and sort) with the strict comparator
 *                       of FR/common/bm25/topk_heap.h:11-16. Generalised: K is per
 *                       query and unbounded (reference: K <= 32), kp1 = k1+1 is a
 *                       parameter (reference: 2.2F).
 *   oracle_bm25_weights standalone/large_corpus.py:148 (idf) and :174 (norm),
 *                       evaluated in double then rounded to FP32.
 *   oracle_contribution cpu_index_search.h:185-186.
 *   oracle_bm25s_search the bm25s-numba retrieve algorithm, restated:
 *                       reference_only/bm25s_original/bm25s/numba/retrieve_utils.py:13-60
 *                       (dense FP32 row per query, scatter-add of precomputed
 *                       impacts in query-token order), scoring.py:345-368, and
 *                       numba/selection.py:31-124 (bounded min-heap top-k, then a
 *                       descending sort). Used as a timed baseline only; ties in
 *                       it follow this file's strict order, upstream's tie order
 *                       is unpinned.
 *
 * Build flags keep FP32 semantics: -ffp-contract=off -fno-fast-math
 * (FR/CMakeLists.txt:22-24).
 *
 * Parity pins (see tests/test_oracle.py): the reference's own known-answer test
 * corpora and independent per-document FP32 oracle of
 * FR/ut/cases/bm25/cpu_index_search_tests.cpp:22-134 restated in Python, a
 * brute-force full sort on small inputs, and differential runs against the
 * reference's own CPU Search compiled from its sources into oracle/_ref.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float score;
    int32_t id;
} cand_t;

/* Better(a, b): a ranks before b (topk_heap.h:11-16). */
static inline int better(cand_t a, cand_t b) { return a.score > b.score || (a.score == b.score && a.id < b.id); }

/* Bounded heap whose root is the worst kept candidate. */
static inline void heap_sift_down(cand_t *h, int64_t size, int64_t pos)
{
    cand_t item = h[pos];
    for (;;) {
        int64_t child = 2 * pos + 1;
        if (child >= size)
            break;
        if (child + 1 < size && better(h[child], h[child + 1]))
            child = child + 1; /* pick the worse child */
        if (better(item, h[child])) {
            h[pos] = h[child];
            pos = child;
        } else {
            break;
        }
    }
    h[pos] = item;
}

static inline void heap_sift_up(cand_t *h, int64_t pos)
{
    cand_t item = h[pos];
    while (pos > 0) {
        int64_t parent = (pos - 1) / 2;
        if (better(h[parent], item)) {
            h[pos] = h[parent];
            pos = parent;
        } else {
            break;
        }
    }
    h[pos] = item;
}

/* KeepCandidate (topk_heap.h:18-28) for any K. */
static inline void keep(cand_t *h, int64_t *size, int64_t k, cand_t item)
{
    if (*size < k) {
        h[*size] = item;
        heap_sift_up(h, *size);
        ++*size;
    } else if (better(item, h[0])) {
        h[0] = item;
        heap_sift_down(h, *size, 0);
    }
}

static int cmp_better(const void *a, const void *b)
{
    const cand_t x = *(const cand_t *)a, y = *(const cand_t *)b;
    if (better(x, y))
        return -1;
    if (better(y, x))
        return 1;
    return 0;
}

float oracle_contribution(float idf, float tf, float norm, float kp1)
{
    const float numerator = tf * kp1;
    return (idf * numerator) / (tf + norm);
}

void oracle_bm25_weights(const int64_t *offsets, int64_t vocab, const int32_t *doc_len, int64_t ndocs, double k1,
                         double b, float *out_idf, float *out_norm)
{
    double total = 0;
    for (int64_t d = 0; d < ndocs; ++d)
        total += (double)doc_len[d];
    const double n = (double)ndocs, average = total / n;
    for (int64_t t = 0; t < vocab; ++t) {
        const double df = (double)(offsets[t + 1] - offsets[t]);
        out_idf[t] = (float)log1p((n - df + .5) / (df + .5));
    }
    for (int64_t d = 0; d < ndocs; ++d)
        out_norm[d] = (float)(k1 * (1 - b + b * (double)doc_len[d] / average));
}

/* std::lower_bound over ids[lo, hi). */
static inline int64_t lower_bound_i32(const int32_t *ids, int64_t lo, int64_t hi, int64_t value)
{
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((int64_t)ids[mid] < value)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/*
 * Full Search of queries [0, nq). Results of query q start at row rows[q]
 * (prefix of k[]). full_scores, when non-NULL, receives the dense FP32 score
 * rows (nq * ndocs). Queries are processed in chunks of at most chunk_q rows
 * of dense scores so memory stays bounded. Returns 0, or -1 on bad input /
 * allocation failure.
 */
int oracle_search(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf, const float *norm,
                  int64_t ndocs, const float *idf, float kp1, const int64_t *qoff, const int32_t *qterms, int32_t nq,
                  const int32_t *k, int32_t shards, int32_t threads, int32_t chunk_q, float *out_scores,
                  int32_t *out_ids, float *full_scores)
{
    if (nq < 1 || shards < 1 || ndocs < 1 || vocab < 1)
        return -1;
    if (shards > ndocs)
        shards = (int32_t)ndocs;
    if (threads > 0) {
        omp_set_dynamic(0);
        omp_set_num_threads(threads);
    }
    if (chunk_q < 1)
        chunk_q = 1;
    int64_t kmax = 0;
    int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * (nq + 1));
    if (!rows)
        return -1;
    rows[0] = 0;
    for (int32_t q = 0; q < nq; ++q) {
        if (k[q] < 1 || k[q] > ndocs) {
            free(rows);
            return -1;
        }
        kmax = k[q] > kmax ? k[q] : kmax;
        rows[q + 1] = rows[q] + k[q];
    }
    const int32_t cq = chunk_q < nq ? chunk_q : nq;
    float *scores = full_scores ? NULL : (float *)malloc(sizeof(float) * (size_t)cq * (size_t)ndocs);
    cand_t *partials = (cand_t *)malloc(sizeof(cand_t) * (size_t)cq * (size_t)shards * (size_t)kmax);
    int64_t *psize = (int64_t *)malloc(sizeof(int64_t) * (size_t)cq * (size_t)shards);
    if ((!full_scores && !scores) || !partials || !psize) {
        free(scores);
        free(partials);
        free(psize);
        free(rows);
        return -1;
    }
    for (int32_t q0 = 0; q0 < nq; q0 += cq) {
        const int32_t qn = (nq - q0) < cq ? (nq - q0) : cq;
        float *base = full_scores ? full_scores + (size_t)q0 * (size_t)ndocs : scores;
        const int64_t tasks = (int64_t)qn * shards;
        /* Scoring: cpu_index_search.h:170-190. */
#pragma omp parallel for schedule(static)
        for (int64_t task = 0; task < tasks; ++task) {
            const int64_t ql = task / shards, shard = task % shards, q = q0 + ql;
            const int64_t begin = ndocs * shard / shards, end = ndocs * (shard + 1) / shards;
            float *row = base + ql * ndocs;
            for (int64_t d = begin; d < end; ++d)
                row[d] = 0.0F;
            for (int64_t w = qoff[q]; w < qoff[q + 1]; ++w) {
                const int32_t term = qterms[w];
                if (term < 0 || (int64_t)term >= vocab)
                    continue;
                const int64_t lo = lower_bound_i32(docs, offsets[term], offsets[term + 1], begin);
                const int64_t hi = lower_bound_i32(docs, lo, offsets[term + 1], end);
                for (int64_t p = lo; p < hi; ++p) {
                    const int32_t doc = docs[p];
                    const float numerator = tf[p] * kp1;
                    const float contribution = (idf[term] * numerator) / (tf[p] + norm[doc]);
                    row[doc] += contribution;
                }
            }
        }
        /* Selection: topk_document_shards.h:13-47. */
#pragma omp parallel for schedule(static)
        for (int64_t task = 0; task < tasks; ++task) {
            const int64_t ql = task / shards, shard = task % shards, q = q0 + ql;
            const int64_t begin = ndocs * shard / shards, end = ndocs * (shard + 1) / shards;
            const float *row = base + ql * ndocs;
            cand_t *heap = partials + task * kmax;
            int64_t size = 0;
            for (int64_t d = begin; d < end; ++d) {
                cand_t c = {row[d], (int32_t)d};
                keep(heap, &size, k[q], c);
            }
            psize[task] = size;
        }
#pragma omp parallel for schedule(static)
        for (int64_t ql = 0; ql < qn; ++ql) {
            const int64_t q = q0 + ql;
            cand_t *heap = (cand_t *)malloc(sizeof(cand_t) * (size_t)k[q]);
            int64_t size = 0;
            if (heap) {
                for (int64_t s = 0; s < shards; ++s) {
                    const cand_t *part = partials + (ql * shards + s) * kmax;
                    for (int64_t i = 0; i < psize[ql * shards + s]; ++i)
                        keep(heap, &size, k[q], part[i]);
                }
                qsort(heap, (size_t)size, sizeof(cand_t), cmp_better);
                for (int64_t r = 0; r < k[q]; ++r) {
                    out_scores[rows[q] + r] = r < size ? heap[r].score : -1.0F;
                    out_ids[rows[q] + r] = r < size ? heap[r].id : -1;
                }
                free(heap);
            }
        }
    }
    free(scores);
    free(partials);
    free(psize);
    free(rows);
    return 0;
}

/* bm25s-numba retrieve algorithm over precomputed impacts (see header). */
int oracle_bm25s_search(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *impact,
                        int64_t ndocs, const int64_t *qoff, const int32_t *qterms, int32_t nq, int32_t k,
                        int32_t threads, float *out_scores, int32_t *out_ids)
{
    if (nq < 1 || k < 1 || k > ndocs)
        return -1;
    if (threads > 0) {
        omp_set_dynamic(0);
        omp_set_num_threads(threads);
    }
    int failed = 0;
#pragma omp parallel
    {
        float *row = (float *)malloc(sizeof(float) * (size_t)ndocs);
        cand_t *heap = (cand_t *)malloc(sizeof(cand_t) * (size_t)k);
        if (!row || !heap) {
#pragma omp atomic write
            failed = 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int32_t q = 0; q < nq; ++q) {
                memset(row, 0, sizeof(float) * (size_t)ndocs);
                for (int64_t w = qoff[q]; w < qoff[q + 1]; ++w) {
                    const int32_t term = qterms[w];
                    if (term < 0 || (int64_t)term >= vocab)
                        continue;
                    for (int64_t p = offsets[term]; p < offsets[term + 1]; ++p)
                        row[docs[p]] += impact[p];
                }
                int64_t size = 0;
                for (int64_t d = 0; d < ndocs; ++d) {
                    cand_t c = {row[d], (int32_t)d};
                    keep(heap, &size, k, c);
                }
                qsort(heap, (size_t)size, sizeof(cand_t), cmp_better);
                for (int64_t r = 0; r < k; ++r) {
                    out_scores[(int64_t)q * k + r] = heap[r].score;
                    out_ids[(int64_t)q * k + r] = heap[r].id;
                }
            }
        }
        free(row);
        free(heap);
    }
    return failed ? -1 : 0;
}
