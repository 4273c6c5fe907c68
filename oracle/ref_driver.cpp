// C entry point over the reference's own header-only CPU Search, compiled from
// the reference sources where they lie (FR/test/ascendhost/bm25/reference/
// cpu_index_search.h and the FR/common/bm25 headers it includes). TEST
// INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile, used by
// tests to check the C oracle bit for bit at the reference's constants
// (k1 = 1.2, K <= 32).
#include "test/ascendhost/bm25/reference/cpu_index_search.h"

#include <cstring>
#include <exception>

extern "C" int ref_cpu_search(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf,
                              int64_t postings, const float *norm, int64_t ndocs, const float *idf,
                              const int64_t *qoff, const int32_t *qterms, int32_t nq, int32_t k, int32_t threads,
                              int32_t shards, float *out_scores, int32_t *out_ids, float *out_full)
{
    try {
        bm25cpu::Index index(std::vector<int64_t>(offsets, offsets + vocab + 1),
                             std::vector<int32_t>(docs, docs + postings), std::vector<float>(tf, tf + postings),
                             std::vector<float>(norm, norm + ndocs), std::vector<float>(idf, idf + vocab));
        bm25cpu::Queries queries(nq);
        for (int32_t q = 0; q < nq; ++q)
            queries[q].assign(qterms + qoff[q], qterms + qoff[q + 1]);
        const size_t words = static_cast<size_t>(qoff[nq]);
        bm25cpu::CpuSearchSession session(index, threads, k, shards, static_cast<size_t>(nq), words);
        const auto result = session.Search(queries);
        std::memcpy(out_scores, result.scores.data(), result.scores.size() * sizeof(float));
        std::memcpy(out_ids, result.ids.data(), result.ids.size() * sizeof(int32_t));
        if (out_full) {
            const auto full = session.DebugScores(static_cast<size_t>(nq));
            std::memcpy(out_full, full.data(), full.size() * sizeof(float));
        }
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// Persistent sessions for timing (bench.py's CPU baseline): the index and the reference's
// CpuSearchSession (variant 0) or DirectoryCpuSearchSession (variant 1,
// FR/test/ascendhost/bm25/reference/cpu_directory_search.h) are built once; ref_cpu_run times
// only Search, as the reference's own benchmark does (benchmark_common.h:51-107).
#include "test/ascendhost/bm25/reference/cpu_directory_search.h"

#include <memory>

struct RefSession {
    bm25cpu::Index index;
    std::unique_ptr<bm25cpu::CpuSearchSession> plain;
    std::unique_ptr<bm25cpu::DirectoryCpuSearchSession> directory;
    RefSession(bm25cpu::Index i) : index(std::move(i)) {}
};

extern "C" void *ref_cpu_create(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf,
                                int64_t postings, const float *norm, int64_t ndocs, const float *idf, int32_t variant,
                                int32_t threads, int32_t shards, int32_t k, int32_t max_queries, int64_t max_words)
{
    try {
        auto s = std::make_unique<RefSession>(bm25cpu::Index(
            std::vector<int64_t>(offsets, offsets + vocab + 1), std::vector<int32_t>(docs, docs + postings),
            std::vector<float>(tf, tf + postings), std::vector<float>(norm, norm + ndocs),
            std::vector<float>(idf, idf + vocab)));
        if (variant == 0)
            s->plain = std::make_unique<bm25cpu::CpuSearchSession>(s->index, threads, k, shards, max_queries,
                                                                   static_cast<size_t>(max_words));
        else
            s->directory = std::make_unique<bm25cpu::DirectoryCpuSearchSession>(
                s->index, threads, k, shards, max_queries, static_cast<size_t>(max_words));
        return s.release();
    } catch (const std::exception &) {
        return nullptr;
    }
}

extern "C" int ref_cpu_run(void *handle, const int64_t *qoff, const int32_t *qterms, int32_t nq, float *out_scores,
                           int32_t *out_ids)
{
    try {
        auto *s = static_cast<RefSession *>(handle);
        bm25cpu::Queries queries(nq);
        for (int32_t q = 0; q < nq; ++q)
            queries[q].assign(qterms + qoff[q], qterms + qoff[q + 1]);
        const auto result = s->plain ? s->plain->Search(queries) : s->directory->Search(queries);
        std::memcpy(out_scores, result.scores.data(), result.scores.size() * sizeof(float));
        std::memcpy(out_ids, result.ids.data(), result.ids.size() * sizeof(int32_t));
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

extern "C" void ref_cpu_destroy(void *handle) { delete static_cast<RefSession *>(handle); }
