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
