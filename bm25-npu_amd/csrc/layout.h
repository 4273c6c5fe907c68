// Host-side description of the resident index layout.
#pragma once

#include <cstdint>
#include <vector>

namespace bm25mi {

// Block directory for long posting lists (df >= longMin). For long term t with
// slot s = termSlot[t]: dir[s*(granules+1) + g] is the number of postings of t
// whose doc ID is below g*granule (relative to the term's first posting), and
// blockMax[s*granules + g] the largest impact of t among docs of granule g.
// termMax[t] is the largest impact of any term. Short lists are located by an
// in-kernel 64-ary search instead, so the directory costs O(long terms x
// granules) rather than the reference's O(vocab x leaves)
// (FR/common/bm25/shard_directory.h:19).
struct Layout {
    int32_t granule = 1024;
    int32_t longMin = 4096;
    int64_t granules = 0;
    int64_t longTerms = 0;
    std::vector<int32_t> termSlot;
    std::vector<float> termMax;
    std::vector<float> termMean; // mean impact of each term (0 for empty lists)
    std::vector<uint32_t> dir;
    std::vector<float> blockMax;
    // The largest impacts of each term in descending order: min(df, 1024) of them for long
    // lists, min(df, 64) for short ones, term t at topVal[topOff[t] ...]. A doc's score is
    // at least each of its impacts (FP32 sums of non-negative values never decrease), so the
    // K-th largest impact of any query term is a lower bound of the query's K-th score.
    std::vector<int64_t> topOff;
    std::vector<float> topVal;
};

constexpr int32_t kTopLong = 1024;
constexpr int32_t kTopShort = 64;

void ValidateIndex(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf, int64_t postings,
                   const float *norm, int64_t ndocs, const float *idf, int threads);
float KPlusOne(double k1);
void ComputeImpacts(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf, const float *norm,
                    const float *idf, double k1, float *out, int threads);
void BuildLayout(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *impact, int64_t ndocs,
                 int32_t granule, int32_t longMin, int threads, Layout &out);

} // namespace bm25mi
