// Index validation, FP32 impact precompute and the device-side block layout
// (host C++).
//
// Validation follows FR/common/bm25/index_data.h:13-41 rule for rule. The
// impact of a posting is the reference's per-query contribution,
// (idf[t] * (tf * (k1+1)f)) / (tf + norm[d]) in FP32
// (FR/test/ascendhost/bm25/reference/cpu_index_search.h:185-186), computed once
// here instead of once per posting per query. Each operation is a single FP32
// rounding, so the value is bit-identical to the reference's.
#include <algorithm>
#include <cmath>
#include <functional>

#include "common.h"
#include "layout.h"

namespace bm25mi {

void ValidateIndex(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf, int64_t postings,
                   const float *norm, int64_t ndocs, const float *idf, int threads)
{
    Require(offsets && norm && idf && (postings == 0 || (docs && tf)), "Null index array");
    Require(ndocs >= 1 && ndocs <= INT32_MAX && vocab >= 1 && vocab < INT32_MAX && postings >= 0 &&
                postings <= INT32_MAX,
            "Empty or oversized index");
    Require(offsets[0] == 0 && offsets[vocab] == postings, "Index dimensions/endpoints differ");
    for (int64_t d = 0; d < ndocs; ++d)
        Require(std::isfinite(norm[d]) && norm[d] > 0, "Invalid article norm");
    std::vector<std::string> errors(Threads(threads));
    ParallelChunks(vocab, Threads(threads), [&](int c, int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi && errors[c].empty(); ++t) {
            if (!(std::isfinite(idf[t]) && idf[t] >= 0)) {
                errors[c] = "Invalid IDF";
                break;
            }
            const int64_t b = offsets[t], e = offsets[t + 1];
            if (!(b >= 0 && b <= e && e <= postings)) {
                errors[c] = "Invalid posting bounds";
                break;
            }
            for (int64_t p = b; p < e; ++p) {
                if (!(docs[p] >= 0 && docs[p] < ndocs && (p == b || docs[p - 1] < docs[p]))) {
                    errors[c] = "Unsorted/duplicate/out-of-range posting ID";
                    break;
                }
                if (!(std::isfinite(tf[p]) && tf[p] >= 1 && tf[p] <= 16777216.0F && std::floor(tf[p]) == tf[p])) {
                    errors[c] = "Invalid posting TF";
                    break;
                }
            }
        }
    });
    for (auto &e : errors)
        Require(e.empty(), e);
}

// FP32 operations on volatile operands so no host compiler can contract or
// reassociate them; each line is exactly one IEEE binary32 rounding.
static inline float Impact(float idf, float tf, float norm, float kp1)
{
    volatile float numerator = tf * kp1;
    volatile float top = idf * numerator;
    volatile float bottom = tf + norm;
    volatile float value = top / bottom;
    return value;
}

float KPlusOne(double k1) { return static_cast<float>(k1 + 1.0); }

void ComputeImpacts(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *tf, const float *norm,
                    const float *idf, double k1, float *out, int threads)
{
    const float kp1 = KPlusOne(k1);
    ParallelChunks(vocab, Threads(threads), [&](int, int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t)
            for (int64_t p = offsets[t]; p < offsets[t + 1]; ++p)
                out[p] = Impact(idf[t], tf[p], norm[docs[p]], kp1);
    });
}

void BuildLayout(const int64_t *offsets, int64_t vocab, const int32_t *docs, const float *impact, int64_t ndocs,
                 int32_t granule, int32_t longMin, int threads, Layout &out)
{
    Require(granule >= 64 && (granule & (granule - 1)) == 0, "Directory granule must be a power of two >= 64");
    out.granule = granule;
    out.longMin = longMin;
    out.granules = (ndocs + granule - 1) / granule;
    out.termSlot.assign(vocab, -1);
    out.termMax.assign(vocab, 0.0F);
    out.termMean.assign(vocab, 0.0F);
    int64_t longTerms = 0;
    for (int64_t t = 0; t < vocab; ++t)
        if (offsets[t + 1] - offsets[t] >= longMin)
            out.termSlot[t] = static_cast<int32_t>(longTerms++);
    out.longTerms = longTerms;
    const int64_t stride = out.granules + 1;
    out.dir.assign(static_cast<size_t>(longTerms * stride), 0);
    out.blockMax.assign(static_cast<size_t>(longTerms * out.granules), 0.0F);
    const int shift = __builtin_ctz(static_cast<unsigned>(granule));
    ParallelChunks(vocab, Threads(threads), [&](int, int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t) {
            const int64_t b = offsets[t], e = offsets[t + 1];
            float tmax = 0.0F;
            double sum = 0.0;
            for (int64_t p = b; p < e; ++p) {
                tmax = std::max(tmax, impact[p]);
                sum += impact[p];
            }
            out.termMax[t] = tmax;
            out.termMean[t] = e > b ? static_cast<float>(sum / static_cast<double>(e - b)) : 0.0F;
            const int32_t slot = out.termSlot[t];
            if (slot < 0)
                continue;
            uint32_t *dir = out.dir.data() + slot * stride;
            float *bmax = out.blockMax.data() + slot * out.granules;
            // dir[g] = postings of t with doc < g * granule, relative to the term start.
            int64_t g = 0;
            for (int64_t p = b; p < e; ++p) {
                const int64_t dg = docs[p] >> shift;
                while (g <= dg)
                    dir[g++] = static_cast<uint32_t>(p - b);
                bmax[dg] = std::max(bmax[dg], impact[p]);
            }
            while (g <= out.granules)
                dir[g++] = static_cast<uint32_t>(e - b);
        }
    });
    out.topOff.assign(static_cast<size_t>(vocab + 1), 0);
    for (int64_t t = 0; t < vocab; ++t) {
        const int64_t df = offsets[t + 1] - offsets[t];
        out.topOff[t + 1] = out.topOff[t] + std::min<int64_t>(df, df >= longMin ? kTopLong : kTopShort);
    }
    out.topVal.assign(static_cast<size_t>(out.topOff[vocab]), 0.0F);
    ParallelChunks(vocab, Threads(threads), [&](int, int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t) {
            float *top = out.topVal.data() + out.topOff[t];
            const int64_t r = out.topOff[t + 1] - out.topOff[t];
            std::partial_sort_copy(impact + offsets[t], impact + offsets[t + 1], top, top + r, std::greater<float>());
        }
    });
}

} // namespace bm25mi

using namespace bm25mi;

extern "C" int bm25mi_validate_index(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids,
                                     const float *tf, int64_t postings, const float *doc_norm, int64_t docs,
                                     const float *idf)
{
    return Invoke([&] { ValidateIndex(term_offsets, vocab, doc_ids, tf, postings, doc_norm, docs, idf, 0); });
}

extern "C" int bm25mi_impacts(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids, const float *tf,
                              const float *doc_norm, const float *idf, double k1, float *out_impact, int32_t threads)
{
    return Invoke([&] {
        Require(term_offsets && idf && doc_norm && out_impact, "Null impact arrays");
        Require(vocab >= 1 && std::isfinite(k1) && k1 > 0, "Invalid vocab or k1");
        ComputeImpacts(term_offsets, vocab, doc_ids, tf, doc_norm, idf, k1, out_impact, threads);
    });
}
