// Checks the C++ facade (include/bm25mi_index.hpp) the way the reference's SDK facade test
// checks its own (FR/ut/cases/bm25/sdk_facade_tests.cpp:29-62, FR = feature_retrieval/src/
// ascendfaiss): state machine, immutable index snapshot, output replaced only on success,
// error classes and messages, Close idempotence.
//   facade_check cpu   checks that need no GPU (run anywhere)
//   facade_check gpu   the full contract on cuda device 0
// Prints PASS and exits 0, or names the first failed check and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "bm25mi_index.hpp"

using namespace bm25mi;

static int failures = 0;
#define CHECK(cond, what)                                                                                              \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            std::fprintf(stderr, "FAIL: %s (line %d)\n", what, __LINE__);                                              \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

// Three terms over five docs: term 0 in docs {0, 2, 4}, term 1 in {1, 2}, term 2 in {3}.
static Index Tiny()
{
    Index ix;
    ix.offsets = {0, 3, 5, 6};
    ix.docs = {0, 2, 4, 1, 2, 3};
    ix.tf = {1, 2, 1, 3, 1, 1};
    ix.norm = {1.0F, 0.9F, 1.1F, 1.3F, 0.8F};
    ix.idf = {0.5F, 1.5F, 2.0F};
    return ix;
}

// The reference's per-posting contribution at k1 = 1.2 (cpu_index_search.h:185-186).
static float Contribution(const Index &ix, int term, int p)
{
    const float tf = ix.tf[p];
    return (ix.idf[term] * (tf * 2.2F)) / (tf + ix.norm[ix.docs[p]]);
}

int main(int argc, char **argv)
{
    const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
    {
        AscendIndexBm25 model;
        Bm25SearchResult out{{9.0F}, {9}};
        CHECK(model.Search({{0}}, out) == APP_ERR_ILLEGAL_OPERATION, "Search before Init accepted");
        CHECK(std::string(model.GetLastError().message.data()) == "BM25 is not initialized", "not-initialized message");
        CHECK(out.scores[0] == 9.0F && out.labels[0] == 9, "rejected Search changed the output");
        Index broken = Tiny();
        broken.tf[0] = 1.5F; // tf must be integral
        Bm25Config cfg;
        cfg.queryCount = 2;
        cfg.maxQueryWords = 8;
        cfg.topK = 3;
        CHECK(model.Init(broken, cfg) == APP_ERR_INVALID_PARAM, "invalid index accepted");
        CHECK(std::strlen(model.GetLastError().message.data()) > 0, "invalid index left no message");
        broken = Tiny();
        broken.offsets.pop_back();
        CHECK(model.Init(broken, cfg) == APP_ERR_INVALID_PARAM, "index with bad offsets accepted");
        cfg.topK = 2000;
        CHECK(model.Init(Tiny(), cfg) == APP_ERR_INVALID_PARAM, "topK above the limit accepted");
        CHECK(model.Close() == APP_ERR_OK && model.Close() == APP_ERR_OK, "Close is not idempotent");
        CHECK(model.GetLastError().message[0] == '\0', "success did not clear the message");
    }
    if (gpu) {
        AscendIndexBm25 model;
        Index ix = Tiny();
        Bm25Config cfg;
        cfg.queryCount = 2;
        cfg.maxQueryWords = 8;
        cfg.topK = 3;
        CHECK(model.Init(ix, cfg) == APP_ERR_OK, "Init failed");
        CHECK(model.Init(ix, cfg) == APP_ERR_ILLEGAL_OPERATION, "double Init replaced the active state");
        ix.idf[0] = 99.0F; // the facade keeps its own snapshot
        Bm25SearchResult out{{7.0F}, {7}};
        CHECK(model.Search({{-1}, {0}}, out) == APP_ERR_INVALID_PARAM, "negative ID accepted");
        CHECK(model.Search({{0}}, out) == APP_ERR_INVALID_PARAM, "wrong query count accepted");
        CHECK(model.Search({{0, 1, 2, 0, 1, 2, 0, 1, 2}, {0}}, out) == APP_ERR_INVALID_PARAM, "word capacity exceeded");
        CHECK(out.scores.size() == 1 && out.scores[0] == 7.0F, "rejected query changed the output");
        CHECK(model.Search({{0, 1}, {2, 7}}, out) == APP_ERR_OK, "valid Search failed after rejections");
        const Index t = Tiny();
        // Query 0 = terms 0, 1 (in order); query 1 = term 2 plus an unknown term (skipped).
        float s0[5] = {0, 0, 0, 0, 0};
        for (int term : {0, 1})
            for (int64_t p = t.offsets[term]; p < t.offsets[term + 1]; ++p)
                s0[t.docs[p]] = s0[t.docs[p]] + Contribution(t, term, static_cast<int>(p));
        // Strict order (score desc, doc asc): doc 1 (tf 3 of term 1), doc 2 (both terms), doc 4.
        const int want0[3] = {1, 2, 4};
        const float c5 = Contribution(t, 2, 5);
        CHECK(out.labels.size() == 6, "result shape");
        bool exact = true;
        for (int r = 0; r < 3; ++r)
            exact = exact && out.labels[r] == want0[r] && out.scores[r] == s0[want0[r]];
        CHECK(s0[1] > s0[2] && s0[2] > s0[4] && s0[4] > s0[0], "hand-computed order assumption");
        CHECK(exact, "query 0 differs from the hand-computed scores");
        CHECK(out.labels[3] == 3 && out.scores[3] == c5, "query 1 top doc");
        CHECK(out.labels[4] == 0 && out.scores[4] == 0.0F && out.labels[5] == 1 && out.scores[5] == 0.0F,
              "zero-score padding by lowest doc ID");
        CHECK(model.Close() == APP_ERR_OK, "Close failed");
        CHECK(model.Search({{0}, {1}}, out) == APP_ERR_ILLEGAL_OPERATION, "Search after Close accepted");
        CHECK(model.Close() == APP_ERR_OK, "second Close failed");
        CHECK(model.Init(Tiny(), cfg) == APP_ERR_OK, "Init after Close failed");
        CHECK(model.Search({{}, {}}, out) == APP_ERR_OK && out.labels[0] == 0 && out.labels[1] == 1 &&
                  out.labels[2] == 2 && out.scores[0] == 0.0F,
              "empty queries return IDs 0..K-1");
    }
    if (failures)
        return 1;
    std::printf("PASS: facade %s checks\n", gpu ? "cpu+gpu" : "cpu");
    return 0;
}
