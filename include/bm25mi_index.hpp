// bm25mi_index.hpp - C++ facade with the shape of the reference's IndexSDK BM25 index,
// header-only over the C ABI of bm25mi.h (link with libbm25mi.so).
//
// Mirrors faiss::ascend::AscendIndexBm25 (FR = feature_retrieval/src/ascendfaiss):
//   FR/ascendhost/include/index/AscendIndexBm25.h:17-57 (Bm25Config, Bm25SearchResult, the
//   class) and FR/ascendhost/src/bm25/AscendIndexBm25.cpp:14-84 (behaviour):
//   * one mutex per object serialises Init / Search / Close / GetLastError;
//   * Init copies and re-validates the index (the library keeps its own snapshot, so later
//     changes to the caller's arrays have no effect); a second Init without Close is
//     APP_ERR_ILLEGAL_OPERATION; a failed Init leaves the object uninitialised;
//   * Search requires exactly the declared query count (FR/common/bm25/fixed_query_packet.h:28),
//     at most 64 terms per query and maxQueryWords in total, IDs in [0, INT32_MAX); it
//     replaces `result` only on success, with int64 labels; input errors keep the session
//     usable, a device error makes later Searches APP_ERR_ILLEGAL_OPERATION;
//   * Close is idempotent and may be retried after a failure;
//   * GetLastError returns {aclCode, message} (FR/common/bm25/error_info.h:8-20); the code is
//     the HIP error of a device failure, the message is cut to 1023 characters and cleared
//     by every successful call.
// Status classes follow FR/ascendhost/include/impl/Bm25ErrorAdapter.h:11-45. The numeric values
// are this library's (bm25mi.h), since IndexSDK's ErrorCode.h is not part of this library.
// Differences by design: no SDK context/resources/vendor directory (the HIP device is named by
// deviceId); topK may be up to 1024 (reference: 32); k1 and the scoring block size are config.
#ifndef BM25MI_INDEX_HPP
#define BM25MI_INDEX_HPP

#include <array>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "bm25mi.h"

namespace bm25mi {

using APP_ERROR = int;
constexpr APP_ERROR APP_ERR_OK = BM25MI_OK;
constexpr APP_ERROR APP_ERR_INVALID_PARAM = BM25MI_ERR_INVALID_PARAM;
constexpr APP_ERROR APP_ERR_ILLEGAL_OPERATION = BM25MI_ERR_ILLEGAL_OPERATION;
constexpr APP_ERROR APP_ERR_INNER_ERROR = BM25MI_ERR_INNER;

// ascend::bm25::Index (FR/common/bm25/index_data.h:9-42): the five resident arrays.
struct Index {
    std::vector<int64_t> offsets; // [vocab + 1]
    std::vector<int32_t> docs;    // [postings], ascending within each term
    std::vector<float> tf;        // [postings], integral in [1, 2^24]
    std::vector<float> norm;      // [docs]
    std::vector<float> idf;       // [vocab]
};

using Queries = std::vector<std::vector<int32_t>>; // bm25search::Queries (search_inputs.h:12)

struct Bm25Config {
    int deviceId = 0;
    uint64_t queryCount = 8;
    uint64_t maxQueryWords = 24;
    uint64_t topK = 10;
    double k1 = 1.2;     // the reference's constant (2.2F = k1 + 1 in cpu_index_search.h:185)
    int32_t blockDocs = 0; // docs per scoring task; 0 = library default
};

struct Bm25SearchResult {
    std::vector<float> scores;   // row-major [queryCount][topK]
    std::vector<int64_t> labels; // doc IDs, same layout
};

struct ErrorInfo {
    int64_t aclCode = 0;
    std::array<char, 1024> message{};
};

class AscendIndexBm25 {
public:
    AscendIndexBm25() = default;
    ~AscendIndexBm25() { (void)Close(); }
    AscendIndexBm25(const AscendIndexBm25 &) = delete;
    AscendIndexBm25 &operator=(const AscendIndexBm25 &) = delete;
    AscendIndexBm25(AscendIndexBm25 &&) = delete;
    AscendIndexBm25 &operator=(AscendIndexBm25 &&) = delete;

    APP_ERROR Init(const Index &index, const Bm25Config &config)
    {
        std::lock_guard<std::mutex> guard(mutex_);
        Clear();
        if (session_)
            return Fail(APP_ERR_ILLEGAL_OPERATION, "BM25 is already initialized; Close before Init", 0);
        if (config.queryCount < 1 || config.queryCount > BM25MI_MAX_QUERIES || config.topK < 1 ||
            config.topK > BM25MI_MAX_K || config.maxQueryWords > 16777216)
            return Fail(APP_ERR_INVALID_PARAM, "Invalid BM25 configuration", 0);
        if (index.idf.empty() || index.offsets.size() != index.idf.size() + 1 || index.docs.size() != index.tf.size())
            return Fail(APP_ERR_INVALID_PARAM, "Index dimensions/endpoints differ", 0);
        bm25mi_config c;
        bm25mi_config_default(&c);
        c.device_id = config.deviceId;
        c.max_queries = static_cast<int32_t>(config.queryCount);
        c.max_words = static_cast<int64_t>(config.maxQueryWords == 0 ? 1 : config.maxQueryWords);
        c.max_k = static_cast<int32_t>(config.topK);
        c.k1 = config.k1;
        c.block_docs = config.blockDocs;
        bm25mi_session *s = nullptr;
        const int rc = bm25mi_create(index.offsets.data(), static_cast<int64_t>(index.idf.size()), index.docs.data(),
                                     index.tf.data(), static_cast<int64_t>(index.docs.size()), index.norm.data(),
                                     static_cast<int64_t>(index.norm.size()), index.idf.data(), &c, &s);
        if (rc != BM25MI_OK)
            return FailFrom(rc, nullptr);
        session_ = s;
        queries_ = config.queryCount;
        words_ = config.maxQueryWords;
        k_ = config.topK;
        return APP_ERR_OK;
    }

    APP_ERROR Search(const Queries &queries, Bm25SearchResult &result)
    {
        std::lock_guard<std::mutex> guard(mutex_);
        Clear();
        if (!session_)
            return Fail(APP_ERR_ILLEGAL_OPERATION, "BM25 is not initialized", 0);
        if (queries.size() != queries_)
            return Fail(APP_ERR_INVALID_PARAM, "Fixed session requires its declared query count", 0);
        std::vector<int64_t> offsets(queries.size() + 1, 0);
        std::vector<int32_t> terms;
        uint64_t total = 0;
        for (size_t q = 0; q < queries.size(); ++q) {
            const auto &row = queries[q];
            if (row.size() > BM25MI_MAX_TERMS_PER_QUERY || row.size() > words_ - total)
                return Fail(APP_ERR_INVALID_PARAM, "Query exceeds fixed word/64-term capacity", 0);
            total += row.size();
            for (int32_t w : row)
                if (w < 0 || w == INT32_MAX)
                    return Fail(APP_ERR_INVALID_PARAM, "Invalid public word ID", 0);
            terms.insert(terms.end(), row.begin(), row.end());
            offsets[q + 1] = static_cast<int64_t>(terms.size());
        }
        std::vector<float> scores(queries.size() * k_);
        std::vector<int32_t> ids(scores.size());
        const int rc = bm25mi_search(session_, offsets.data(), terms.empty() ? nullptr : terms.data(),
                                     static_cast<int32_t>(queries.size()), static_cast<int32_t>(k_), scores.data(),
                                     ids.data());
        if (rc != BM25MI_OK)
            return FailFrom(rc, session_);
        Bm25SearchResult next;
        next.scores = std::move(scores);
        next.labels.assign(ids.begin(), ids.end());
        result = std::move(next);
        return APP_ERR_OK;
    }

    APP_ERROR Close()
    {
        std::lock_guard<std::mutex> guard(mutex_);
        Clear();
        if (!session_)
            return APP_ERR_OK;
        const int rc = bm25mi_destroy(session_);
        if (rc != BM25MI_OK)
            return FailFrom(rc, session_); // the session stays, so Close can be retried
        session_ = nullptr;
        return APP_ERR_OK;
    }

    ErrorInfo GetLastError() const
    {
        std::lock_guard<std::mutex> guard(mutex_);
        return error_;
    }

private:
    void Clear() { error_ = ErrorInfo{}; }
    APP_ERROR Fail(APP_ERROR code, const char *message, int64_t device)
    {
        error_.aclCode = device;
        error_.message.fill('\0');
        std::strncpy(error_.message.data(), message, error_.message.size() - 1);
        return code;
    }
    APP_ERROR FailFrom(int rc, const bm25mi_session *s)
    {
        char buf[1024] = {0};
        const int64_t device = bm25mi_last_error(s, buf, sizeof(buf));
        return Fail(rc, buf, device);
    }

    mutable std::mutex mutex_;
    bm25mi_session *session_ = nullptr;
    uint64_t queries_ = 0, words_ = 0, k_ = 0;
    ErrorInfo error_;
};

} // namespace bm25mi

#endif // BM25MI_INDEX_HPP
