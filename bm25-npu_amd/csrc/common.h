// Shared host-side helpers of libbm25mi: error classes, thread-local last error,
// a deterministic parallel-for and checked byte sizes.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bm25mi.h"

namespace bm25mi {

// Error classes of FR/common/bm25/error.h:8-31, mapped to status codes the way
// FR/ascendhost/include/impl/Bm25ErrorAdapter.h:11-45 maps them to APP_ERR.
struct InvalidArgument : std::runtime_error {
    explicit InvalidArgument(const std::string &m) : std::runtime_error(m) {}
};
struct InvalidState : std::runtime_error {
    explicit InvalidState(const std::string &m) : std::runtime_error(m) {}
};
struct DeviceError : std::runtime_error {
    int64_t code;
    DeviceError(int64_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void Require(bool ok, const std::string &message)
{
    if (!ok)
        throw InvalidArgument(message);
}

// Last error of the calling thread; sessions keep their own copy as well.
struct ErrorInfo {
    int64_t code = 0;
    char message[1024] = {0};
    void Set(const char *text, int64_t c)
    {
        code = c;
        std::memset(message, 0, sizeof(message));
        if (text)
            std::strncpy(message, text, sizeof(message) - 1);
    }
};
ErrorInfo &ThreadError();

// Runs op, converts exceptions to a status, records the message in *info and
// in the thread error. A success clears both.
template <class Op> int Invoke(Op &&op, ErrorInfo *info = nullptr) noexcept
{
    auto set = [&](const char *m, int64_t c) {
        ThreadError().Set(m, c);
        if (info)
            info->Set(m, c);
    };
    set(nullptr, 0);
    try {
        op();
        return BM25MI_OK;
    } catch (const InvalidArgument &e) {
        set(e.what(), 0);
        return BM25MI_ERR_INVALID_PARAM;
    } catch (const InvalidState &e) {
        set(e.what(), 0);
        return BM25MI_ERR_ILLEGAL_OPERATION;
    } catch (const DeviceError &e) {
        set(e.what(), e.code);
        return BM25MI_ERR_INNER;
    } catch (const std::exception &e) {
        set(e.what(), 0);
        return BM25MI_ERR_INNER;
    } catch (...) {
        set("Unknown BM25 failure", 0);
        return BM25MI_ERR_INNER;
    }
}

inline int Threads(int32_t requested)
{
    if (requested > 0)
        return requested;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? static_cast<int>(hw > 64 ? 64 : hw) : 1;
}

// Static partition of [0, n) into `threads` contiguous chunks; fn(chunk, begin, end).
template <class Fn> void ParallelChunks(int64_t n, int threads, Fn &&fn)
{
    if (threads <= 1 || n < 2) {
        fn(0, int64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> errors(threads);
    for (int t = 0; t < threads; ++t) {
        const int64_t b = n * t / threads, e = n * (t + 1) / threads;
        pool.emplace_back([&, t, b, e] {
            try {
                fn(t, b, e);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        });
    }
    for (auto &th : pool)
        th.join();
    for (auto &err : errors)
        if (err)
            std::rethrow_exception(err);
}

} // namespace bm25mi
