// Serving loop with continuous batching over a resident-index session (no reference
// counterpart: the reference makes no concurrency or service claims,
// docs/FINAL_DELIVERY.md:30).
//
// Requests (term IDs, K) are queued by any thread. One worker thread keeps up to
// `slots` batches in flight: whenever a slot is free and requests wait, it admits the
// oldest ones into a batch, up to max_batch queries, the session's word capacity and
// the slot's candidate-pool memory (each query needs PoolCapacity(K) entries), packs
// and launches it on the slot's stream without waiting. Requests that arrive while
// other batches run on the GPU are admitted as soon as a slot frees up, and their
// kernels overlap the running ones on the device. Finished slots are collected in
// launch order; each request gets its own rows. A query's result does not depend on the
// batch it ran in: every query's top-K is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "session.h"

namespace bm25mi {

using Clock = std::chrono::steady_clock;

struct Request {
    std::vector<int32_t> terms;
    int32_t k = 0;
    float thr = 0.0F;       // pruning threshold (0 = plain); a pruned run that keeps fewer than K
                            // docs is requeued at the head as a plain request
    bool learned = false;   // threshold predicted on the device by the session's model
    bool fellBack = false;
    Clock::time_point submitted, finished;
    std::vector<float> scores;
    std::vector<int32_t> ids;
    int status = BM25MI_OK; // BM25MI_ERR_INNER when the device failed
    std::string error;
    bool done = false;
    bool collected = false; // a Wait has claimed it
};

class Server {
public:
    Server(Session &s, int32_t maxBatch) : session_(s), maxBatch_(maxBatch)
    {
        Require(maxBatch >= 1 && maxBatch <= s.cfg.max_queries, "max_batch must be in [1, max_queries]");
        session_.serving = true;
        worker_ = std::thread([this] { Loop(); });
    }
    ~Server()
    {
        {
            std::lock_guard<std::mutex> g(mutex_);
            stop_ = true;
        }
        work_.notify_all();
        if (worker_.joinable())
            worker_.join();
        // Threads still inside Wait (woken by the last results) leave before the members go.
        {
            std::unique_lock<std::mutex> lock(mutex_);
            idle_.wait(lock, [&] { return waiters_ == 0; });
        }
        session_.serving = false;
    }

    uint64_t Submit(const int32_t *terms, int32_t n, int32_t k, float thr = 0.0F, bool learned = false)
    {
        Require(std::isfinite(thr) && thr >= 0.0F, "Threshold must be finite and >= 0");
        if (learned && !session_.dModel)
            throw InvalidState("Learned requests need a threshold model (bm25mi_set_threshold_model)");
        Require(n >= 0 && n <= kMaxTerms && n <= session_.cfg.max_words, "Query exceeds word capacity or 64 terms");
        Require(n == 0 || terms, "Null query terms");
        Require(k >= 1 && k <= session_.cfg.max_k && k <= session_.ndocs, "K out of range");
        Require(session_.PoolCapacity(k) <= session_.poolEntries, "K needs more pool memory than a slot holds");
        // A request the session would score on the plain path anyway (K above kLiveMaxK), or
        // whose threshold pool would not fit a slot, is queued as a plain request: admission
        // then reserves only the plain pool for it.
        if (k > session_.liveMaxK || session_.PoolCapacity(k, true) > session_.poolEntries) {
            thr = 0.0F;
            learned = false;
        }
        for (int32_t i = 0; i < n; ++i)
            Require(terms[i] >= 0 && terms[i] < INT32_MAX, "Invalid public word ID");
        auto req = std::make_unique<Request>();
        req->terms.assign(terms, terms + n);
        req->k = k;
        req->thr = thr;
        req->learned = learned;
        uint64_t ticket;
        {
            std::lock_guard<std::mutex> g(mutex_);
            Require(!stop_ && !failed_, failed_ ? "Server stopped after a device failure" : "Server is stopping");
            ticket = next_++;
            req->submitted = Clock::now();
            queue_.push_back(ticket);
            requests_.emplace(ticket, std::move(req));
        }
        work_.notify_one();
        return ticket;
    }

    // Blocks until the request is done; copies its K rows and removes it. The server is not
    // destroyed while a thread is inside (waiters_): the last member it touches is idle_,
    // notified with the lock held, before it copies the rows it now owns.
    void Wait(uint64_t ticket, float *scores, int32_t *ids, double *latencyUs)
    {
        std::unique_ptr<Request> req;
        {
            std::unique_lock<std::mutex> lock(mutex_);
            auto it = requests_.find(ticket);
            Require(it != requests_.end() && !it->second->collected, "Unknown or already collected ticket");
            // Wait on the request itself: map iterators do not survive inserts by Submit.
            Request *pending = it->second.get();
            pending->collected = true;
            ++waiters_;
            done_.wait(lock, [&] { return pending->done; });
            auto at = requests_.find(ticket);
            req = std::move(at->second);
            requests_.erase(at);
            if (--waiters_ == 0)
                idle_.notify_all();
        }
        if (req->status == BM25MI_ERR_INVALID_PARAM)
            throw InvalidArgument(req->error);
        if (req->status != BM25MI_OK)
            throw DeviceError(0, req->error);
        if (scores)
            std::memcpy(scores, req->scores.data(), req->scores.size() * 4);
        if (ids)
            std::memcpy(ids, req->ids.data(), req->ids.size() * 4);
        if (latencyUs)
            *latencyUs = std::chrono::duration<double, std::micro>(req->finished - req->submitted).count();
    }

    void Counters(int64_t *batches, int64_t *queries, int32_t *maxInflight, int64_t *fallbacks)
    {
        std::lock_guard<std::mutex> g(mutex_);
        if (batches)
            *batches = batches_;
        if (queries)
            *queries = queries_;
        if (maxInflight)
            *maxInflight = maxInflight_;
        if (fallbacks)
            *fallbacks = fallbacks_;
    }

private:
    struct InFlight {
        Slot *slot;
        std::vector<uint64_t> tickets;
        std::vector<int32_t> ks;
    };

    // Admits the oldest waiting requests into a batch for a free slot (called with the lock).
    // A batch holds learned and plain requests, or threshold and plain requests: a learned
    // request and one with a threshold of its own wait for different batches.
    bool TakeBatch(std::vector<uint64_t> &tickets, std::vector<int64_t> &offsets, std::vector<int32_t> &terms,
                   std::vector<int32_t> &ks, std::vector<float> &thr, std::vector<uint8_t> &predict, bool &learned)
    {
        tickets.clear();
        offsets.assign(1, 0);
        terms.clear();
        ks.clear();
        thr.clear();
        predict.clear();
        learned = false;
        bool withThr = false;
        int64_t pool = 0;
        while (!queue_.empty() && static_cast<int32_t>(tickets.size()) < maxBatch_) {
            const Request &r = *requests_.at(queue_.front());
            const int64_t need = session_.PoolCapacity(r.k, r.thr > 0.0F || r.learned);
            if (static_cast<int64_t>(terms.size() + r.terms.size()) > session_.cfg.max_words ||
                pool + need > session_.poolEntries)
                break; // memory admission: the rest waits for the next free slot
            if ((r.learned && withThr) || (r.thr > 0.0F && learned))
                break;
            learned = learned || r.learned;
            withThr = withThr || r.thr > 0.0F;
            predict.push_back(r.learned ? 1 : 0);
            pool += need;
            tickets.push_back(queue_.front());
            terms.insert(terms.end(), r.terms.begin(), r.terms.end());
            offsets.push_back(static_cast<int64_t>(terms.size()));
            ks.push_back(r.k);
            thr.push_back(r.thr);
            queue_.pop_front();
        }
        return !tickets.empty();
    }

    // Hands each request its rows; a pruned request with fewer than K survivors (status 1)
    // goes back to the head of the queue as a plain request instead.
    void Finish(InFlight &f, const std::vector<float> &scores, const std::vector<int32_t> &ids,
                const std::vector<int32_t> *status, int code, const std::string &error)
    {
        const auto now = Clock::now();
        std::lock_guard<std::mutex> g(mutex_);
        int64_t row = 0;
        for (size_t i = 0; i < f.tickets.size(); ++i) {
            Request &r = *requests_.at(f.tickets[i]);
            if (code == BM25MI_OK && status && (*status)[i]) {
                r.thr = 0.0F;
                r.learned = false;
                r.fellBack = true;
                ++fallbacks_;
                queue_.push_front(f.tickets[i]);
                row += f.ks[i];
                continue;
            }
            if (code == BM25MI_OK) {
                r.scores.assign(scores.begin() + row, scores.begin() + row + f.ks[i]);
                r.ids.assign(ids.begin() + row, ids.begin() + row + f.ks[i]);
            }
            row += f.ks[i];
            r.status = code;
            r.error = error;
            r.finished = now;
            r.done = true;
        }
    }

    void Loop()
    {
        std::vector<Slot *> free;
        for (auto &s : session_.slots)
            free.push_back(&s);
        std::deque<InFlight> inflight;
        std::vector<uint64_t> tickets;
        std::vector<int64_t> offsets;
        std::vector<int32_t> terms, ks, status;
        std::vector<float> scores, thr;
        std::vector<uint8_t> predict;
        std::vector<int32_t> ids;
        bool learned = false;
        for (;;) {
            bool launched = false;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                if (inflight.empty())
                    work_.wait(lock, [&] { return stop_ || !queue_.empty(); });
                if (failed_ && !queue_.empty()) {
                    // After a device failure nothing more runs: fail the waiting requests.
                    for (uint64_t t : queue_) {
                        Request &r = *requests_.at(t);
                        r.status = BM25MI_ERR_INNER;
                        r.error = "Server stopped after a device failure";
                        r.finished = Clock::now();
                        r.done = true;
                    }
                    queue_.clear();
                    done_.notify_all();
                }
                if (stop_ && queue_.empty() && inflight.empty())
                    return;
                if (queue_.empty() && inflight.empty())
                    continue;
                if (!free.empty() && !failed_ && TakeBatch(tickets, offsets, terms, ks, thr, predict, learned)) {
                    lock.unlock();
                    Slot *slot = free.back();
                    free.pop_back();
                    const bool anyThr = std::any_of(thr.begin(), thr.end(), [](float t) { return t > 0.0F; });
                    Query q{offsets.data(), terms.data(), static_cast<int32_t>(ks.size()), ks.data(),
                            anyThr ? thr.data() : nullptr};
                    q.learned = learned;
                    q.predict = learned ? predict.data() : nullptr;
                    InFlight f{slot, tickets, ks};
                    try {
                        session_.Pack(q, *slot);
                        session_.Launch(*slot, false);
                        inflight.push_back(std::move(f));
                        launched = true;
                        std::lock_guard<std::mutex> g(mutex_);
                        ++batches_;
                        queries_ += static_cast<int64_t>(ks.size());
                        maxInflight_ = std::max<int32_t>(maxInflight_, static_cast<int32_t>(inflight.size()));
                    } catch (const InvalidArgument &e) {
                        // A request the session refuses fails alone; the server keeps serving.
                        free.push_back(slot);
                        Fail(f, e.what(), false);
                    } catch (const std::exception &e) {
                        free.push_back(slot);
                        Fail(f, e.what());
                    }
                }
            }
            // Collect the oldest batch once it is done (poll while new work may be admitted).
            if (!inflight.empty()) {
                InFlight &f = inflight.front();
                bool ready = false;
                try {
                    ready = session_.Ready(*f.slot);
                } catch (const std::exception &e) {
                    Fail(f, e.what());
                    free.push_back(f.slot);
                    inflight.pop_front();
                    continue;
                }
                if (!ready) {
                    if (!launched)
                        std::this_thread::yield();
                    continue;
                }
                int64_t rows = 0;
                for (int32_t k : f.ks)
                    rows += k;
                scores.resize(static_cast<size_t>(rows));
                ids.resize(static_cast<size_t>(rows));
                status.resize(f.ks.size());
                try {
                    session_.Complete(*f.slot, scores.data(), ids.data(), status.data());
                    Finish(f, scores, ids, &status, BM25MI_OK, std::string());
                    work_.notify_one();
                } catch (const std::exception &e) {
                    Fail(f, e.what());
                }
                free.push_back(f.slot);
                inflight.pop_front();
                done_.notify_all();
            }
        }
    }

    // Fails the batch's requests. A device failure (device = true) also ends the server and the
    // session (large_search.cpp:208-210); a rejected request only fails its batch.
    void Fail(InFlight &f, const char *what, bool device = true)
    {
        if (device) {
            {
                std::lock_guard<std::mutex> g(mutex_);
                failed_ = true;
            }
            session_.failed = true;
        }
        Finish(f, {}, {}, nullptr, device ? BM25MI_ERR_INNER : BM25MI_ERR_INVALID_PARAM, what);
        done_.notify_all();
    }

    Session &session_;
    const int32_t maxBatch_;
    std::mutex mutex_;
    std::condition_variable work_, done_, idle_;
    int32_t waiters_ = 0; // threads inside Wait
    std::deque<uint64_t> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests_;
    uint64_t next_ = 1;
    bool stop_ = false, failed_ = false;
    int64_t batches_ = 0, queries_ = 0, fallbacks_ = 0;
    int32_t maxInflight_ = 0;
    std::thread worker_;
};

} // namespace bm25mi

using namespace bm25mi;

struct bm25mi_server {
    std::unique_ptr<Server> impl;
    bm25mi_session *session = nullptr;
};

extern "C" int bm25mi_server_start(bm25mi_session *s, int32_t max_batch, bm25mi_server **out)
{
    return Invoke([&] {
        Require(s && out, "Null session or server output");
        std::lock_guard<std::mutex> guard(s->mutex);
        Require(s->impl && !s->impl->closed && !s->impl->failed, "BM25 session closed/failed");
        Require(!s->impl->serving, "The session already has a server");
        auto srv = std::make_unique<bm25mi_server>();
        srv->session = s;
        srv->impl = std::make_unique<Server>(*s->impl, max_batch);
        *out = srv.release();
    });
}

extern "C" int bm25mi_server_submit(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k,
                                    uint64_t *ticket)
{
    return Invoke([&] {
        Require(srv && srv->impl && ticket, "Null server or ticket");
        *ticket = srv->impl->Submit(terms, n, k);
    });
}

extern "C" int bm25mi_server_submit_pruned(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k,
                                           float threshold, uint64_t *ticket)
{
    return Invoke([&] {
        Require(srv && srv->impl && ticket, "Null server or ticket");
        *ticket = srv->impl->Submit(terms, n, k, threshold);
    });
}

extern "C" int bm25mi_server_submit_learned(bm25mi_server *srv, const int32_t *terms, int32_t n, int32_t k,
                                            uint64_t *ticket)
{
    return Invoke([&] {
        Require(srv && srv->impl && ticket, "Null server or ticket");
        *ticket = srv->impl->Submit(terms, n, k, 0.0F, true);
    });
}

extern "C" int bm25mi_server_wait(bm25mi_server *srv, uint64_t ticket, float *scores, int32_t *ids,
                                  double *latency_us)
{
    return Invoke([&] {
        Require(srv && srv->impl, "Null server");
        srv->impl->Wait(ticket, scores, ids, latency_us);
    });
}

// Open-loop load: request i (terms offsets[i]..offsets[i+1], K ks[i]) is submitted at the
// i-th arrival of a Poisson process of rate qps (seeded), by this thread, and collected by a
// second one. latency_us[i] = submit to completion; scores/ids (optional) receive the rows
// of request i at offset sum_{j<i} ks[j].
extern "C" int bm25mi_server_run_poisson(bm25mi_server *srv, const int64_t *offsets, const int32_t *terms,
                                         int32_t n, const int32_t *ks, const float *thresholds, int32_t learned,
                                         double qps, uint64_t seed, double *latency_us, float *scores, int32_t *ids,
                                         double *elapsed_s)
{
    return Invoke([&] {
        Require(srv && srv->impl && offsets && ks && latency_us && n >= 1 && qps > 0, "Invalid Poisson run");
        Require(!(learned && thresholds), "A learned run predicts its own thresholds");
        std::vector<int64_t> rowOf(static_cast<size_t>(n) + 1, 0);
        for (int32_t i = 0; i < n; ++i)
            rowOf[i + 1] = rowOf[i] + ks[i];
        std::vector<uint64_t> tickets(static_cast<size_t>(n), 0);
        std::atomic<int32_t> submitted{0};
        std::string failure;
        std::mutex failMutex;
        std::thread collector([&] {
            for (int32_t i = 0; i < n; ++i) {
                while (submitted.load(std::memory_order_acquire) <= i)
                    std::this_thread::yield();
                if (!tickets[i])
                    continue;
                try {
                    srv->impl->Wait(tickets[i], scores ? scores + rowOf[i] : nullptr, ids ? ids + rowOf[i] : nullptr,
                                    latency_us + i);
                } catch (const std::exception &e) {
                    std::lock_guard<std::mutex> g(failMutex);
                    if (failure.empty())
                        failure = e.what();
                }
            }
        });
        std::mt19937_64 rng(seed);
        std::exponential_distribution<double> gap(qps);
        const auto start = Clock::now();
        double at = 0;
        for (int32_t i = 0; i < n; ++i) {
            at += gap(rng);
            const auto due = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(at));
            while (Clock::now() < due)
                std::this_thread::yield();
            try {
                tickets[i] = srv->impl->Submit(terms + offsets[i], static_cast<int32_t>(offsets[i + 1] - offsets[i]),
                                               ks[i], thresholds ? thresholds[i] : 0.0F, learned != 0);
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> g(failMutex);
                if (failure.empty())
                    failure = e.what();
            }
            submitted.store(i + 1, std::memory_order_release);
        }
        collector.join();
        if (elapsed_s)
            *elapsed_s = std::chrono::duration<double>(Clock::now() - start).count();
        if (!failure.empty())
            throw DeviceError(0, failure);
    });
}

extern "C" int bm25mi_server_counters(bm25mi_server *srv, int64_t *batches, int64_t *queries, int32_t *max_inflight,
                                      int64_t *fallbacks)
{
    return Invoke([&] {
        Require(srv && srv->impl, "Null server");
        srv->impl->Counters(batches, queries, max_inflight, fallbacks);
    });
}

// Stops the server after the submitted requests finish (uncollected results are dropped).
extern "C" int bm25mi_server_stop(bm25mi_server *srv)
{
    if (!srv)
        return BM25MI_OK;
    return Invoke([&] {
        {
            std::lock_guard<std::mutex> guard(srv->session->mutex);
            srv->impl.reset();
        }
        delete srv;
    });
}
