// The resident-index session behind the C ABI, shared by the Search entry points
// (session.hip) and the serving loop (server.hip).
//
// Mirrors bm25large::Session (FR/ascendhost/src/bm25/large_search.cpp:107-212, FR =
// feature_retrieval/src/ascendfaiss): the index is validated and uploaded once; every
// per-Search buffer is allocated up front; H2D packet copy -> kernels is launched directly or,
// with use_graph, replayed from a captured hipGraph; an empty bootstrap Search must return IDs
// 0..K-1 with score 0 (:184-187); a request is validated whole before the reusable packet is
// touched
// (FR/common/bm25/fixed_query_packet.h:26-49); results are checked the way Session::Receive
// does (:64-85); a device error marks the session failed for good (:208-210).
//
// Differences by design: impacts are precomputed at create time; the directory covers only
// long posting lists; results land in mapped pinned host memory written by the last kernel
// (no D2H copy node); Q and K may vary per call up to the configured capacity, K per query;
// and the per-Search state is replicated in `slots`, each with its own stream and graph, so
// that the serving loop can keep several batches in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace bm25mi {

#define BM25MI_HIP(expr, what)                                                                                         \
    do {                                                                                                               \
        const hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            throw DeviceError(static_cast<int64_t>(e_), std::string(what) + ": " + hipGetErrorString(e_));         \
    } while (0)

struct Query {
    const int64_t *offsets;
    const int32_t *terms;
    int32_t count;
    const int32_t *k;   // per query
    const float *thr;   // per query or nullptr
    bool learned = false; // thresholds predicted on the device by the resident model
    const uint8_t *predict = nullptr; // learned batches: per query 1 = predict, 0 = plain (nullptr: all)
};

// One in-flight Search: request packet, candidate pools, per-query bound state, per-wave
// counters, result block, stream and captured graph.
struct Slot {
    void *dPacket = nullptr, *dPool = nullptr, *dCount = nullptr, *dTau = nullptr, *dStats = nullptr;
    void *dBlockCount = nullptr;
    void *dHist = nullptr, *dItems = nullptr, *dDead = nullptr, *dCursor = nullptr, *dLive = nullptr;
    void *hPacket = nullptr, *hPacketDev = nullptr, *hResult = nullptr, *hResultDev = nullptr;
    StageRanges ranges{}; // the packet ranges the staging kernel copies for the packed batch
    BatchView btHost{};   // the packet's arrays in its mapped host memory (staging kernel)
    BatchView bt{};
    WorkView wk{};
    hipStream_t stream = nullptr;
    // Captured Searches by batch mode: [0] plain queries only; [2] threshold queries only
    // (planning, live scoring); [3] learned (captured once the session has a model).
    // Mode 1 (both kinds) is launched directly, [1] stays empty.
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // The launched batch: query count, output row offsets, plain-path postings, timing mode.
    int32_t nq = 0;
    int mode = 0; // 0 plain, 1 mixed, 2 threshold queries only, 3 learned (thresholds predicted on the
                  // device); modes 0, 2 and 3 may replay a captured graph (exec[mode])
    // A learned batch none of whose queries the device can route to the plain path: each one's
    // start bound is at least kLiveMinRatio of its score upper bound (the threshold's clamp keeps
    // it there) and its K at most liveMaxK. Its direct launch leaves out the plain kernel.
    bool allLive = false;
    std::vector<int64_t> out; // [maxQ+1]
    int64_t total = 0;
    int32_t finalGrid = 0; // final-kernel workgroups of the launched Search (its counter rows)
    // Per-query host state of Pack and the status rows of Complete, sized to the session's
    // max_queries at create time so that a Search allocates no host memory.
    std::vector<float> thrUse;
    std::vector<int64_t> perQuery;
    std::vector<uint8_t> listed;
    std::vector<int32_t> status;
    bool timed = false;
    bm25mi_stats last{};
};

class Session {
public:
    bm25mi_config cfg{};
    int64_t ndocs = 0, vocab = 0, postings = 0;
    int32_t blockDocs = 2048, granule = kGranule, blocks = 0, grid = 0, cus = 0;
    int32_t gridCapacity = 0; // the packet and counter rows hold this many scoring waves
    int32_t liveGrid = 0;     // live-kernel waves (as many as are resident at once)
    int64_t poolEntries = 0; // candidate-pool entries of each slot
    int64_t taskCost = kTaskCost;
    std::vector<int64_t> toffHost;
    std::vector<float> idfHost, tmaxHost, tmeanHost;
    std::vector<int32_t> slotHost;
    std::vector<int64_t> topOffHost;
    std::vector<float> topValHost;
    int64_t granules = 0;

    // Resident index in HBM.
    void *dPost = nullptr, *dSamples = nullptr, *dDir = nullptr, *dBmax = nullptr;
    IndexView ix{};
    size_t packetBytes = 0, resultBytes = 0;
    size_t offQoff = 0, offK = 0, offThr = 0, offTau0 = 0, offPool = 0, offOut = 0, offTerms = 0, offCost = 0;
    size_t offFeat = 0, offUpper = 0, offPlanCost = 0;
    int64_t itemsPerQuery = 0; // 64-granule items per query (planning kernel)
    int32_t liveMaxK = kLiveMaxK; // largest K a threshold query takes the live path with
    // The resident threshold model of learned Searches (bm25mi_set_threshold_model).
    void *dModel = nullptr;
    float modelYMean = 0.0F, modelYScale = 1.0F, modelSafety = 1.0F;
    size_t resStatus = 0, resScores = 0, resIds = 0;
    std::vector<Slot> slots;
    int device = 0;
    void *scratchBase = nullptr; // device address of the no-allocation scratch on this device

    bool failed = false, closed = false, serving = false;
    bm25mi_stats last{};

    Session(const int64_t *offsets, int64_t v, const int32_t *docs, const float *tf, int64_t m, const float *norm,
            int64_t n, const float *idf, const bm25mi_config &config);
    ~Session() { Release(); }
    Session(const Session &) = delete;
    Session &operator=(const Session &) = delete;

    void Release() noexcept;
    // Candidate pool entries one query may need: every block of a plain query appends at most
    // BlockKeepCap(K) docs and no doc is appended twice. A threshold query's pool also holds the
    // appends of up to kLiveReserveGranules live granules (min(K, kGranule) each); an overflow
    // reruns it on the plain path.
    int64_t PoolCapacity(int32_t k, bool threshold = false) const;
    // Score bits of the largest K-th term impact among the query's terms: every doc scores at
    // least each of its impacts, so the query's K-th score is at least this (0 if unknown).
    uint32_t StartBound(const int32_t *terms, int64_t n, int32_t k) const;
    // The wave plan: the prefix of per-query task costs (a task of query q costing taskCost +
    // ceil(postings of q / blocks), threshold queries nothing); every scoring wave cuts its own
    // contiguous range of equal estimated work out of the query-major task order (WaveRange).
    // The learned model's kFeatures inputs for one query (index statistics of its terms), and
    // the FP32 sum of its terms' largest impacts in query order (upperOut, optional).
    void Features(const int32_t *terms, int64_t n, int32_t k, float *out, float *upperOut) const;
    void PlanWaves(const int64_t *postingsPerQuery, const uint8_t *listed, int32_t nq, int64_t *costPrefix) const;
    // Validates a whole request (throws InvalidArgument, touching nothing, when it is not
    // acceptable), then writes it into the slot's packet. The batch's candidate pools must fit
    // poolBudget entries (default: the slot's pool).
    void Pack(const Query &query, Slot &slot, int64_t poolBudget = -1);
    // A learned batch's query q whose threshold the device predicts (the others take the plain path).
    bool Predicted(const Query &query, int32_t q) const;
    // Queues the slot's Search on its stream (graph replay, or direct launches with events
    // around each kernel when timed). Returns at once.
    void Launch(Slot &slot, bool timed);
    // True once the slot's launched Search has finished on the device.
    bool Ready(Slot &slot);
    // Waits for the slot's Search, checks it and copies the results (rows of query q start at
    // sum of the K of earlier queries); status[q] = 1 marks a threshold query with fewer than
    // K survivors, whose rows are left untouched.
    void Complete(Slot &slot, float *outScores, int32_t *outIds, int32_t *status);
    // Pack + Launch + Complete on slot 0; any device failure marks the session failed.
    void Search(const Query &query, float *outScores, int32_t *outIds, int32_t *status, bool timed);
    // The same plain Search with its device scratch (packet, counters, histograms, counter rows,
    // candidate pools) in the library's static scratch array instead of slot 0's buffers, and
    // no allocation of any kind; a batch whose pools do not fit is rejected.
    void SearchScratch(const Query &query, float *outScores, int32_t *outIds, int32_t *status);
    // The thresholds a learned Search of the batch would use, predicted on the device (0 = the
    // query takes the plain path).
    void PredictThresholds(const Query &query, float *out);
    // Changes the scoring launch (waves, schedule task cost) and recaptures the slots' graphs.
    void SetLaunch(int32_t waves, int64_t cost);

    // Captures the slot's graphs that do not exist yet (modes 0 and 2; 3 once a model is set).
    void Capture(Slot &slot);

private:
    void InitSlot(Slot &slot);
    void Enqueue(Slot &slot, bool timed, int mode, bool capture = false);
    void Bootstrap();
    ModelView Model() const;
    StageRanges FullRanges(bool learned) const;
    BatchView PacketView(const unsigned char *packet) const;
    int32_t FinalGrid(int32_t nq) const;
    template <class T> T *Alloc(size_t count, const char *what)
    {
        void *p = nullptr;
        BM25MI_HIP(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)), what);
        return static_cast<T *>(p);
    }
};

void LaunchInit(const WorkView &wk, int32_t maxQ, hipStream_t stream);
// Copies the ranges of a request packet from mapped host memory to its device copy.
// Learned batches (predictQueries > 0) also get their thresholds predicted, from the host packet
// view `host`, into thrOut (one wave per query).
void LaunchStage(const uint4 *src, uint4 *dst, const StageRanges &ranges, const BatchView &host, float *thrOut,
                 const ModelView &mv, int32_t predictQueries, hipStream_t stream);
// start/stop (optional): events stamped at the kernel's own start and end.
// addRows: add the counters to the rows of the live kernel that ran before it in the same Search.
// keepCap: the session's keep_cap (BlockKeepCap). nqKnown: the batch's query count, or -1 for
// the kernel to read it from the packet (a captured graph).
bool LaunchScore(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t grid, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr,
                 int32_t addRows = 0, int32_t keepCap = 0, int32_t nqKnown = -1);
// Live granules of threshold queries, after LaunchPrune with pruneWaves item waves; counter rows
// from rowBase on.
bool LaunchLive(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                int32_t grid, int32_t rowBase, int32_t pruneWaves, hipStream_t stream);
int LiveBlocksPerCu(int blockDocs);
int PruneBlocksPerCu(int blockDocs);
// Planning of threshold queries: `waves` item waves (a multiple of 256 up to kMaxPruneWaves) and
// the learned batches' header wave; nqKnown = the batch's query count, or -1 to read it from the
// packet (captured graphs).
bool LaunchPrune(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t waves, int32_t nqKnown, hipStream_t stream);
// Item waves of the planning kernel for a batch of `items` items (64 granules of one query each).
int32_t PruneWaves(int64_t items);
int ScoreBlocksPerCu(int blockDocs);
// One workgroup per query slot (workgroups beyond the batch's queries only sum counter rows).
void LaunchFinal(const BatchView &bt, const WorkView &wk, int32_t workgroups, unsigned long long *hostStats,
                 int32_t scoreGrid, int32_t blocks, int32_t blockDocs, int32_t keepCap, hipStream_t stream);

} // namespace bm25mi

struct bm25mi_session {
    std::unique_ptr<bm25mi::Session> impl;
    bm25mi::ErrorInfo error;
    std::mutex mutex;
};
