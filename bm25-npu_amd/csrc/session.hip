// Host runtime of libbm25mi: the resident-index session behind the C ABI (declared in
// session.h, which describes how it mirrors bm25large::Session of the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "layout.h"
#include "session.h"

namespace bm25mi {

static size_t Align(size_t n, size_t a) { return (n + a - 1) / a * a; }

// One-shot failures injected by bm25mi_inject_failure, per stage: the next call that reaches
// the stage throws a DeviceError with the injected code and message (the reference injects the
// same failures through mock ACL functions: FR/ut/cases/bm25/sdk_facade_tests.cpp:53-62).
namespace {
struct Injected {
    bool pending = false;
    int64_t code = 0;
    std::string message;
};
std::mutex injectMutex;
constexpr int kStages = 4;
Injected injected[kStages]; // create, search, close, result
int StageIndex(const char *stage)
{
    static const char *names[kStages] = {"create", "search", "close", "result"};
    for (int i = 0; i < kStages; ++i)
        if (std::strcmp(stage, names[i]) == 0)
            return i;
    return -1;
}
// Takes a pending injection of the stage, if any.
bool TakeInjected(const char *stage)
{
    std::lock_guard<std::mutex> g(injectMutex);
    Injected &f = injected[StageIndex(stage)];
    const bool pending = f.pending;
    f.pending = false;
    return pending;
}
} // namespace

// Device scratch of bm25mi_search_batch_noalloc: the request packet, per-query counters, score
// histograms and bounds, the per-wave counter rows and the candidate pools of one Search, at
// most kScratchBytes. Each GPU holds its own copy from the library's load until the process ends;
// its device address is taken once per device (ScratchAddress) and only that address is used.
constexpr size_t kScratchBytes = BM25MI_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratchMutex;
void *scratchAddress[kMaxDevices] = {}; // per device, from hipGetSymbolAddress
} // namespace

// The device address of `scratch` on `device`, which must be the current device. Called at
// session creation.
static void *ScratchAddress(int device)
{
    Require(device >= 0 && device < kMaxDevices, "device_id beyond the library's device table");
    std::lock_guard<std::mutex> g(scratchMutex);
    if (!scratchAddress[device]) {
        void *p = nullptr;
        BM25MI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "no-allocation scratch address");
        Require(p != nullptr, "no-allocation scratch has no device address");
        scratchAddress[device] = p;
    }
    return scratchAddress[device];
}

static void FailIfInjected(const char *stage)
{
    std::lock_guard<std::mutex> g(injectMutex);
    Injected &f = injected[StageIndex(stage)];
    if (!f.pending)
        return;
    f.pending = false;
    throw DeviceError(f.code, f.message);
}

int64_t Session::PoolCapacity(int32_t k, bool threshold) const
{
    const int64_t plain =
        std::min<int64_t>(static_cast<int64_t>(blocks) * BlockKeepCap(k, blockDocs, cfg.keep_cap), ndocs);
    // A threshold query may be scored on the plain path (Pack, or the staging kernel's routing),
    // so its pool covers the plain path; on the live path each live granule appends at most
    // min(K, kGranule) docs, reserved for up to kLiveReserveGranules granules (a live pool that
    // still overflows reruns the query on the plain path: the final kernel flags it).
    if (!threshold)
        return plain;
    const int64_t live = std::min<int64_t>(granules, kLiveReserveGranules) * std::min<int64_t>(k, kGranule);
    return std::max(plain, std::min<int64_t>(live, ndocs));
}

uint32_t Session::StartBound(const int32_t *terms, int64_t n, int32_t k) const
{
    float best = 0.0F;
    for (int64_t w = 0; w < n; ++w) {
        const int32_t t = terms[w];
        if (t >= vocab || topOffHost[t + 1] - topOffHost[t] < k)
            continue;
        best = std::max(best, topValHost[topOffHost[t] + k - 1]);
    }
    uint32_t bits;
    std::memcpy(&bits, &best, 4);
    return bits;
}

void Session::Features(const int32_t *terms, int64_t n, int32_t k, float *out, float *upperOut) const
{
    // Per-term statistics of the query's known terms (IDs >= vocab skipped, repeats counted),
    // pooled per query. kth = the term's K-th largest impact when the session keeps it (0
    // otherwise); the largest kth is the provable lower bound of the K-th score.
    constexpr int kTop = 3;
    float m[kTop] = {0.0F, 0.0F, 0.0F}; // largest term maxima, descending
    float upper = 0.0F, lower = 0.0F, second = 0.0F;
    double sumKth = 0, sumMean = 0, sumIdf = 0, sumDf = 0, maxDf = 0, minDf = -1, density = 0, logAll = 0;
    int known = 0;
    const double nd = static_cast<double>(ndocs);
    for (int64_t w = 0; w < n; ++w) {
        const int32_t t = terms[w];
        if (t >= vocab)
            continue;
        ++known;
        const float mx = tmaxHost[t];
        upper += mx; // sequential FP32 sum in query order: no doc scores above it
        for (int i = 0; i < kTop; ++i)
            if (mx > m[i]) {
                for (int j = kTop - 1; j > i; --j)
                    m[j] = m[j - 1];
                m[i] = mx;
                break;
            }
        const float kth = topOffHost[t + 1] - topOffHost[t] >= k ? topValHost[topOffHost[t] + k - 1] : 0.0F;
        if (kth > lower) {
            second = lower;
            lower = kth;
        } else if (kth > second) {
            second = kth;
        }
        const double df = static_cast<double>(toffHost[t + 1] - toffHost[t]);
        sumKth += kth;
        sumMean += tmeanHost[t];
        sumIdf += idfHost[t];
        sumDf += df;
        maxDf = std::max(maxDf, df);
        minDf = minDf < 0 ? df : std::min(minDf, df);
        density += mx * df / nd;
        logAll += std::log(std::max(df, 0.5) / nd);
    }
    const auto lg = [](double v) { return std::log(std::max(v, 1e-30)); };
    const double vals[kFeatures] = {std::log(static_cast<double>(k)),
                                    static_cast<double>(known),
                                    lg(upper),
                                    lg(lower),
                                    lg(second),
                                    m[0],
                                    m[1],
                                    m[2],
                                    sumKth,
                                    sumMean,
                                    sumIdf,
                                    std::log1p(sumDf),
                                    std::log1p(maxDf),
                                    std::log1p(std::max(minDf, 0.0)),
                                    density,
                                    std::max(-40.0, logAll)};
    for (int i = 0; i < kFeatures; ++i)
        out[i] = static_cast<float>(vals[i]);
    if (upperOut)
        *upperOut = upper;
}

void Session::PlanWaves(const int64_t *postingsPerQuery, const uint8_t *listed, int32_t nq, int64_t *costPrefix) const
{
    costPrefix[0] = 0;
    for (int32_t q = 0; q < nq; ++q) {
        const int64_t cost = listed[q] ? 0 : taskCost + (postingsPerQuery[q] + blocks - 1) / blocks;
        costPrefix[q + 1] = costPrefix[q] + cost * blocks;
    }
    // The kernels multiply the total by wave indices below gridCapacity in 64 bits.
    Require(costPrefix[nq] < (int64_t(1) << 62) / gridCapacity, "Batch too large for the wave plan");
}

Session::Session(const int64_t *offsets, int64_t v, const int32_t *docs, const float *tf, int64_t m,
                 const float *norm, int64_t n, const float *idf, const bm25mi_config &config)
    : cfg(config), ndocs(n), vocab(v), postings(m)
{
    Require(cfg.max_queries >= 1 && cfg.max_queries <= BM25MI_MAX_QUERIES, "max_queries must be in [1, 10000]");
    Require(cfg.max_words >= 0 && cfg.max_words <= 16777216, "max_words must be in [0, 2^24]");
    Require(cfg.max_k >= 1 && cfg.max_k <= BM25MI_MAX_K, "max_k must be in [1, 1024]");
    Require(std::isfinite(cfg.k1) && cfg.k1 > 0, "k1 must be finite and > 0");
    Require(cfg.task_cost >= 0, "task_cost must be >= 0");
    Require(cfg.slots >= 0 && cfg.slots <= 16, "slots must be in [0, 16]");
    Require(cfg.pool_entries >= 0, "pool_entries must be >= 0");
    Require(cfg.live_max_k >= 0 && cfg.live_max_k <= BM25MI_MAX_K, "live_max_k must be in [0, 1024]");
    Require(cfg.keep_cap >= 0 && cfg.keep_cap <= 8192, "keep_cap must be in [0, 8192]");
    if (cfg.task_cost > 0)
        taskCost = cfg.task_cost;
    blockDocs = cfg.block_docs ? cfg.block_docs : 2048;
    Require(blockDocs == 2048 || blockDocs == 4096 || blockDocs == 8192, "block_docs must be 2048, 4096 or 8192");
    ValidateIndex(offsets, v, docs, tf, m, norm, n, idf, 0);
    Require(m >= 1, "Index needs at least one posting");
    // Posting indices are 32-bit on the device.
    Require(m < 0xFFFFFFFFLL, "Index has too many postings (at most 2^32 - 2)");
    Require(cfg.max_k <= n, "max_k exceeds the document count");
    device = cfg.device_id;

    // Host-side tables: impacts, directory, block maxima, term top impacts.
    blocks = static_cast<int32_t>((n + blockDocs - 1) / blockDocs);
    Layout layout;
    std::vector<float> impact(static_cast<size_t>(m));
    ComputeImpacts(offsets, v, docs, tf, norm, idf, cfg.k1, impact.data(), 0);
    BuildLayout(offsets, v, docs, impact.data(), n, granule, kLongMin, 0, layout);
    granules = layout.granules;
    toffHost.assign(offsets, offsets + v + 1);
    idfHost.assign(idf, idf + v);
    tmaxHost = layout.termMax;
    tmeanHost = layout.termMean;
    slotHost = layout.termSlot;
    topOffHost.swap(layout.topOff);
    topValHost.swap(layout.topVal);

    int count = 0;
    BM25MI_HIP(hipGetDeviceCount(&count), "count devices");
    Require(device >= 0 && device < count, "device_id out of range");
    BM25MI_HIP(hipSetDevice(device), "set device");
    // Spin while waiting for a Search (lowest wake-up latency). The runtime refuses the flag once
    // the device is already active in this process; the Search is then only waited on differently.
    (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
    (void)hipGetLastError();
    hipDeviceProp_t prop;
    BM25MI_HIP(hipGetDeviceProperties(&prop, device), "device properties");
    cus = prop.multiProcessorCount;
    scratchBase = ScratchAddress(device);
    const int perCu = ScoreBlocksPerCu(blockDocs);
    Require(perCu >= 1, "The scoring kernel does not fit this device");
    grid = cfg.grid_blocks > 0 ? cfg.grid_blocks : cus * perCu;
    Require(grid >= 1 && grid <= (1 << 20), "grid_blocks out of range");
    Require(PruneBlocksPerCu(blockDocs) >= 1, "The planning kernel does not fit this device");
    const int livePerCu = LiveBlocksPerCu(blockDocs);
    Require(livePerCu >= 1, "The live kernel does not fit this device");
#ifdef BM25MI_LIVE_PER_CU
    liveGrid = cus * std::min(livePerCu, BM25MI_LIVE_PER_CU);
#else
    liveGrid = cus * livePerCu;
#endif
    // Counter rows: 2 x the largest grid.
    gridCapacity = std::max(std::max(grid, liveGrid), 2 * cus * perCu);
    Require(static_cast<int64_t>(cfg.max_queries) * blocks < (int64_t(1) << 32), "Too many scoring tasks per Search");
    itemsPerQuery = (granules + kWave - 1) / kWave;
    // Live granules are ranked in 32 bits (planning and live kernels).
    Require(static_cast<int64_t>(cfg.max_queries) * itemsPerQuery * kWave < (int64_t(1) << 32),
            "Too many granules per Search");
    liveMaxK = cfg.live_max_k > 0 ? cfg.live_max_k : kLiveMaxK;

    // Candidate pools: by default room for max_queries queries of K = max_k in every slot;
    // a smaller pool_entries budget makes callers admit fewer large-K queries per batch.
    const int64_t fullPool =
        static_cast<int64_t>(cfg.max_queries) * std::max(PoolCapacity(cfg.max_k), PoolCapacity(cfg.max_k, true));
    poolEntries = cfg.pool_entries > 0 ? std::min(cfg.pool_entries, fullPool) : fullPool;
    Require(poolEntries >= PoolCapacity(cfg.max_k), "pool_entries cannot hold one query of K = max_k");
    const int32_t nslots = cfg.slots > 0 ? cfg.slots : 1;
    size_t freeBytes = 0, totalBytes = 0;
    BM25MI_HIP(hipMemGetInfo(&freeBytes, &totalBytes), "memory info");
    const uint64_t known = static_cast<uint64_t>(m) * 8 + static_cast<uint64_t>(m) / 4 + 4 +
                           layout.dir.size() * 4 + layout.blockMax.size() * 4 +
                           static_cast<uint64_t>(nslots) * static_cast<uint64_t>(poolEntries) * 8 +
                           static_cast<uint64_t>(nslots) * static_cast<uint64_t>(cfg.max_queries) * blocks * 5 +
                           static_cast<uint64_t>(nslots) * static_cast<uint64_t>(std::max<int64_t>(1, cfg.max_words)) * blocks * 4;
    // MemoryGate of large_search.cpp:87-93: known bytes must fit 80% of free HBM.
    Require(known <= static_cast<uint64_t>(freeBytes) / 5 * 4, "Index, pools and arenas exceed 80% of free HBM");

    // Postings as {doc, impact bits}: one 8-byte load per posting.
    {
        std::vector<uint2> packed(static_cast<size_t>(m));
        ParallelChunks(m, Threads(0), [&](int, int64_t b, int64_t e) {
            for (int64_t p = b; p < e; ++p) {
                uint32_t bitsv;
                std::memcpy(&bitsv, &impact[p], 4);
                packed[p] = make_uint2(static_cast<uint32_t>(docs[p]), bitsv);
            }
        });
        std::vector<float>().swap(impact);
        dPost = Alloc<uint2>(m, "alloc postings");
        BM25MI_HIP(hipMemcpy(dPost, packed.data(), m * sizeof(uint2), hipMemcpyHostToDevice), "upload postings");
    }
    {
        // Doc ID of every 16th posting: short lists are located through these samples.
        const int64_t ns = (m + (1 << kSampleShift) - 1) >> kSampleShift;
        std::vector<int32_t> samples(static_cast<size_t>(ns));
        for (int64_t i = 0; i < ns; ++i)
            samples[i] = docs[i << kSampleShift];
        dSamples = Alloc<int32_t>(ns, "alloc posting samples");
        BM25MI_HIP(hipMemcpy(dSamples, samples.data(), ns * 4, hipMemcpyHostToDevice), "upload posting samples");
    }
    dDir = Alloc<uint32_t>(layout.dir.size(), "alloc directory");
    if (!layout.dir.empty())
        BM25MI_HIP(hipMemcpy(dDir, layout.dir.data(), layout.dir.size() * 4, hipMemcpyHostToDevice), "upload dir");
    dBmax = Alloc<float>(layout.blockMax.size(), "alloc block max");
    if (!layout.blockMax.empty())
        BM25MI_HIP(hipMemcpy(dBmax, layout.blockMax.data(), layout.blockMax.size() * 4, hipMemcpyHostToDevice),
                   "upload block max");
    ix = IndexView{static_cast<const uint2 *>(dPost), static_cast<const int32_t *>(dSamples),
                   static_cast<const uint32_t *>(dDir), static_cast<const float *>(dBmax), granules,
                   static_cast<int32_t>(n)};

    // Request packet layout (the same in every slot); every array starts on 16 bytes (staging).
    const size_t q = static_cast<size_t>(cfg.max_queries);
    size_t at = 16;
    offQoff = at;
    at = Align(at + (q + 1) * 8, 16);
    offK = at;
    at = Align(at + q * 4, 16);
    offThr = at;
    at = Align(at + q * 4, 16);
    offTau0 = at;
    at = Align(at + q * 4, 16);
    offPool = at;
    at = Align(at + (q + 1) * 8, 16);
    offOut = at;
    at = Align(at + (q + 1) * 8, 16);
    offTerms = at;
    at = Align(at + std::max<size_t>(1, static_cast<size_t>(cfg.max_words)) * 16, 16);
    offCost = at;
    at = Align(at + (q + 1) * 8, 16);
    // Learned batches only: per-query model features, score upper bounds and plain-path costs.
    offFeat = at;
    at = Align(at + q * kFeatures * 4, 16);
    offUpper = at;
    at = Align(at + q * 4, 16);
    offPlanCost = at;
    at = Align(at + q * 8, 256);
    packetBytes = at;
    // Result block layout: per final-workgroup counter rows, status, scores, IDs.
    resStatus = q * kStats * 8;
    resScores = Align(resStatus + q * 4, 16);
    resIds = Align(resScores + q * cfg.max_k * 4, 16);
    resultBytes = Align(resIds + q * cfg.max_k * 4, 256);

    slots.resize(static_cast<size_t>(nslots));
    for (auto &slot : slots)
        InitSlot(slot);
    Bootstrap();
}

void Session::InitSlot(Slot &slot)
{
    const size_t q = static_cast<size_t>(cfg.max_queries);
    slot.out.assign(q + 1, 0);
    slot.thrUse.assign(q, 0.0F);
    slot.perQuery.assign(q, 0);
    slot.listed.assign(q, 0);
    slot.status.assign(q, 0);
    BM25MI_HIP(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking), "create stream");
    for (auto &e : slot.ev)
        BM25MI_HIP(hipEventCreate(&e), "create event");
    // Coherent (fine-grained) and mapped: the staging kernel reads the packet from it directly.
    BM25MI_HIP(hipHostMalloc(&slot.hPacket, packetBytes, hipHostMallocMapped | hipHostMallocCoherent),
               "alloc pinned packet");
    std::memset(slot.hPacket, 0, packetBytes);
    BM25MI_HIP(hipHostGetDevicePointer(&slot.hPacketDev, slot.hPacket, 0), "map pinned packet");
    slot.ranges = FullRanges(false);
    slot.dPacket = Alloc<unsigned char>(packetBytes, "alloc device packet");
    BM25MI_HIP(hipHostMalloc(&slot.hResult, resultBytes, hipHostMallocMapped), "alloc pinned result");
    std::memset(slot.hResult, 0, resultBytes);
    BM25MI_HIP(hipHostGetDevicePointer(&slot.hResultDev, slot.hResult, 0), "map pinned result");
    slot.dPool = Alloc<unsigned long long>(static_cast<size_t>(poolEntries), "alloc candidate pools");
    slot.dCount = Alloc<uint32_t>(q, "alloc pool counts");
    slot.dBlockCount = Alloc<uint32_t>(q * static_cast<size_t>(BlockCountStride(blocks)), "alloc task candidate counts");
    slot.dHist = Alloc<uint32_t>(q * kHistBins, "alloc score histograms");
    slot.dTau = Alloc<uint32_t>(q, "alloc thresholds");
    slot.dStats = Alloc<unsigned long long>(2 * static_cast<size_t>(gridCapacity) * kStats, "alloc stats");
    slot.dItems = Alloc<unsigned long long>(q * static_cast<size_t>(itemsPerQuery), "alloc live item masks");
    slot.dDead = Alloc<uint32_t>(kMaxPruneWaves, "alloc dead counts");
    slot.dLive = Alloc<uint32_t>(kMaxPruneWaves, "alloc live counts");
    slot.dCursor = Alloc<uint32_t>(std::max<size_t>(1, static_cast<size_t>(cfg.max_words)) * static_cast<size_t>(blocks),
                                   "alloc short-list cursors");
    slot.bt = PacketView(static_cast<unsigned char *>(slot.dPacket));
    slot.btHost = PacketView(static_cast<unsigned char *>(slot.hPacketDev));
    auto *rs = static_cast<unsigned char *>(slot.hResultDev);
    slot.wk = WorkView{static_cast<unsigned long long *>(slot.dPool),
                       static_cast<uint32_t *>(slot.dCount),
                       static_cast<uint32_t *>(slot.dBlockCount),
                       static_cast<uint32_t *>(slot.dHist),
                       static_cast<uint32_t *>(slot.dTau),
                       static_cast<unsigned long long *>(slot.dStats),
                       static_cast<unsigned long long *>(slot.dItems),
                       static_cast<uint32_t *>(slot.dDead),
                       static_cast<uint32_t *>(slot.dLive),
                       static_cast<uint32_t *>(slot.dCursor),
                       reinterpret_cast<int32_t *>(rs + resStatus),
                       reinterpret_cast<float *>(rs + resScores),
                       reinterpret_cast<int32_t *>(rs + resIds)};
    LaunchInit(slot.wk, cfg.max_queries, slot.stream);
    BM25MI_HIP(hipStreamSynchronize(slot.stream), "initialise counters");
    if (cfg.use_graph)
        Capture(slot);
}

StageRanges Session::FullRanges(bool learned) const
{
    // The whole packet; for learned batches without the thresholds, which the staging kernel
    // predicts on the device.
    StageRanges r{};
    if (!learned) {
        r.off[0] = 0;
        r.len[0] = static_cast<uint32_t>(packetBytes / 16);
        r.n = 1;
    } else {
        r.off[0] = 0;
        r.len[0] = static_cast<uint32_t>(offThr / 16);
        r.off[1] = static_cast<uint32_t>(offTau0 / 16);
        r.len[1] = static_cast<uint32_t>((packetBytes - offTau0) / 16);
        r.n = 2;
    }
    return r;
}

BatchView Session::PacketView(const unsigned char *pk) const
{
    return BatchView{reinterpret_cast<const int32_t *>(pk),
                     reinterpret_cast<const int64_t *>(pk + offQoff),
                     reinterpret_cast<const int32_t *>(pk + offK),
                     reinterpret_cast<const float *>(pk + offThr),
                     reinterpret_cast<const uint32_t *>(pk + offTau0),
                     reinterpret_cast<const int64_t *>(pk + offPool),
                     reinterpret_cast<const int64_t *>(pk + offOut),
                     reinterpret_cast<const uint4 *>(pk + offTerms),
                     reinterpret_cast<const int64_t *>(pk + offCost),
                     reinterpret_cast<const float *>(pk + offFeat),
                     reinterpret_cast<const float *>(pk + offUpper),
                     reinterpret_cast<const int64_t *>(pk + offPlanCost)};
}

int32_t Session::FinalGrid(int32_t nq) const
{
    return std::min(cfg.max_queries, std::max(nq, 8));
}

ModelView Session::Model() const
{
    return ModelView{static_cast<const float *>(dModel), modelYMean, modelYScale, modelSafety, kLiveMinRatio,
                     liveMaxK};
}

void Session::Enqueue(Slot &slot, bool timed, int mode, bool capture)
{
    // Learned batches: the staging kernel also predicts the thresholds (one wave per query).
    const int32_t predict = mode == 3 ? (capture ? cfg.max_queries : slot.nq) : 0;
    // Mixed and learned batches also run the plain kernel, unless no query of a learned batch can
    // be routed to the plain path (a captured graph keeps it).
    const bool plainToo = mode == 1 || (mode == 3 && (capture || !slot.allLive));
    LaunchStage(static_cast<const uint4 *>(slot.hPacketDev), static_cast<uint4 *>(slot.dPacket),
                capture ? FullRanges(mode == 3) : slot.ranges, slot.btHost,
                reinterpret_cast<float *>(static_cast<unsigned char *>(slot.dPacket) + offThr), Model(), predict,
                slot.stream);
    // Timed plain batches: events stamped by the scoring kernel itself; otherwise recorded
    // around the scoring kernels (planning, plain and live scoring).
    if (timed && mode != 0)
        BM25MI_HIP(hipEventRecord(slot.ev[1], slot.stream), "record score start");
    if (mode == 0) {
        Require(LaunchScore(blockDocs, ix, slot.bt, slot.wk, blocks, grid, slot.stream, timed ? slot.ev[1] : nullptr,
                            timed ? slot.ev[2] : nullptr, 0, cfg.keep_cap, capture ? -1 : slot.nq),
                "Unsupported block size");
    } else {
        // One stream: planning (with the plain queries' wave plan of learned batches), live
        // scoring, then plain scoring of the queries routed to the plain path (mixed and learned
        // batches), which adds its counters to the live kernel's rows.
        const int32_t pruneWaves =
            PruneWaves(static_cast<int64_t>(capture ? cfg.max_queries : slot.nq) * itemsPerQuery);
        Require(LaunchPrune(blockDocs, ix, slot.bt, slot.wk, blocks, pruneWaves, capture ? -1 : slot.nq, slot.stream),
                "Unsupported block size");
        Require(LaunchLive(blockDocs, ix, slot.bt, slot.wk, blocks, liveGrid, 0, pruneWaves, slot.stream),
                "Unsupported block size");
        if (plainToo)
            Require(LaunchScore(blockDocs, ix, slot.bt, slot.wk, blocks, grid, slot.stream, nullptr, nullptr, liveGrid,
                                cfg.keep_cap, capture ? -1 : slot.nq),
                    "Unsupported block size");
    }
    if (timed && mode != 0)
        BM25MI_HIP(hipEventRecord(slot.ev[2], slot.stream), "record score end");
    // Counter rows: [0, grid) of the plain kernel, [0, liveGrid) of the live kernel (the plain
    // kernel adds into the live kernel's rows where both wrote).
    const int32_t rows = mode == 0 ? grid : plainToo ? std::max(grid, liveGrid) : liveGrid;
    // One final workgroup per query of the batch, and at least 8 (up to the session's capacity)
    // to share the counter rows; a captured graph has one per query of the session's capacity.
    // Workgroups beyond the batch's queries only sum counter rows.
    slot.finalGrid = capture ? cfg.max_queries : FinalGrid(slot.nq);
    LaunchFinal(slot.bt, slot.wk, slot.finalGrid, static_cast<unsigned long long *>(slot.hResultDev), rows, blocks,
                blockDocs, cfg.keep_cap, slot.stream);
}

void Session::Capture(Slot &slot)
{
    for (int g = 0; g < 4; ++g) {
        // Mixed batches launch directly (Launch); learned ones are captured once the session has
        // a model (its address and output transform are part of the graph).
        if (g == 1 || (g == 3 && !dModel) || slot.exec[g])
            continue;
        BM25MI_HIP(hipStreamBeginCapture(slot.stream, hipStreamCaptureModeThreadLocal), "begin graph capture");
        try {
            Enqueue(slot, false, g, true);
        } catch (...) {
            hipGraph_t discarded = nullptr;
            (void)hipStreamEndCapture(slot.stream, &discarded);
            if (discarded)
                (void)hipGraphDestroy(discarded);
            throw;
        }
        BM25MI_HIP(hipStreamEndCapture(slot.stream, &slot.graph[g]), "end graph capture");
        BM25MI_HIP(hipGraphInstantiate(&slot.exec[g], slot.graph[g], nullptr, nullptr, 0), "instantiate graph");
    }
}

bool Session::Predicted(const Query &query, int32_t q) const
{
    return query.learned && (!query.predict || query.predict[q]) && query.k[q] <= liveMaxK;
}

void Session::Pack(const Query &query, Slot &slot, int64_t poolBudget)
{
    if (poolBudget < 0)
        poolBudget = poolEntries;
    // Validate the complete request before modifying the reusable packet.
    const int32_t nq = query.count;
    Require(nq >= 1 && nq <= cfg.max_queries, "Query batch exceeds declared capacity or is empty");
    Require(query.offsets && query.k, "Null query or K array");
    Require(query.offsets[0] == 0, "query_offsets[0] must be 0");
    int64_t words = 0;
    for (int32_t q = 0; q < nq; ++q) {
        const int64_t len = query.offsets[q + 1] - query.offsets[q];
        Require(len >= 0 && len <= kMaxTerms && len <= cfg.max_words - words,
                "Query exceeds word capacity or the 64-term limit");
        words += len;
        Require(query.k[q] >= 1 && query.k[q] <= cfg.max_k && query.k[q] <= ndocs, "K out of range");
        if (query.thr)
            Require(std::isfinite(query.thr[q]) && query.thr[q] >= 0, "Threshold must be finite and >= 0");
    }
    Require(words == 0 || query.terms, "Null query terms");
    for (int64_t w = 0; w < words; ++w)
        Require(query.terms[w] >= 0 && query.terms[w] < INT32_MAX, "Invalid public word ID");
    // A threshold below kLiveMinRatio of the query's score upper bound (the FP32 sum of its
    // terms' largest impacts) leaves most granules live, and above K = liveMaxK the live path's
    // per-granule cut costs more than it saves: such a query is scored on the plain path, with
    // the same results. Learned batches are routed on the device (staging kernel) by the same rules;
    // their queries with K <= liveMaxK get pools that fit either path.
    // Per-query host state lives in the slot's buffers, sized at create time: packing allocates
    // nothing.
    float *thrUse = slot.thrUse.data();
    std::fill_n(thrUse, nq, 0.0F);
    int64_t poolNeed = 0;
    for (int32_t q = 0; q < nq; ++q) {
        if (query.thr && query.thr[q] > 0.0F && query.k[q] <= liveMaxK) {
            float ub = 0.0F;
            for (int64_t w = query.offsets[q]; w < query.offsets[q + 1]; ++w)
                if (query.terms[w] < vocab)
                    ub += tmaxHost[query.terms[w]];
            if (query.thr[q] >= kLiveMinRatio * ub)
                thrUse[q] = query.thr[q];
        }
        poolNeed += PoolCapacity(query.k[q], thrUse[q] > 0.0F || Predicted(query, q));
    }
    Require(poolNeed <= poolBudget, "Batch needs more candidate-pool memory than a slot (or the no-allocation scratch) holds");

    auto *pk = static_cast<unsigned char *>(slot.hPacket);
    reinterpret_cast<int32_t *>(pk)[0] = nq;
    auto *qoff = reinterpret_cast<int64_t *>(pk + offQoff);
    auto *kk = reinterpret_cast<int32_t *>(pk + offK);
    auto *thr = reinterpret_cast<float *>(pk + offThr);
    auto *tau0 = reinterpret_cast<uint32_t *>(pk + offTau0);
    auto *pool = reinterpret_cast<int64_t *>(pk + offPool);
    auto *out = reinterpret_cast<int64_t *>(pk + offOut);
    std::fill_n(slot.out.begin(), nq + 1, int64_t(0));
    qoff[0] = pool[0] = out[0] = 0;
    for (int32_t q = 0; q < nq; ++q) {
        qoff[q + 1] = query.offsets[q + 1];
        kk[q] = query.k[q];
        thr[q] = thrUse[q];
        pool[q + 1] = pool[q] + PoolCapacity(query.k[q], thr[q] > 0.0F || Predicted(query, q));
        out[q + 1] = out[q] + query.k[q];
        slot.out[q + 1] = out[q + 1];
        tau0[q] = StartBound(query.terms + query.offsets[q], query.offsets[q + 1] - query.offsets[q], query.k[q]);
    }
    auto *info = reinterpret_cast<uint4 *>(pk + offTerms);
    int64_t *perQuery = slot.perQuery.data();
    std::fill_n(perQuery, nq, int64_t(0));
    int64_t total = 0;
    for (int32_t q = 0; q < nq; ++q)
        for (int64_t w = query.offsets[q]; w < query.offsets[q + 1]; ++w) {
            const int32_t t = query.terms[w];
            if (t < vocab) {
                const int64_t len = toffHost[t + 1] - toffHost[t];
                uint32_t ub;
                std::memcpy(&ub, &tmaxHost[t], 4);
                info[w] = make_uint4(static_cast<uint32_t>(toffHost[t]), static_cast<uint32_t>(len),
                                     static_cast<uint32_t>(slotHost[t]), ub);
                total += len;
                perQuery[q] += len;
            } else {
                info[w] = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
            }
        }
    // Threshold queries are planned by the planning kernel and scored from its live masks.
    uint8_t *listed = slot.listed.data();
    int32_t withThr = 0;
    for (int32_t q = 0; q < nq; ++q) {
        listed[q] = thr[q] > 0.0F ? 1 : 0;
        withThr += listed[q] ? 1 : 0;
    }
    // A learned batch none of whose queries may take the live path (every K above liveMaxK, or
    // all excluded by the caller) is launched as a plain batch: no prediction or planning.
    bool learnedLaunch = false;
    for (int32_t q = 0; q < nq && query.learned && !learnedLaunch; ++q)
        learnedLaunch = Predicted(query, q);
    reinterpret_cast<int32_t *>(pk)[1] = withThr;
    reinterpret_cast<int32_t *>(pk)[2] = nq - withThr;
    reinterpret_cast<int32_t *>(pk)[3] = learnedLaunch ? 1 : 0;
    slot.mode = learnedLaunch ? 3 : withThr == 0 ? 0 : withThr < nq ? 1 : 2;
    // Learned batches: the plan with every query plain, the most the device-side routing can
    // keep (checked here), and what the threshold kernel needs to route and re-plan.
    PlanWaves(perQuery, listed, nq, reinterpret_cast<int64_t *>(pk + offCost));
    // The packet bytes the kernels read: each array's first nq (or words) entries; a learned
    // batch's thresholds are predicted on the device instead. A captured graph stages the whole
    // packet (FullRanges: its ranges are fixed at capture).
    {
        StageRanges &r = slot.ranges;
        r.n = 0;
        const auto add = [&](size_t off, size_t bytes) {
            if (bytes)
                r.off[r.n] = static_cast<uint32_t>(off / 16), r.len[r.n++] = static_cast<uint32_t>((bytes + 15) / 16);
        };
        const size_t n = static_cast<size_t>(nq);
        add(0, 16);
        add(offQoff, (n + 1) * 8);
        add(offK, n * 4);
        if (!learnedLaunch)
            add(offThr, n * 4);
        add(offTau0, n * 4);
        add(offPool, (n + 1) * 8);
        add(offOut, (n + 1) * 8);
        add(offTerms, static_cast<size_t>(words) * 16);
        add(offCost, (n + 1) * 8);
        if (learnedLaunch) {
            add(offFeat, n * kFeatures * 4);
            add(offUpper, n * 4);
            add(offPlanCost, n * 8);
        }
    }
    slot.allLive = false;
    if (learnedLaunch) {
        auto *feat = reinterpret_cast<float *>(pk + offFeat);
        auto *upper = reinterpret_cast<float *>(pk + offUpper);
        auto *planCost = reinterpret_cast<int64_t *>(pk + offPlanCost);
        bool allLive = true;
        for (int32_t q = 0; q < nq; ++q) {
            Features(query.terms + query.offsets[q], query.offsets[q + 1] - query.offsets[q], query.k[q],
                     feat + static_cast<int64_t>(q) * kFeatures, upper + q);
            if (!Predicted(query, q))
                upper[q] = 0.0F; // routed to the plain path by the staging kernel
            // The predicted threshold is clamped to at least the start bound, so a start bound
            // of at least kLiveMinRatio of the upper bound (the same FP32 product as on the
            // device) keeps the query on the live path whatever the model predicts.
            float lower;
            std::memcpy(&lower, &tau0[q], 4);
            allLive = allLive && upper[q] > 0.0F && query.k[q] <= liveMaxK && lower >= kLiveMinRatio * upper[q];
            planCost[q] = (taskCost + (perQuery[q] + blocks - 1) / blocks) * blocks;
        }
        slot.allLive = allLive;
    }
    slot.nq = nq;
    slot.total = total;
}

void Session::Launch(Slot &slot, bool timed)
{
    slot.timed = timed;
    // Device-time events only for timed Searches: the two records cost about 2 us per call.
    if (timed)
        BM25MI_HIP(hipEventRecord(slot.ev[0], slot.stream), "record start");
    // Mixed batches (mode 1) launch directly; the others replay their graph when the session
    // captured one (use_graph) and the Search is not timed.
    if (slot.mode != 1 && slot.exec[slot.mode] && !timed) {
        BM25MI_HIP(hipGraphLaunch(slot.exec[slot.mode], slot.stream), "replay Search graph");
        slot.finalGrid = cfg.max_queries;
    } else {
        Enqueue(slot, timed, slot.mode);
    }
    if (timed)
        BM25MI_HIP(hipEventRecord(slot.ev[3], slot.stream), "record end");
}

bool Session::Ready(Slot &slot)
{
    const hipError_t e = hipStreamQuery(slot.stream);
    if (e == hipErrorNotReady)
        return false;
    BM25MI_HIP(e, "query Search stream");
    return true;
}

void Session::Complete(Slot &slot, float *outScores, int32_t *outIds, int32_t *status)
{
    BM25MI_HIP(hipStreamSynchronize(slot.stream), "complete Search");
    bm25mi_stats &st = slot.last;
    float ms = -1;
    if (slot.timed)
        BM25MI_HIP(hipEventElapsedTime(&ms, slot.ev[0], slot.ev[3]), "device time");
    st.device_ms = ms;
    st.score_kernel_ms = st.final_ms = -1;
    if (slot.timed) {
        BM25MI_HIP(hipEventElapsedTime(&ms, slot.ev[1], slot.ev[2]), "score kernel time");
        st.score_kernel_ms = ms;
        BM25MI_HIP(hipEventElapsedTime(&ms, slot.ev[2], slot.ev[3]), "final time");
        st.final_ms = ms;
    }
    auto *res = static_cast<unsigned char *>(slot.hResult);
    const auto *rows = reinterpret_cast<const unsigned long long *>(res);
    unsigned long long stats[kStats] = {};
    for (int32_t r = 0; r < slot.finalGrid; ++r)
        for (int i = 0; i < kStats; ++i)
            stats[i] = i >= kStatMax ? std::max(stats[i], rows[r * kStats + i]) : stats[i] + rows[r * kStats + i];
    const auto *stq = reinterpret_cast<const int32_t *>(res + resStatus);
    const auto *sc = reinterpret_cast<const float *>(res + resScores);
    const auto *id = reinterpret_cast<const int32_t *>(res + resIds);
    if (stats[kStatOverflow])
        throw DeviceError(0, "Candidate pool overflow");
    int32_t fallbacks = 0;
    const auto &out = slot.out;
    if (TakeInjected("result") && slot.nq > 0 && !stq[0]) {
        // Injected corruption of the first query's rows, for the checks below to reject.
        auto *idw = reinterpret_cast<int32_t *>(res + resIds);
        auto *scw = reinterpret_cast<float *>(res + resScores);
        if (out[1] - out[0] >= 2) {
            std::swap(idw[0], idw[1]);
            std::swap(scw[0], scw[1]);
        } else {
            idw[0] = static_cast<int32_t>(ndocs);
        }
    }
    for (int32_t q = 0; q < slot.nq; ++q) {
        status[q] = stq[q];
        fallbacks += stq[q] ? 1 : 0;
        if (stq[q])
            continue;
        // Receive checks of large_search.cpp:64-85.
        for (int64_t r = out[q]; r < out[q + 1]; ++r) {
            if (!(id[r] >= 0 && id[r] < ndocs && std::isfinite(sc[r]) && sc[r] >= 0))
                throw DeviceError(0, "Invalid real-document result");
            if (r > out[q] && !(sc[r - 1] > sc[r] || (sc[r - 1] == sc[r] && id[r - 1] < id[r])))
                throw DeviceError(0, "Strict global ranking failure");
        }
        std::memcpy(outScores + out[q], sc + out[q], (out[q + 1] - out[q]) * 4);
        std::memcpy(outIds + out[q], id + out[q], (out[q + 1] - out[q]) * 4);
    }
    st.postings_scanned = static_cast<int64_t>(stats[kStatPostings]);
    st.postings_in_range = static_cast<int64_t>(stats[kStatUseful]);
    st.emitting_blocks = static_cast<int64_t>(stats[kStatEmitting]);
    st.exact_blocks = static_cast<int64_t>(stats[kStatExact]);
    st.postings_total = slot.total;
    st.score_grid = grid;
    for (int i = 0; i < 4; ++i)
        st.phase_cycles[i] = static_cast<int64_t>(stats[kStatPhase + i]);
    st.wave_busy_max = static_cast<int64_t>(stats[kStatMax]);
    st.wave_span = static_cast<int64_t>(stats[kStatMax + 1] - ~stats[kStatMax + 2]);
    st.wave_start_spread = static_cast<int64_t>(stats[kStatMax + 3] - ~stats[kStatMax + 2]);
    st.block_docs = blockDocs;
    st.blocks_scored = static_cast<int64_t>(stats[kStatScored]);
    st.blocks_skipped = static_cast<int64_t>(stats[kStatSkipped]);
    st.candidates = static_cast<int64_t>(stats[kStatCandidates]);
    st.fallbacks = fallbacks;
    st.queries = slot.nq;
}

void Session::Search(const Query &query, float *outScores, int32_t *outIds, int32_t *status, bool timed)
{
    Require(outScores && outIds, "Null output array");
    Slot &slot = slots[0];
    Pack(query, slot); // input errors leave the session usable
    try {
        FailIfInjected("search");
        Launch(slot, timed);
        Complete(slot, outScores, outIds, status);
        last = slot.last;
    } catch (...) {
        failed = true;
        throw;
    }
}

void Session::SearchScratch(const Query &query, float *outScores, int32_t *outIds, int32_t *status)
{
    Require(outScores && outIds && status, "Null output array");
    Require(!query.thr, "The no-allocation Search takes no thresholds");
    // Scratch layout of this Search: packet | count | task counts | tau | histograms | counter rows
    // | pools. Every task of a plain query writes its count, so the task counts need no clearing.
    const size_t q = static_cast<size_t>(cfg.max_queries);
    size_t at = Align(packetBytes, 256);
    const size_t offCount = at;
    at = Align(at + q * 4, 256);
    const size_t offBlockCount = at;
    at = Align(at + q * static_cast<size_t>(BlockCountStride(blocks)) * 4, 256);
    const size_t offTau = at;
    at = Align(at + q * 4, 256);
    const size_t offHist = at;
    at = Align(at + q * kHistBins * 4, 256);
    const size_t offRows = at;
    at = Align(at + static_cast<size_t>(grid) * kStats * 8, 256);
    Require(at < kScratchBytes, "The session's packet and counters exceed the no-allocation scratch");
    const int64_t poolCap = static_cast<int64_t>((kScratchBytes - at) / 8);
    Slot &slot = slots[0];
    Pack(query, slot, poolCap); // input errors (a batch whose pools exceed the scratch too) keep the session
    auto *base = static_cast<unsigned char *>(scratchBase);
    BatchView bt = slot.bt;
    const auto rebase = [&](const void *p) {
        return base + (static_cast<const unsigned char *>(p) - static_cast<const unsigned char *>(slot.dPacket));
    };
    bt.header = reinterpret_cast<const int32_t *>(rebase(bt.header));
    bt.qoff = reinterpret_cast<const int64_t *>(rebase(bt.qoff));
    bt.k = reinterpret_cast<const int32_t *>(rebase(bt.k));
    bt.thr = reinterpret_cast<const float *>(rebase(bt.thr));
    bt.tau0 = reinterpret_cast<const uint32_t *>(rebase(bt.tau0));
    bt.poolOff = reinterpret_cast<const int64_t *>(rebase(bt.poolOff));
    bt.outOff = reinterpret_cast<const int64_t *>(rebase(bt.outOff));
    bt.words = reinterpret_cast<const uint4 *>(rebase(bt.words));
    bt.costPrefix = reinterpret_cast<const int64_t *>(rebase(bt.costPrefix));
    WorkView wk = slot.wk; // status, scores and IDs: the slot's mapped result block
    wk.pool = reinterpret_cast<unsigned long long *>(base + at);
    wk.count = reinterpret_cast<uint32_t *>(base + offCount);
    wk.blockCount = reinterpret_cast<uint32_t *>(base + offBlockCount);
    wk.tau = reinterpret_cast<uint32_t *>(base + offTau);
    wk.hist = reinterpret_cast<uint32_t *>(base + offHist);
    wk.stats = reinterpret_cast<unsigned long long *>(base + offRows);
    wk.itemMask = nullptr; // threshold planning only
    wk.dead = wk.cursor = nullptr;
    wk.live = nullptr;
    try {
        slot.timed = false;
        LaunchStage(static_cast<const uint4 *>(slot.hPacketDev), reinterpret_cast<uint4 *>(base), slot.ranges,
                    slot.btHost, nullptr, Model(), 0, slot.stream);
        // The scratch keeps the last call's contents, possibly another session's: the counters
        // are set by a fill kernel before scoring.
        LaunchInit(wk, cfg.max_queries, slot.stream);
        Require(LaunchScore(blockDocs, ix, bt, wk, blocks, grid, slot.stream, nullptr, nullptr, 0, cfg.keep_cap, slot.nq),
                "Unsupported block size");
        slot.finalGrid = FinalGrid(slot.nq);
        LaunchFinal(bt, wk, slot.finalGrid, static_cast<unsigned long long *>(slot.hResultDev), grid, blocks, blockDocs,
                    cfg.keep_cap, slot.stream);
        BM25MI_HIP(hipGetLastError(), "launch Search kernels (scratch)");
        Complete(slot, outScores, outIds, status);
        last = slot.last;
    } catch (...) {
        failed = true;
        throw;
    }
}

void Session::PredictThresholds(const Query &query, float *out)
{
    Require(out != nullptr, "Null threshold output");
    Require(dModel != nullptr, "Learned thresholds need a threshold model (bm25mi_set_threshold_model)");
    Slot &slot = slots[0];
    Pack(query, slot);
    if (slot.mode != 3) {
        std::fill_n(out, query.count, 0.0F); // no query may take the live path
        return;
    }
    try {
        LaunchStage(static_cast<const uint4 *>(slot.hPacketDev), static_cast<uint4 *>(slot.dPacket), slot.ranges,
                    slot.btHost, reinterpret_cast<float *>(static_cast<unsigned char *>(slot.dPacket) + offThr),
                    Model(), query.count, slot.stream);
        BM25MI_HIP(hipMemcpyAsync(out, slot.bt.thr, static_cast<size_t>(query.count) * sizeof(float),
                                  hipMemcpyDeviceToHost, slot.stream),
                   "thresholds D2H");
        BM25MI_HIP(hipStreamSynchronize(slot.stream), "predict thresholds");
    } catch (...) {
        failed = true;
        throw;
    }
}

void Session::SetLaunch(int32_t waves, int64_t cost)
{
    Require(waves >= 1 && waves <= gridCapacity, "waves out of range for this session");
    Require(cost >= 1, "task cost must be >= 1");
    for (auto &slot : slots)
        BM25MI_HIP(hipStreamSynchronize(slot.stream), "drain before relaunch");
    grid = waves;
    taskCost = cost;
    for (auto &slot : slots) {
        for (int g = 0; g < 4; ++g) {
            if (slot.exec[g])
                BM25MI_HIP(hipGraphExecDestroy(slot.exec[g]), "destroy graph exec");
            if (slot.graph[g])
                BM25MI_HIP(hipGraphDestroy(slot.graph[g]), "destroy graph");
            slot.exec[g] = nullptr;
            slot.graph[g] = nullptr;
        }
        if (cfg.use_graph)
            Capture(slot);
    }
}

void Session::Bootstrap()
{
    const int32_t q = cfg.max_queries, k = cfg.max_k;
    std::vector<int64_t> offsets(q + 1, 0);
    std::vector<int32_t> ks(q, k), status(q);
    std::vector<float> scores(static_cast<size_t>(q) * k);
    std::vector<int32_t> ids(scores.size());
    // Each query needs PoolCapacity(max_k); a smaller pool budget bootstraps fewer at a time.
    const int32_t per = static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(q, poolEntries / PoolCapacity(k))));
    for (auto &slot : slots) {
        for (int32_t q0 = 0; q0 < q; q0 += per) {
            const int32_t nq = std::min(per, q - q0);
            Query query{offsets.data(), nullptr, nq, ks.data(), nullptr};
            Pack(query, slot);
            Launch(slot, false);
            Complete(slot, scores.data(), ids.data(), status.data());
            for (size_t i = 0; i < static_cast<size_t>(nq) * k; ++i)
                if (!(scores[i] == 0 && ids[i] == static_cast<int32_t>(i % k)))
                    throw DeviceError(0, "Empty bootstrap Search failed");
        }
    }
}

void Session::Release() noexcept
{
    for (auto &slot : slots) {
        for (int g = 0; g < 4; ++g) {
            if (slot.exec[g])
                (void)hipGraphExecDestroy(slot.exec[g]);
            if (slot.graph[g])
                (void)hipGraphDestroy(slot.graph[g]);
            slot.exec[g] = nullptr;
            slot.graph[g] = nullptr;
        }
        for (auto &e : slot.ev)
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        if (slot.stream) {
            (void)hipStreamSynchronize(slot.stream);
            (void)hipStreamDestroy(slot.stream);
            slot.stream = nullptr;
        }
        for (void **p : {&slot.dPacket, &slot.dPool, &slot.dCount, &slot.dBlockCount, &slot.dTau, &slot.dStats,
                         &slot.dHist, &slot.dItems, &slot.dDead, &slot.dCursor, &slot.dLive})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        for (void **p : {&slot.hPacket, &slot.hResult})
            if (*p) {
                (void)hipHostFree(*p);
                *p = nullptr;
            }
    }
    for (void **p : {&dPost, &dSamples, &dDir, &dBmax, &dModel})
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
}

} // namespace bm25mi

using namespace bm25mi;

extern "C" void bm25mi_config_default(bm25mi_config *cfg)
{
    if (!cfg)
        return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device_id = 0;
    cfg->max_queries = 8;
    cfg->max_words = 24;
    cfg->max_k = 10;
    cfg->block_docs = 2048;
    cfg->k1 = 1.2;
    cfg->use_graph = 0;
    cfg->grid_blocks = 0;
    cfg->task_cost = 0;
    cfg->slots = 1;
    cfg->pool_entries = 0;
    cfg->live_max_k = 0;
    cfg->keep_cap = 0;
}

extern "C" int bm25mi_create(const int64_t *term_offsets, int64_t vocab, const int32_t *doc_ids, const float *tf,
                             int64_t postings, const float *doc_norm, int64_t docs, const float *idf,
                             const bm25mi_config *cfg, bm25mi_session **out)
{
    return Invoke([&] {
        Require(out != nullptr, "Null session output");
        *out = nullptr;
        bm25mi_config c;
        bm25mi_config_default(&c);
        if (cfg)
            c = *cfg;
        FailIfInjected("create");
        auto s = std::make_unique<bm25mi_session>();
        s->impl = std::make_unique<Session>(term_offsets, vocab, doc_ids, tf, postings, doc_norm, docs, idf, c);
        *out = s.release();
    });
}

static int SearchCommon(bm25mi_session *s, const Query &query, float *scores, int32_t *ids, int32_t *fallback,
                        bool pruned, bool timed)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            const auto start = std::chrono::steady_clock::now();
            if (!s->impl || s->impl->closed || s->impl->failed)
                throw InvalidState("BM25 session closed/failed");
            if (s->impl->serving)
                throw InvalidState("BM25 session is serving: submit through its server");
            Session &ss = *s->impl;
            if (query.learned && !ss.dModel)
                throw InvalidState("Learned Search needs a threshold model (bm25mi_set_threshold_model)");
            std::vector<int32_t> status(std::max(0, query.count));
            ss.Search(query, scores, ids, status.data(), timed);
            int32_t reruns = 0;
            if (pruned) {
                // Plain-path rerun of every query that kept fewer than K docs above its threshold.
                std::vector<int32_t> which;
                for (int32_t q = 0; q < query.count; ++q) {
                    if (fallback)
                        fallback[q] = status[q];
                    if (status[q])
                        which.push_back(q);
                }
                reruns = static_cast<int32_t>(which.size());
                if (!which.empty()) {
                    const bm25mi_stats first = ss.last;
                    std::vector<int64_t> offs(1, 0);
                    std::vector<int32_t> terms, ks, st(which.size());
                    std::vector<int64_t> rowOf;
                    int64_t rows = 0;
                    std::vector<int64_t> prefix(query.count + 1, 0);
                    for (int32_t q = 0; q < query.count; ++q)
                        prefix[q + 1] = prefix[q] + query.k[q];
                    for (int32_t q : which) {
                        terms.insert(terms.end(), query.terms + query.offsets[q], query.terms + query.offsets[q + 1]);
                        offs.push_back(static_cast<int64_t>(terms.size()));
                        ks.push_back(query.k[q]);
                        rowOf.push_back(rows);
                        rows += query.k[q];
                    }
                    std::vector<float> sc(rows);
                    std::vector<int32_t> id(rows);
                    Query sub{offs.data(), terms.data(), static_cast<int32_t>(which.size()), ks.data(), nullptr};
                    ss.Search(sub, sc.data(), id.data(), st.data(), timed);
                    for (size_t i = 0; i < which.size(); ++i) {
                        const int32_t q = which[i];
                        std::memcpy(scores + prefix[q], sc.data() + rowOf[i], query.k[q] * 4);
                        std::memcpy(ids + prefix[q], id.data() + rowOf[i], query.k[q] * 4);
                    }
                    if (first.device_ms >= 0 && ss.last.device_ms >= 0)
                        ss.last.device_ms += first.device_ms;
                    if (first.score_kernel_ms >= 0 && ss.last.score_kernel_ms >= 0)
                        ss.last.score_kernel_ms += first.score_kernel_ms;
                    ss.last.postings_scanned += first.postings_scanned;
                    ss.last.postings_in_range += first.postings_in_range;
                    ss.last.postings_total = first.postings_total;
                    ss.last.blocks_scored += first.blocks_scored;
                    ss.last.blocks_skipped += first.blocks_skipped;
                    ss.last.candidates += first.candidates;
                    ss.last.queries = query.count;
                }
            } else if (fallback) {
                for (int32_t q = 0; q < query.count; ++q)
                    fallback[q] = 0;
            }
            ss.last.fallbacks = reruns;
            ss.last.host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        },
        &s->error);
}

extern "C" int bm25mi_search(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                             int32_t queries, int32_t k, float *out_scores, int32_t *out_ids)
{
    std::vector<int32_t> ks(queries > 0 ? queries : 0, k);
    Query q{query_offsets, query_terms, queries, ks.data(), nullptr};
    return SearchCommon(s, q, out_scores, out_ids, nullptr, false, false);
}

extern "C" int bm25mi_search_batch(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                   int32_t queries, const int32_t *k_per_query, float *out_scores, int32_t *out_ids)
{
    Query q{query_offsets, query_terms, queries, k_per_query, nullptr};
    return SearchCommon(s, q, out_scores, out_ids, nullptr, false, false);
}

extern "C" int bm25mi_search_learned(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                     int32_t queries, const int32_t *k_per_query, int32_t timed, float *out_scores,
                                     int32_t *out_ids, int32_t *fallback_out)
{
    Query q{query_offsets, query_terms, queries, k_per_query, nullptr};
    q.learned = true;
    return SearchCommon(s, q, out_scores, out_ids, fallback_out, true, timed != 0);
}

extern "C" int bm25mi_learned_thresholds(bm25mi_session *s, const int64_t *query_offsets,
                                         const int32_t *query_terms, int32_t queries, const int32_t *k_per_query,
                                         float *out_thresholds)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            if (!s->impl || s->impl->closed || s->impl->failed)
                throw InvalidState("BM25 session closed/failed");
            if (s->impl->serving)
                throw InvalidState("BM25 session is serving: submit through its server");
            Query q{query_offsets, query_terms, queries, k_per_query, nullptr};
            q.learned = true;
            s->impl->PredictThresholds(q, out_thresholds);
        },
        &s->error);
}

extern "C" int bm25mi_set_threshold_model(bm25mi_session *s, const float *model, int64_t count, float y_mean,
                                          float y_scale, float safety)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            Require(s->impl && !s->impl->closed && !s->impl->failed, "BM25 session closed/failed");
            Require(!s->impl->serving, "Stop the session's server before changing its model");
            constexpr int64_t kParams = int64_t(kHidden) * kFeatures + kHidden + kHidden * kHidden + kHidden + kHidden + 1;
            Require(model && count == kParams + 2 * kFeatures, "Model must hold the MLP parameters, mean and scale");
            for (int64_t i = 0; i < count; ++i)
                Require(std::isfinite(model[i]), "Model values must be finite");
            for (int i = 0; i < kFeatures; ++i)
                Require(model[kParams + kFeatures + i] > 0.0F, "Feature scales must be positive");
            Require(std::isfinite(y_mean) && std::isfinite(y_scale) && y_scale > 0.0F && std::isfinite(safety) &&
                        safety > 0.0F,
                    "Invalid model output transform");
            Session &ss = *s->impl;
            // Device layout: both weight matrices regrouped so that hidden unit j's four inputs
            // 4v .. 4v+3 are float4 v * 64 + j (one coalesced 16-byte load per lane and v).
            std::vector<float> dev(model, model + count);
            const auto regroup = [&](int64_t at, int inputs) {
                for (int j = 0; j < kHidden; ++j)
                    for (int i = 0; i < inputs; ++i)
                        dev[at + (static_cast<int64_t>(i / 4) * kHidden + j) * 4 + i % 4] = model[at + j * inputs + i];
            };
            regroup(0, kFeatures);
            regroup(int64_t(kHidden) * kFeatures + kHidden, kHidden);
            try {
                BM25MI_HIP(hipSetDevice(ss.device), "set device");
                if (!ss.dModel)
                    BM25MI_HIP(hipMalloc(&ss.dModel, static_cast<size_t>(count) * sizeof(float)), "alloc model");
                BM25MI_HIP(hipMemcpy(ss.dModel, dev.data(), static_cast<size_t>(count) * sizeof(float),
                                     hipMemcpyHostToDevice),
                           "upload model");
            } catch (const DeviceError &) {
                ss.failed = true;
                throw;
            }
            ss.modelYMean = y_mean;
            ss.modelYScale = y_scale;
            ss.modelSafety = safety;
            if (ss.cfg.use_graph) {
                try {
                    for (auto &slot : ss.slots) {
                        if (slot.exec[3])
                            BM25MI_HIP(hipGraphExecDestroy(slot.exec[3]), "destroy learned graph exec");
                        if (slot.graph[3])
                            BM25MI_HIP(hipGraphDestroy(slot.graph[3]), "destroy learned graph");
                        slot.exec[3] = nullptr;
                        slot.graph[3] = nullptr;
                        ss.Capture(slot);
                    }
                } catch (const DeviceError &) {
                    ss.failed = true;
                    throw;
                }
            }
        },
        &s->error);
}

extern "C" int bm25mi_search_batch_noalloc(bm25mi_session *s, const int64_t *query_offsets,
                                           const int32_t *query_terms, int32_t queries, const int32_t *k_per_query,
                                           float *out_scores, int32_t *out_ids)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            const auto start = std::chrono::steady_clock::now();
            if (!s->impl || s->impl->closed || s->impl->failed)
                throw InvalidState("BM25 session closed/failed");
            if (s->impl->serving)
                throw InvalidState("BM25 session is serving: submit through its server");
            Session &ss = *s->impl;
            Query q{query_offsets, query_terms, queries, k_per_query, nullptr};
            ss.SearchScratch(q, out_scores, out_ids, ss.slots[0].status.data());
            ss.last.fallbacks = 0;
            ss.last.host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        },
        &s->error);
}

extern "C" int bm25mi_search_pruned(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                    int32_t queries, const int32_t *k_per_query, const float *thresholds,
                                    float *out_scores, int32_t *out_ids, int32_t *fallback_out)
{
    Query q{query_offsets, query_terms, queries, k_per_query, thresholds};
    return SearchCommon(s, q, out_scores, out_ids, fallback_out, true, false);
}

// Same as bm25mi_search_batch / bm25mi_search_pruned, but runs the kernels without the
// graph and with HIP events around the scoring kernel (bm25mi_stats.score_kernel_ms).
extern "C" int bm25mi_search_timed(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                   int32_t queries, const int32_t *k_per_query, const float *thresholds,
                                   float *out_scores, int32_t *out_ids, int32_t *fallback_out)
{
    Query q{query_offsets, query_terms, queries, k_per_query, thresholds};
    return SearchCommon(s, q, out_scores, out_ids, fallback_out, thresholds != nullptr, true);
}

// Request checks shared by the per-query statistics entry points (features, bounds): the rules
// of Session::Pack that do not depend on the session's capacity (offsets start at 0 and ascend,
// at most 64 terms per query, term IDs in [0, INT32_MAX), K >= 1). They read only tables that
// are fixed at create time, so they may run while a server owns the session.
static void CheckQueries(const int64_t *offsets, const int32_t *terms, int32_t queries, const int32_t *ks)
{
    Require(queries >= 1 && offsets && ks, "Null query offsets or K array");
    Require(offsets[0] == 0, "query_offsets[0] must be 0");
    for (int32_t q = 0; q < queries; ++q) {
        const int64_t len = offsets[q + 1] - offsets[q];
        Require(len >= 0 && len <= kMaxTerms, "Query exceeds the 64-term limit or has negative length");
        Require(ks[q] >= 1, "K out of range");
    }
    Require(offsets[queries] == 0 || terms, "Null query terms");
    for (int64_t w = 0; w < offsets[queries]; ++w)
        Require(terms[w] >= 0 && terms[w] < INT32_MAX, "Invalid public word ID");
}

// The K-th largest term impact among the query's terms (a lower bound of the K-th score).
static double LowerBound(const Session &ss, const int32_t *terms, int64_t n, int32_t k)
{
    uint32_t bits = ss.StartBound(terms, n, k);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// Per-query features of the learned K-th-score model (Session::Features); they depend only on
// index statistics of the query's terms.
extern "C" int bm25mi_query_features(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                     int32_t queries, const int32_t *k_per_query, float *out_features)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            Require(s->impl && !s->impl->closed, "Session closed");
            Require(out_features != nullptr, "Null feature output");
            CheckQueries(query_offsets, query_terms, queries, k_per_query);
            const Session &ss = *s->impl;
            for (int32_t q = 0; q < queries; ++q)
                ss.Features(query_terms + query_offsets[q], query_offsets[q + 1] - query_offsets[q], k_per_query[q],
                            out_features + static_cast<int64_t>(q) * BM25MI_FEATURES, nullptr);
        },
        &s->error);
}

extern "C" int bm25mi_query_bounds(bm25mi_session *s, const int64_t *query_offsets, const int32_t *query_terms,
                                   int32_t queries, const int32_t *k_per_query, float *out_lower, float *out_upper)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            Require(s->impl && !s->impl->closed, "Session closed");
            Require(out_lower || out_upper, "Null bound outputs");
            CheckQueries(query_offsets, query_terms, queries, k_per_query);
            const Session &ss = *s->impl;
            for (int32_t q = 0; q < queries; ++q) {
                const int64_t n = query_offsets[q + 1] - query_offsets[q];
                const int32_t *t = query_terms + query_offsets[q];
                if (out_lower)
                    out_lower[q] = static_cast<float>(LowerBound(ss, t, n, k_per_query[q]));
                if (out_upper) {
                    // Sequential FP32 sum of the term maxima in query order: FP32 addition is
                    // monotone, so no doc's score exceeds it.
                    float ub = 0.0F;
                    for (int64_t w = 0; w < n; ++w)
                        if (t[w] < ss.vocab)
                            ub += ss.tmaxHost[t[w]];
                    out_upper[q] = ub;
                }
            }
        },
        &s->error);
}

extern "C" int bm25mi_set_launch(bm25mi_session *s, int32_t grid_waves, int32_t task_cost)
{
    if (!s)
        return Invoke([] { throw InvalidState("BM25 session is not initialized"); });
    std::lock_guard<std::mutex> guard(s->mutex);
    return Invoke(
        [&] {
            Require(s->impl && !s->impl->closed && !s->impl->failed, "BM25 session closed/failed");
            Require(!s->impl->serving, "Stop the session's server before changing its launch");
            Session &ss = *s->impl;
            try {
                ss.SetLaunch(grid_waves > 0 ? grid_waves : ss.grid, task_cost > 0 ? task_cost : ss.taskCost);
            } catch (const DeviceError &) {
                ss.failed = true;
                throw;
            }
        },
        &s->error);
}

#ifdef BM25MI_PHASES
// Measurement builds only: the per-wave counter rows of slot 0's last Search
// ([max(scoring grid, live grid)][16]).
extern "C" int bm25mi_debug_wave_rows(bm25mi_session *s, unsigned long long *out, int64_t rows)
{
    return Invoke([&] {
        Require(s && s->impl && out, "Null session or output");
        std::lock_guard<std::mutex> guard(s->mutex);
        Session &ss = *s->impl;
        const int64_t n = std::max(ss.grid, ss.liveGrid);
        Require(rows >= n, "Output holds fewer rows than the scoring and live grids");
        BM25MI_HIP(hipMemcpy(out, ss.slots[0].dStats, sizeof(unsigned long long) * kStats * n, hipMemcpyDeviceToHost),
                   "copy wave rows");
    });
}
#endif

extern "C" int bm25mi_last_stats(const bm25mi_session *s, bm25mi_stats *out)
{
    return Invoke([&] {
        Require(s && out, "Null session or stats output");
        auto *ms = const_cast<bm25mi_session *>(s);
        std::lock_guard<std::mutex> guard(ms->mutex);
        Require(ms->impl != nullptr, "Session closed");
        *out = ms->impl->last;
    });
}

extern "C" int bm25mi_destroy(bm25mi_session *s)
{
    if (!s)
        return BM25MI_OK; // Close is idempotent
    int rc;
    {
        std::lock_guard<std::mutex> guard(s->mutex);
        rc = Invoke(
            [&] {
                if (s->impl) {
                    if (s->impl->serving)
                        throw InvalidState("Stop the session's server before Close");
                    FailIfInjected("close"); // the session stays intact: Close can be retried
                    s->impl->closed = true;
                    BM25MI_HIP(hipSetDevice(s->impl->device), "set device for close");
                    s->impl.reset();
                }
            },
            &s->error);
    }
    if (rc == BM25MI_OK)
        delete s;
    return rc;
}

extern "C" int bm25mi_inject_failure(const char *stage, int64_t device_code, const char *message)
{
    return Invoke([&] {
        std::lock_guard<std::mutex> g(injectMutex);
        if (!stage) {
            for (auto &f : injected)
                f = Injected{};
            return;
        }
        const int i = StageIndex(stage);
        Require(i >= 0, "Unknown failure stage (create, search, close or result)");
        injected[i] = Injected{true, device_code, message ? message : ""};
    });
}

extern "C" int64_t bm25mi_last_error(const bm25mi_session *s, char *buf, size_t size)
{
    const ErrorInfo &info = s ? s->error : ThreadError();
    if (buf && size) {
        std::strncpy(buf, info.message, size - 1);
        buf[size - 1] = '\0';
    }
    return info.code;
}

extern "C" int bm25mi_device_info(int32_t device_id, int32_t *cus, int32_t *lds_per_block, int32_t *lds_per_cu)
{
    return Invoke([&] {
        int count = 0;
        BM25MI_HIP(hipGetDeviceCount(&count), "count devices");
        Require(device_id >= 0 && device_id < count, "device_id out of range");
        hipDeviceProp_t prop;
        BM25MI_HIP(hipGetDeviceProperties(&prop, device_id), "device properties");
        if (cus)
            *cus = prop.multiProcessorCount;
        if (lds_per_block)
            *lds_per_block = static_cast<int32_t>(prop.sharedMemPerBlock);
        if (lds_per_cu)
            *lds_per_cu = static_cast<int32_t>(prop.maxSharedMemoryPerMultiProcessor);
    });
}
