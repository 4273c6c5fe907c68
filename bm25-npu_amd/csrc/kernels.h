// Device-side data structures shared by the Search kernels and the host session.
#pragma once

#include <cstdint>

namespace bm25mi {

constexpr int kWave = 64;                  // CDNA wavefront
constexpr int kMaxTerms = 64;              // terms per query (fixed_query_packet.h:32): one lane per term
// Docs per directory / block-max granule: the unit the threshold path skips or scores. 128 docs:
// on the 10M-doc corpus at thresholds of 0.93 x the K-th score, the live granules of Q8 x 3-term
// batches hold 5.3% (seed 1) / 2.6% (seed 2) of the query terms' postings, against 13% / 6.6%
// with 256-doc granules, and are fewer (25.8K against 31.9K).
constexpr int kGranule = 128;
constexpr int kGranuleShift = 7;
constexpr int kLongMin = 4096;             // lists this long use the granule directory; shorter ones
                                           // are located through the posting samples
constexpr int kSampleShift = 4;            // samples[i] = doc ID of posting 16 * i (global CSC index)
constexpr int kFinalThreads = 1024;        // selection workgroup of the final kernel
constexpr int kSortCap = 16384;            // pool entries the final kernel sorts in LDS (128 KB)
constexpr int kTaskCost = 1024;            // estimated fixed cost of a scoring task, in postings
// A threshold query takes the live path only when its threshold is at least this fraction of the
// FP32 sum of its terms' largest impacts (below it, most granules stay live; measured on the
// 10M-doc corpus: 0.64 of the postings streamed under 0.3, 0.13-0.30 above it).
constexpr float kLiveMinRatio = 0.3F;
// ... and only when its K is at most kLiveMaxK: at K = 100 the live path (a cut of up to K docs per
// live granule) measured 0.92-1.0x of the plain path on the 10M-doc corpus.
constexpr int32_t kLiveMaxK = 64;
// Live granules whose appends a threshold query's candidate pool reserves room for beyond the
// plain path's capacity (which covers many more on large indexes: 312K entries at K = 10 on the
// 10M-doc corpus, against 78K granules x 10 for every granule). The learned thresholds keep a
// few hundred docs per batch there; a live pool that overflows reruns its query on the plain path.
constexpr int64_t kLiveReserveGranules = 4096;
#ifndef BM25MI_COLUMNS
#define BM25MI_COLUMNS 14
#endif
// 64-posting columns a scoring wave keeps in flight per round (14: 0.1175 ms end to end on the
// 10M-doc Q8 x 3-term K10 8-batch rotation against 0.1205 for 16, which needs one register more
// than the 128 of four waves per SIMD and spills it), and the live kernel's.
constexpr int kColumns = BM25MI_COLUMNS;
#ifndef BM25MI_LIVE_COLUMNS
#define BM25MI_LIVE_COLUMNS 12
#endif
constexpr int kLiveColumns = BM25MI_LIVE_COLUMNS;
constexpr int kStats = 16;
// Item waves of the planning kernel of threshold queries: a multiple of 256 up to this many (the
// live kernel reads their live counts, 64 per lane at most, in one round of 16-byte loads).
constexpr int kMaxPruneWaves = 4096;
constexpr int kPruneItems = 4; // items (64 granules of one query) a planning wave bounds together
// Score histogram of kept candidates, per query: bucket = (score bits >> 19) - kHistBase,
// i.e. 16 buckets per octave over [2^-4, 2^8) (lower scores fall in bucket 0, higher ones
// in the last bucket, which only makes the bound it yields smaller).
constexpr int kHistBins = 192;
// The scoring kernel reads a query's score histogram every kHistEvery-th task of the query, or
// every kHistEveryLargeK-th one when K > kHistLargeK (a block may then append up to 2K docs, so
// a fresher bound pays: uneven 8.84M-passage batches at Q64, 1.63 -> 1.29 ms).
constexpr uint32_t kHistEvery = 16;
constexpr uint32_t kHistEveryLargeK = 4;
constexpr uint32_t kHistLargeK = 64;
constexpr uint32_t kHistBase = (127u - 4u) << 4;
// The scoring kernel adds a wave's appends to the query's histogram at once when the batch has at
// most this many scoring waves per query (measured on the 10M-doc corpus, Q64 x 2-12 terms: fewer
// candidates and a faster kernel than adding them when the wave's own count doubles); with more
// (Q8: 512 per query), the atomics on a query's histogram lines contend, and a wave adds its
// appends when its own count has doubled and when it leaves the query.
constexpr int kEagerWaves = 128;

// Candidates a block may append: all of them when at most this many score at or above the
// query's floor, else a cut keeping between K and this many. It sizes the candidate pools.
// keepCap > 0 (the session's keep_cap) replaces the default max(2K, 64) with max(K, keepCap).
__host__ __device__ constexpr int32_t BlockKeepCap(int32_t k, int32_t blockDocs, int32_t keepCap = 0)
{
    const int32_t c = keepCap > 0 ? (keepCap > k ? keepCap : k) : (2 * k > 64 ? 2 * k : 64);
    return c < blockDocs ? c : blockDocs;
}

// Row stride of the per-task candidate counts: blocks rounded up to 8, so that each row starts on
// 32 bytes and the final kernel loads 8 counts as two 16-byte words.
__host__ __device__ constexpr int64_t BlockCountStride(int32_t blocks)
{
    return (static_cast<int64_t>(blocks) + 7) & ~int64_t(7);
}

// Stats slots (unsigned long long).
enum {
    kStatPostings = 0, // postings streamed (including masked window postings of short lists)
    kStatScored = 1,   // (query, block) tasks scored
    kStatSkipped = 2,  // tasks skipped: no postings, or every granule bound below the threshold
    kStatCandidates = 3,
    kStatOverflow = 4,
    kStatEmitting = 5, // tasks that appended at least one candidate
    kStatExact = 6,    // tasks that needed a cut (more candidates than they may append)
    kStatUseful = 7,   // postings inside their task's doc range (the plain-path work)
    kStatPhase = 8,    // 8..11: scoring time per phase (locate, stream and add, select, append)
    kStatMax = 12      // 12..15 (maxima over waves): busy time, end, ~start, start
                       // (8..15 in measurement builds, -DBM25MI_PHASES, only; 100 MHz ticks)
};

// Read-only resident index.
struct IndexView {
    const uint2 *post;      // {doc ID, impact bits} per posting, CSC order
    const int32_t *samples; // doc ID of every 16th posting (global CSC index)
    const uint32_t *dir;    // [longTerms * (granules+1)]
    const float *bmax;      // [longTerms * granules]
    int64_t granules;
    int32_t ndocs;
};

// One Search batch, laid out in a single device packet uploaded per Search.
struct BatchView {
    const int32_t *header;  // [0] = queries, [1] = with a threshold (live path), [2] = plain ones
    const int64_t *qoff;    // [maxQ+1]
    const int32_t *k;       // [maxQ]
    const float *thr;       // [maxQ] pruning threshold (0 = plain): the live blocks of a threshold
                            // query are scored after the wave ranges below, strided over the waves
    const uint32_t *tau0;   // [maxQ] lower bound of the K-th score known before the Search
                            // (score bits): the largest K-th term impact over the query's terms
    const int64_t *poolOff; // [maxQ+1]
    const int64_t *outOff;  // [maxQ+1]
    // Per query word, resolved on the host from the term ID: {first posting, posting
    // count (0 for IDs >= vocab), long-term slot or -1, largest impact bits}.
    const uint4 *words;     // [maxWords]
    // Wave plan: costPrefix[q] = sum over queries p < q of (task cost of p) * blocks, in posting
    // equivalents (0 for threshold queries). Scoring wave w takes the tasks of the query-major
    // task order (task = q * blocks + blk) from cost costPrefix[nq] * w / grid on.
    const int64_t *costPrefix; // [maxQ+1]
    // Learned batches (header[3] = 1), whose thresholds the threshold kernel predicts and writes
    // into thr[], together with costPrefix and header[1..2]: per query its model features, the
    // FP32 sum of its terms' largest impacts (an upper bound of every score) and its plain-path
    // cost (task cost * blocks), used when the query is routed to the plain path.
    const float *feat;         // [maxQ][kFeatures]
    const float *upper;        // [maxQ]
    const int64_t *planCost;   // [maxQ]
};

// Ranges of a request packet to stage, in 16-byte units: the packet's arrays, each up to the
// batch's queries (or words).
constexpr int kStageRanges = 12;
struct StageRanges {
    uint32_t off[kStageRanges];
    uint32_t len[kStageRanges];
    int32_t n;
};

constexpr int kFeatures = 16;  // model inputs per query
constexpr int kHidden = 64;    // width of the model's two hidden layers

// The resident K-th-score model of learned batches and the routing rules of threshold queries.
// model: W1 b1[64] W2 b2[64] W3[64] b3[1], then the features' mean[16] and scale[16]; nullptr
// when the session has none. The weight matrices W1[64][16] and W2[64][64] (hidden unit j's
// inputs in row j) are regrouped on the device: float4 v * 64 + j holds row j's inputs
// 4v .. 4v+3, so that a wave's lanes (one per hidden unit) load them coalesced.
struct ModelView {
    const float *model;
    float yMean, yScale, safety; // threshold = safety * upper * exp(z * yScale + yMean), clamped
    float minRatio;              // a threshold below minRatio * upper routes the query to the plain path,
    int32_t liveMaxK;            // as does a K above liveMaxK
};

// Per-Search device scratch and results.
struct WorkView {
    unsigned long long *pool; // composite keys (score bits << 32 | ~doc)
    uint32_t *count;          // [maxQ] pool fill of threshold queries (live path)
    // Per plain task (q * BlockCountStride(blocks) + blk): the candidates it appended to its pool
    // region, entries [blk * BlockKeepCap, ...) of the query's pool. Every plain task writes it.
    uint32_t *blockCount;     // [maxQ * BlockCountStride(blocks)]
    uint32_t *hist;           // [maxQ * kHistBins] histogram of kept candidates' scores
    uint32_t *tau;            // [maxQ] running lower bound of the K-th score (bits)
    unsigned long long *stats;// [grid][kStats]: one counter row per scoring wave
    // Written by the planning kernel for batches with a threshold query: per item (q * items per
    // query + c, the query's granules 64c .. 64c+63) the mask of granules whose upper bound
    // reaches the query's threshold (0 for plain queries), and per planning wave the dead blocks
    // and the live granules it found (the live kernel's waves cut their shares from the latter).
    unsigned long long *itemMask; // [maxQ * ceil(granules / 64)]
    uint32_t *dead;           // [kMaxPruneWaves]
    uint32_t *live;           // [kMaxPruneWaves]
    // Also per (query word w of the batch, block) for the short words of threshold queries: the
    // index of the word's first posting with doc at or after the block's first doc.
    uint32_t *cursor;         // [maxWords * blocks]
    int32_t *status;          // [maxQ] 0 ok, 1 fewer than K survivors under a threshold
    float *scores;            // [maxQ*maxK]
    int32_t *ids;             // [maxQ*maxK]
};

} // namespace bm25mi
