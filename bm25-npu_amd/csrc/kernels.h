// Device-side data structures shared by the Search kernels and the host session.
#pragma once

#include <cstdint>

namespace bm25mi {

constexpr int kThreads = 256;              // scoring workgroup: 4 wave64
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTerms = 64;              // terms per query (fixed_query_packet.h:32)
constexpr int kMaxSegs = 64;               // posting segments per (query, block) task
constexpr int kGranule = 256;              // docs per directory / block-max granule
constexpr int kLongMin = 4096;             // lists this long use the block directory; shorter
                                           // ones a two-round 64-ary in-kernel search (64*64)
constexpr int kFinalThreads = 1024;        // selection workgroup (threshold and final kernels)
constexpr int kSortCap = 16384;            // pool entries a selection workgroup sorts in LDS (128 KB)
constexpr int kPilotStride = 16;           // every 16th block is scored first (spreads the first round)
constexpr int kEmitCap = 512;              // pool entries per block reserved beyond K (sizing only)
// Score histogram of kept candidates, per query: bucket = (score bits >> 19) - kHistBase,
// i.e. 16 buckets per octave over [2^-4, 2^8) (lower scores fall in bucket 0, higher ones
// in the last bucket, which only makes the bound it yields smaller).
constexpr int kHistBins = 192;
constexpr uint32_t kHistBase = (127u - 4u) << 4;
#ifndef BM25MI_LOADS
#define BM25MI_LOADS 8
#endif
constexpr int kLoadsPerThread = BM25MI_LOADS; // posting loads a thread keeps in flight per round
constexpr int kStats = 16;
constexpr int kCachedQueries = 256;      // per-query headers kept in LDS by the scoring kernel

// Stats slots (unsigned long long).
enum {
    kStatPostings = 0,
    kStatScored = 1,
    kStatSkipped = 2,
    kStatCandidates = 3,
    kStatOverflow = 4,
    kStatEmitting = 5, // tasks that appended at least one candidate
    kStatExact = 6,    // tasks that needed the exact K-th-score bisection
    kStatPilot = 7,    // postings streamed by the pilot pass
    kStatPhase = 8     // 8..15: scoring-kernel cycles per phase, main then pilot pass (BM25MI_PHASES builds)
};

// Read-only resident index.
struct IndexView {
    const uint2 *post;      // {doc ID, impact bits} per posting, CSC order
    const int64_t *toff;    // term offsets [vocab+1]
    const int32_t *slot;    // long-term slot or -1 [vocab]
    const uint32_t *dir;    // [longTerms * (granules+1)]
    const float *bmax;      // [longTerms * granules]
    const float *tmax;      // [vocab]
    int64_t vocab;
    int64_t granules;
    int32_t ndocs;
    int32_t granuleShift;
};

// One Search batch, laid out in a single device packet uploaded per Search.
struct BatchView {
    const int32_t *header;  // [0] = queries
    const int64_t *qoff;    // [maxQ+1]
    const int32_t *k;       // [maxQ]
    const float *thr;       // [maxQ] pruning threshold (0 = plain)
    const int64_t *poolOff; // [maxQ+1]
    const int64_t *outOff;  // [maxQ+1]
    // Per query word, resolved on the host from the term ID: {first posting, posting
    // count (0 for IDs >= vocab), long-term slot or -1, largest impact bits}.
    const uint4 *words;     // [maxWords]
};

// Per-Search device scratch and results.
struct WorkView {
    unsigned long long *pool; // composite keys (score bits << 32 | ~doc)
    uint32_t *count;          // [maxQ] pool fill
    uint32_t *hist;           // [maxQ * kHistBins] histogram of kept candidates' scores
    uint32_t *tau;            // [maxQ] running lower bound of the K-th score (bits)
    unsigned long long *stats;// [kStats]
    int32_t *status;          // [maxQ] 0 ok, 1 fewer than K survivors under a threshold
    float *scores;            // [maxQ*maxK]
    int32_t *ids;             // [maxQ*maxK]
    // Posting slice {first, end} of every (query word, block), written by the planning
    // kernel: entry blocks * qoff[q] + blk * n_q + j for word j of query q.
    // Posting segments of every (query, block) task, written by the planning kernel:
    // task (q, blk) owns segs[(q * blocks + blk) * kMaxSegs ...], {first posting,
    // (count << 8) | word}, in word order; segCount[q * blocks + blk] of them.
    uint2 *segs;
    uint32_t *segCount;
};

} // namespace bm25mi
