// BM25 full-Search kernels for gfx950 (CDNA4, wave64).
//
// Replaces the reference's two Ascend C operators
// (FR = feature_retrieval/src/ascendfaiss):
//   * Bm25FusedScoreTopK, FR/ops/ascendc/bm25/fused/op_kernel/bm25_fused_score_top_k.cpp:
//     per (query, 8192-doc leaf) FP32 score accumulation in on-chip memory and an exact
//     local top-K (:637-793 scoring, :518-635 threshold, :795-900 collection);
//   * Bm25TopKMerge, FR/ops/ascendc/bm25/merge/op_kernel/bm25_top_k_merge.cpp:10-98:
//     a 32-way tiered merge of sorted K-lists.
//
// Design here (one GPU, HBM-bound scan/scatter work, no MFMA):
//   KScore  - every wave is its own workgroup and walks (query, doc-block) tasks on its own,
//             with no barriers: it locates the query terms' postings in its block (granule
//             directory for long lists, posting samples for short ones), streams them with
//             8-byte {doc, impact} loads and adds impacts into its LDS FP32 accumulator row
//             in query-term order, so every doc's score is the same sequential FP32 sum as
//             the CPU reference (cpu_index_search.h:176-189). It keeps the docs scoring at
//             least a per-query running lower bound of the global K-th score, cut to at most
//             BlockKeepCap(K) per block, as 64-bit keys (score bits << 32 | ~doc) in the
//             query's candidate pool.
//   KFinal  - one workgroup per query selects the exact top-K of its pool (composite keys are
//             unique per doc, so ties need no extra pass), bitonic-sorts it (score desc, doc
//             asc: topk_heap.h:11-16) and writes it to pinned host memory. Fewer than K
//             positive docs are padded with the lowest-ID zero-score docs, as the reference does.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "../../include/bm25mi.h"
#include "kernels.h"

namespace bm25mi {

__device__ __forceinline__ uint32_t ReadLane(uint32_t v, int lane)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
}

// Wave-wide reductions (every lane active, every lane gets the result): the device library's
// DPP row operations and two lane reads, instead of six dependent LDS permutes.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_or_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) float __ockl_wfred_add_f32(float);

__device__ __forceinline__ uint32_t WaveSum(uint32_t v)
{
    return __ockl_wfred_add_u32(v);
}

__device__ __forceinline__ uint32_t WaveMax(uint32_t v)
{
    return __ockl_wfred_max_u32(v);
}

// Inclusive prefix sum over the 64 lanes (every lane active): row_shr 1, 2, 4, 8 within each
// 16-lane row (lanes without a source add 0), then row 0's total into row 1 and row 2's into
// row 3 (row_bcast:15), then rows 0-1's total into rows 2-3 (row_bcast:31). Six DPP adds instead
// of six dependent LDS permutes.
__device__ __forceinline__ uint32_t WaveInclusiveSum(uint32_t v)
{
    int x = static_cast<int>(v);
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true); // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true); // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true); // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true); // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15, rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31, rows 2 and 3
    return static_cast<uint32_t>(x);
}

// Block-wide sum/max over NT threads; red must hold NT/64 words. All threads call.
template <int NT> __device__ __forceinline__ uint32_t BlockSum(uint32_t v, uint32_t *red)
{
    v = WaveSum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i)
        s += red[i];
    return s;
}

template <int NT> __device__ __forceinline__ uint32_t BlockMax(uint32_t v, uint32_t *red)
{
    v = WaveMax(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i)
        s = max(s, red[i]);
    return s;
}

// Exclusive prefix of v over the block; *total receives the block sum.
template <int NT> __device__ __forceinline__ uint32_t BlockExclusive(uint32_t v, uint32_t *red, uint32_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = WaveInclusiveSum(v);
    __syncthreads();
    if (lane == 63)
        red[wave] = x;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        before += i < wave ? red[i] : 0;
        sum += red[i];
    }
    *total = sum;
    return before + x - v;
}

// Wave-aggregated append: returns this lane's slot (counter += lanes with pred), or
// 0xFFFFFFFF when pred is false. Every lane of the wave calls.
__device__ __forceinline__ uint32_t WaveAppend(bool pred, uint32_t *counter)
{
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = __lane_id();
    const uint32_t leader = mask ? static_cast<uint32_t>(__ffsll(static_cast<long long>(mask)) - 1) : 0u;
    uint32_t base = 0;
    if (mask && lane == leader)
        base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, static_cast<int>(leader));
    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
    return pred ? base + rank : 0xFFFFFFFFu;
}

// Bitonic sort of s[0, n) (n a power of two), descending. All NT threads call.
template <int NT, class T> __device__ void BitonicDesc(T *s, int n)
{
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += NT) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool desc = (lo & size) == 0;
                const T a = s[lo], b = s[hi];
                if ((a < b) == desc) {
                    s[lo] = b;
                    s[hi] = a;
                }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ uint32_t HistBucket(uint32_t bits)
{
    const uint32_t b = bits >> 19;
    return b <= kHistBase ? 0u : min(b - kHistBase, static_cast<uint32_t>(kHistBins - 1));
}

// Smallest score bits of a bucket (bucket 0 also holds scores below 2^-10: its edge is the
// smallest positive score).
__device__ __forceinline__ uint32_t HistEdge(uint32_t bucket)
{
    return bucket == 0 ? 1u : (bucket + kHistBase) << 19;
}

constexpr int kHistPerLane = (kHistBins + 63) / 64;
constexpr int kShortSlots = 4; // short lists per query with cached doc chunks in the scoring kernel

// Lower bound of a query's K-th score from its kept-candidate histogram, computed by one
// wave from the bins it holds (lane l: bins l*kHistPerLane ...): the edge of the highest
// bucket b with at least k counted scores at or above it (those k candidates score at least
// that edge). 1 when fewer than k are counted.
__device__ __forceinline__ uint32_t WaveHistBoundRegs(const uint32_t (&h)[kHistPerLane], uint32_t k)
{
    const int lane = static_cast<int>(__lane_id());
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kHistPerLane; ++i)
        sum += h[i];
    // above = counts in buckets of higher lanes (suffix over lanes).
    const uint32_t incl = WaveInclusiveSum(sum);
    uint32_t above = ReadLane(incl, kWave - 1) - incl;
    uint32_t best = 0; // bucket + 1, or 0
#pragma unroll
    for (int i = kHistPerLane - 1; i >= 0; --i) {
        above += h[i];
        if (best == 0 && above >= k)
            best = static_cast<uint32_t>(lane * kHistPerLane + i) + 1;
    }
    const uint32_t top = WaveMax(best);
    return top ? HistEdge(top - 1) : 1u;
}

__device__ __forceinline__ uint32_t WaveHistBound(const uint32_t *hist, uint32_t k)
{
    uint32_t h[kHistPerLane];
#pragma unroll
    for (int i = 0; i < kHistPerLane; ++i) {
        const int b = static_cast<int>(__lane_id()) * kHistPerLane + i;
        h[i] = b < kHistBins ? __hip_atomic_load(hist + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0u;
    }
    return WaveHistBoundRegs(h, k);
}

__device__ __forceinline__ unsigned long long WaveSum64(unsigned long long v)
{
    return __ockl_wfred_add_u64(v);
}


// k-th largest (1 <= k <= 64) of one value per lane: a bitonic sort across the 64 lanes
// through shuffles, descending, then lane k-1.
__device__ __forceinline__ uint32_t WaveKthLargest(uint32_t v, uint32_t k)
{
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t size = 2; size <= 64; size <<= 1)
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), static_cast<int>(stride)));
            const bool lower = (lane & stride) == 0, desc = (lane & size) == 0;
            v = lower == desc ? max(v, o) : min(v, o);
        }
    return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(k - 1)));
}

constexpr int Log2(int x) { return x <= 1 ? 0 : 1 + Log2(x / 2); }

typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t U32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const U32x4 LdsUint4; // 16-byte LDS words
typedef __attribute__((address_space(3))) const U32x2 LdsUint2; // 8-byte LDS words
typedef __attribute__((address_space(3))) const uint32_t LdsU32;

// Keys (score bits << log2(B) | (B-1-e)) of accumulator entries e (B of them, lane l holding
// entries l * B / 64 ...) at least (xb, xe), counted by each lane over its entries.
template <int B>
__device__ __forceinline__ uint32_t CountKeys(LdsU32 *bits, uint32_t xb, uint32_t xe)
{
    const uint32_t lane = __lane_id();
    uint32_t c = 0;
    if constexpr (B == 2 * kWave) {
        const U32x2 v = reinterpret_cast<LdsUint2 *>(bits)[lane];
        const uint32_t e0 = static_cast<uint32_t>(B - 1) - lane * 2;
        c = (v.x > xb || (v.x == xb && e0 >= xe) ? 1u : 0u) + (v.y > xb || (v.y == xb && e0 - 1 >= xe) ? 1u : 0u);
    } else {
        constexpr int kVec = B / (4 * kWave);
        static_assert(kVec >= 1 && B % (4 * kWave) == 0, "accumulator size");
        LdsUint4 *bits4 = reinterpret_cast<LdsUint4 *>(bits);
#pragma unroll 1
        for (int r = 0; r < kVec; ++r) {
            const U32x4 v = bits4[r * kWave + lane];
            const uint32_t e0 = static_cast<uint32_t>(B - 1) - (r * kWave + lane) * 4;
            c += (v.x > xb || (v.x == xb && e0 >= xe) ? 1u : 0u) + (v.y > xb || (v.y == xb && e0 - 1 >= xe) ? 1u : 0u) +
                 (v.z > xb || (v.z == xb && e0 - 2 >= xe) ? 1u : 0u) + (v.w > xb || (v.w == xb && e0 - 3 >= xe) ? 1u : 0u);
        }
    }
    return c;
}

struct CutResult {
    uint32_t tb, te, kept, mine, tauNew;
};

// The rare part of a task's selection, kept out of line so that its registers do not
// constrain the streaming loop (a call drains the wave's memory counters, so it touches
// only LDS and runs only when needed): more than cap = BlockKeepCap(K) docs score at least
// floor, so the kept set is cut to between K and cap docs, the keys at least (tb, te): at
// the K-th largest per-lane maximum when K <= 64, else by a bisection over composite keys.
// tauNew (when nonzero) is a lower bound of the query's K-th score: at least K docs of the
// block score at least it. kept/mine are the wave and lane counts of docs scoring at least
// floor, mx the lane's largest score bits.
template <int B>
__device__ __attribute__((noinline)) CutResult SelectCut(LdsU32 *bits4, uint32_t floor, uint32_t k, uint32_t cap,
                                                         uint32_t mx)
{
    constexpr int kIdxBits = Log2(B);
    static_assert(kIdxBits > 0, "block size");
    CutResult res{floor, 0, 0, 0, 0};
    uint32_t kept = cap + 1;
    if (k <= static_cast<uint32_t>(kWave)) {
        const uint32_t f = max(WaveKthLargest(mx, k), floor);
        res.mine = CountKeys<B>(bits4, f, 0);
        kept = WaveSum(res.mine);
        res.tb = f;
        if (kept <= cap)
            res.tauNew = f;
    }
    if (kept > cap) {
        // Bisection over composite keys: count(>= lo) > cap, count(>= hi) < k.
        unsigned long long lok = static_cast<unsigned long long>(res.tb) << kIdxBits;
        unsigned long long hik = (static_cast<unsigned long long>(WaveMax(mx)) + 1) << kIdxBits;
        for (int it = 0; it < 64 && hik - lok > 1; ++it) {
            const unsigned long long mid = lok + (hik - lok) / 2;
            const uint32_t mb = static_cast<uint32_t>(mid >> kIdxBits);
            const uint32_t me = static_cast<uint32_t>(mid) & static_cast<uint32_t>(B - 1);
            const uint32_t cm = WaveSum(CountKeys<B>(bits4, mb, me));
            if (cm > cap) {
                lok = mid;
            } else if (cm < k) {
                hik = mid;
            } else {
                lok = mid;
                break;
            }
        }
        res.tb = static_cast<uint32_t>(lok >> kIdxBits);
        res.te = static_cast<uint32_t>(lok) & static_cast<uint32_t>(B - 1);
        res.mine = CountKeys<B>(bits4, res.tb, res.te);
        kept = WaveSum(res.mine);
        if (kept >= k)
            res.tauNew = res.tb;
    }
    res.kept = kept;
    return res;
}

// First task of the query-major order at plan cost `target` (costPrefix: see BatchView); c =
// queries whose cost prefix ends at or below target.
__device__ __forceinline__ uint32_t PlanTask(const int64_t *prefix, int32_t nq, int32_t blocks, uint64_t target,
                                             uint32_t c)
{
    if (c >= static_cast<uint32_t>(nq))
        return static_cast<uint32_t>(nq) * static_cast<uint32_t>(blocks);
    const uint64_t p0 = static_cast<uint64_t>(prefix[c]), cost = (static_cast<uint64_t>(prefix[c + 1]) - p0) / blocks;
    const uint64_t in = (target - p0 + cost - 1) / cost;
    return c * static_cast<uint32_t>(blocks) + static_cast<uint32_t>(min(in, static_cast<uint64_t>(blocks)));
}

// This wave's contiguous range [t0, t1) of the task order: an equal share of the plan's total
// cost, found by comparing the cost prefix with the share's ends, 64 queries per load.
__device__ __forceinline__ void WaveRange(const BatchView &bt, int32_t nq, int32_t blocks, uint32_t &t0, uint32_t &t1)
{
    const uint64_t total = static_cast<uint64_t>(bt.costPrefix[nq]);
    const uint64_t lo = total * blockIdx.x / gridDim.x, hi = total * (blockIdx.x + 1) / gridDim.x;
    uint32_t c0 = 0, c1 = 0;
    for (int32_t base = 0; base < nq; base += kWave) {
        const int32_t i = base + static_cast<int32_t>(__lane_id()) + 1;
        const uint64_t v = i <= nq ? static_cast<uint64_t>(bt.costPrefix[i]) : ~0ULL;
        c0 += static_cast<uint32_t>(__popcll(__ballot(v <= lo)));
        c1 += static_cast<uint32_t>(__popcll(__ballot(v <= hi)));
        if (__ballot(v > hi))
            break;
    }
    t0 = PlanTask(bt.costPrefix, nq, blocks, lo, c0);
    t1 = PlanTask(bt.costPrefix, nq, blocks, hi, c1);
}

template <int B>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4))) void KScore(IndexView ix, BatchView bt, WorkView wk, int32_t blocks, int32_t addRows, int32_t keepCap, int32_t nqKnown)
{
    // addRows: the live kernel ran before this one in the same Search and wrote counter rows
    // [0, addRows); a wave below that adds its counters to its row. keepCap: BlockKeepCap's.
    // nqKnown: the batch's query count when the launch knows it (-1: read from the packet, as a
    // captured graph does), which saves a dependent load.
    constexpr int kVec = B / (4 * kWave);   // float4 accumulators per lane
    constexpr int kGpb = B / kGranule;      // granules per block
    constexpr int U = kColumns;
    constexpr int kSampleRounds = (kLongMin >> kSampleShift) / kWave;
    static_assert(kGpb <= kWave && kVec >= 1, "block size");
    static_assert(kHistBins == kWave * kHistPerLane, "every lane loads its histogram bins");
    __shared__ __align__(16) float acc[B];
    __shared__ uint32_t sHist[kHistBins]; // this wave's appends per histogram bucket
    // Doc IDs of 64 consecutive postings of each of the query's first kShortSlots short lists
    // (slot c: the c-th short word in query order), entry i at posting cbase[c] + i, INT32_MAX
    // past the list's end. A wave's tasks walk the blocks of a query in ascending order, so
    // a short list's postings in a block are found from the cached docs, without memory
    // traffic, until the cursor leaves the chunk.
    __shared__ int32_t sChunk[kShortSlots * kWave];

    const uint32_t lane = __lane_id();
    unsigned long long scanned = 0, useful = 0, scored = 0, skipped = 0, emitted = 0, overflow = 0, emitting = 0,
                       cuts = 0;
#ifdef BM25MI_PHASES
    // Shader-clock cycles per phase (descriptors, postings, selection, append), measurement builds only.
    unsigned long long phase[4] = {0, 0, 0, 0}, stamp = wall_clock64();
    const unsigned long long waveStart = stamp;
    int cur = 0;
#define BM25MI_PHASE(next)                                                                                             \
    do {                                                                                                               \
        const unsigned long long now_ = wall_clock64();                                                                     \
        phase[cur] += now_ - stamp;                                                                                    \
        stamp = now_;                                                                                                  \
        cur = (next);                                                                                                  \
    } while (0)
#else
#define BM25MI_PHASE(next)                                                                                             \
    do {                                                                                                               \
    } while (0)
#endif

    // This wave's tasks: a contiguous range of the query-major task order (task = q * blocks
    // + blk), cut by the host so that every wave gets the same estimated work. In batches of fewer
    // than 64 queries, lane i first loads query i's state, in flight with the plan's loads: the
    // wave's first query then needs no further round trip but its words.
    const int32_t nq = nqKnown >= 0 ? nqKnown : bt.header[0];
    const bool fewQueries = nq < kWave;
    int64_t preOff = 0, prePool = 0;
    float preThr = 0.0F;
    int32_t preK = 1;
    uint32_t preTau0 = 0;
    if (fewQueries && static_cast<int32_t>(lane) <= nq) {
        preOff = bt.qoff[lane];
        prePool = bt.poolOff[lane];
        if (static_cast<int32_t>(lane) < nq) {
            preThr = bt.thr[lane];
            preK = bt.k[lane];
            preTau0 = bt.tau0[lane];
        }
    }
    uint32_t tBegin, tEnd;
    WaveRange(bt, nq, blocks, tBegin, tEnd);
    // Per-query state, reloaded only when the task's query changes.
    int qCur = -1, n = 0;
    float thr = 0.0F;
    uint32_t k = 1, tau0q = 0;
    uint4 info = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
    // Long-list cursor (plain queries): a task's slice starts where the previous block's
    // ended, and the next block's directory entry is loaded one task ahead (lane j: word j).
    // The loaded entry is kept as loaded and only used by the next task, so that its load
    // stays in flight through this task's streaming.
    uint32_t hiPrev = 0, dirNext = 0;
    uint32_t cursorTask = 0xFFFFFFFFu; // the task hiPrev / dirNext were prepared for
    uint32_t cbaseV = 0;     // lane c: posting index of slot c's first cached entry
    bool chunksValid = false; // the slots hold chunks of the current query's lists
    // The query's pool (offset, capacity) and its score-histogram bound as last read.
    int64_t poolBase = 0, poolCap = 0;
    uint32_t histBound = 1, histAge = 0;
    // This wave's own appends to the query's pool, per histogram bucket (lane l: buckets
    // l*kHistPerLane ...): k of them at or above a bucket's edge bound the K-th score from below
    // as well. The bound is taken each time the count has doubled (a few times per query) and
    // published through tau, so that every wave of the query starts its selections from it.
    // pendHist holds the appends not yet added to the query's histogram in global memory: they are
    // added at the query's next task with postings, after that task's first round of posting
    // loads, and when the wave leaves the query, not right after the append (global atomics there
    // stall the wave's next posting loads; measured on the 10M-doc Q8 x 3-term rotation, 102 ->
    // 85 us without them).
    uint32_t ownHist[kHistPerLane], pendHist[kHistPerLane];
    bool flushDue = false;
    // With at most kEagerWaves waves per query of the batch, the atomics on a query's histogram
    // contend little, and each append is added at once.
    const bool eagerFlush = static_cast<int64_t>(gridDim.x) <= static_cast<int64_t>(kEagerWaves) * nq;
#pragma unroll
    for (int i = 0; i < kHistPerLane; ++i)
        ownHist[i] = pendHist[i] = 0;
    uint32_t ownCount = 0, ownAt = 0;
    auto flushHist = [&](int qq) {
        uint32_t *histw = wk.hist + static_cast<int64_t>(qq) * kHistBins;
#pragma unroll
        for (int r = 0; r < kHistPerLane; ++r) {
            if (pendHist[r])
                atomicAdd(histw + lane * kHistPerLane + r, pendHist[r]);
            pendHist[r] = 0;
        }
        flushDue = false;
    };
    // Candidates of a plain query go to this wave's own part of the query's pool, with no
    // atomic: the regions of the wave's blocks of the query, entries [segBlock * cap, (last block
    // + 1) * cap) (cap = BlockKeepCap(K) per block), filled contiguously from the first (ownCount
    // of them). The final kernel reads per-block counts: every plain task writes 0 for its block,
    // and when the wave leaves the query it writes the counts that cover its entries, region by
    // region (cap for each full region, then the rest), so that the final kernel's loads of a
    // wave's entries are contiguous.
    int32_t segBlock = 0;
    uint32_t capQ = 1;
    uint32_t floorSeen = 1; // the largest floor the wave used for the query
    auto setCount = [&](int qq, int bb, uint32_t v) {
        if (lane == 0)
            wk.blockCount[static_cast<int64_t>(qq) * BlockCountStride(blocks) + bb] = v;
    };
    // Before writing them, the wave drops the entries below its latest floor (kept while its
    // bounds were still low, e.g. at its first tasks): such a doc scores below a lower bound of the
    // query's K-th score, so it cannot be in the top K. The survivors are compacted in place
    // (entry n + j is written after entry n + j was read: n never exceeds the entries read).
    auto finishSegment = [&](int qq) {
        uint32_t n = 0;
        if (ownCount > 0) {
            // The floor: the largest the wave used, and its own histogram's bound.
            const uint32_t fl = max(floorSeen, histBound);
            unsigned long long *seg = wk.pool + poolBase + static_cast<int64_t>(segBlock) * capQ;
            // The entries written (the pool's capacity bounds them; it always holds them all).
            const uint32_t have = static_cast<uint32_t>(
                min(static_cast<int64_t>(ownCount), max(poolCap - static_cast<int64_t>(segBlock) * capQ, int64_t(0))));
            // The wave's lanes read each other's appends.
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            for (uint32_t i0 = 0; i0 < have; i0 += kWave) {
                const uint32_t i = i0 + lane;
                const unsigned long long key = i < have ? seg[i] : 0ULL;
                const bool keep = i < have && static_cast<uint32_t>(key >> 32) >= fl;
                const unsigned long long m = __ballot(keep);
                if (keep)
                    seg[n + static_cast<uint32_t>(__popcll(m & ((1ULL << lane) - 1ULL)))] = key;
                n += static_cast<uint32_t>(__popcll(m));
            }
        }
        if (lane == 0)
            for (uint32_t done = 0, bb = static_cast<uint32_t>(segBlock); done < n; done += capQ, ++bb)
                wk.blockCount[static_cast<int64_t>(qq) * BlockCountStride(blocks) + bb] = min(capQ, n - done);
    };
    // Enters query q at block blk (its state as loaded).
    auto enterQuery = [&](int q, int blk, int64_t w0, int64_t w1, float thrQ, int32_t kQ, uint32_t tau0Q, int64_t p0,
                          int64_t p1) {
        qCur = q;
        segBlock = blk;
        n = static_cast<int>(w1 - w0);
        thr = thrQ;
        k = static_cast<uint32_t>(kQ);
        tau0q = tau0Q;
        info = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
        if (static_cast<int>(lane) < n)
            info = bt.words[w0 + lane];
        chunksValid = false;
        poolBase = p0;
        poolCap = p1 - p0;
        histBound = 1;
        histAge = kHistEvery;
#pragma unroll
        for (int i = 0; i < kHistPerLane; ++i)
            ownHist[i] = 0;
        ownCount = ownAt = 0;
        capQ = static_cast<uint32_t>(BlockKeepCap(static_cast<int32_t>(k), B, keepCap));
        floorSeen = 1;
    };
    if (fewQueries && tBegin < tEnd) {
        // The first query's state from the lanes that loaded it (q + 1 <= nq < 64).
        const int q = static_cast<int>(tBegin / static_cast<uint32_t>(blocks));
        const auto lane64 = [](int64_t v, int i) {
            return static_cast<int64_t>((static_cast<uint64_t>(ReadLane(static_cast<uint32_t>(v >> 32), i)) << 32) |
                                        ReadLane(static_cast<uint32_t>(v), i));
        };
        enterQuery(q, static_cast<int>(tBegin - static_cast<uint32_t>(q) * static_cast<uint32_t>(blocks)),
                   lane64(preOff, q), lane64(preOff, q + 1), __uint_as_float(ReadLane(__float_as_uint(preThr), q)),
                   static_cast<int32_t>(ReadLane(static_cast<uint32_t>(preK), q)), ReadLane(preTau0, q),
                   lane64(prePool, q), lane64(prePool, q + 1));
    }
    for (uint32_t t = tBegin; t < tEnd; ++t) {
        BM25MI_PHASE(0);
        const int q = static_cast<int>(t / static_cast<uint32_t>(blocks));
        const int blk = static_cast<int>(t - static_cast<uint32_t>(q) * static_cast<uint32_t>(blocks));
        const int32_t base = blk * B;
        const int32_t end = min(base + B, ix.ndocs);
        if (q != qCur) {
            if (qCur >= 0) {
                flushHist(qCur);
                finishSegment(qCur);
            }
            enterQuery(q, blk, bt.qoff[q], bt.qoff[q + 1], bt.thr[q], bt.k[q], bt.tau0[q], bt.poolOff[q],
                       bt.poolOff[q + 1]);
        }
        if (thr > 0.0F) {
            // A threshold query's live blocks are scored by KLive.
            t = max(t, min(tEnd, static_cast<uint32_t>(q + 1) * static_cast<uint32_t>(blocks)) - 1);
            continue;
        }

        // Lower bounds of the query's K-th score, needed only by the selection: the score
        // histogram, read at the query's first task here and then every kHistEvery-th task, and
        // tau, every task; both loaded after the task's last round of posting loads, so that no
        // posting load waits for them (a load's completion is counted in order with the loads
        // issued before it, and the appends' atomics make these lines slow to read). These bounds
        // only grow, so a value read early is still a lower bound.
        const bool readHist = histAge >= (k > kHistLargeK ? kHistEveryLargeK : kHistEvery);
        uint32_t hregs[kHistPerLane];
#pragma unroll
        for (int i = 0; i < kHistPerLane; ++i)
            hregs[i] = 0;
        if (readHist)
            histAge = 0;
        ++histAge;
        // The query's tau (device scope: the raises the other waves published, not a cached copy).
        uint32_t tauLoaded = 1;

        const int64_t g0 = static_cast<int64_t>(blk) * kGpb;
        const int gpb = static_cast<int>(min(static_cast<int64_t>(kGpb), ix.granules - g0));
        const int32_t sl = static_cast<int32_t>(info.z);
        const bool isLong = info.y > 0 && sl >= 0;
        const uint32_t *dirRow = ix.dir + static_cast<int64_t>(sl) * (ix.granules + 1);
        float4 *acc4 = reinterpret_cast<float4 *>(acc);

        // Window of short list j (a doc-ascending run of < kLongMin postings) covering docs
        // [lo, hi): postings between the samples that bracket lo and hi. Executed by the whole
        // wave; lane j receives it.
        auto shortWindow = [&](int j, int32_t wlo, int32_t whi, uint32_t &oLo, uint32_t &oLen) {
            const uint32_t first = ReadLane(info.x, j), cnt = ReadLane(info.y, j);
            const uint32_t i0 = (first + 15u) >> kSampleShift, i1 = (first + cnt + 15u) >> kSampleShift;
            // Every lane loads (indices clamped into the list's samples, i1 > i0 - 1 >= 0 since
            // the list is not empty), so that all loads are in flight before the first use.
            int32_t smp[kSampleRounds];
#pragma unroll
            for (int r = 0; r < kSampleRounds; ++r)
                smp[r] = ix.samples[min(i0 + r * kWave + lane, i1 - 1u)];
            uint32_t ca = 0, cb = 0;
#pragma unroll
            for (int r = 0; r < kSampleRounds; ++r) {
                const bool in = i0 + r * kWave + lane < i1;
                ca += in && smp[r] < wlo ? 1u : 0u;
                cb += in && smp[r] < whi ? 1u : 0u;
            }
            ca = WaveSum(ca);
            cb = WaveSum(cb);
            const uint32_t ws = ca > 0 ? ((i0 + ca - 1) << kSampleShift) + 1 : first;
            const uint32_t we = cb < i1 - i0 ? (i0 + cb) << kSampleShift : first + cnt;
            if (static_cast<int>(lane) == j) {
                oLo = ws;
                oLen = we > ws ? we - ws : 0u;
            }
        };

        // Short words in query order; word j's cache slot is its rank among them.
        const unsigned long long shorts = __ballot(static_cast<int>(lane) < n && info.y > 0 && sl < 0);
        // Loads every slot's chunk so that it holds the list's first posting with doc >= at
        // (or the list's end): the samples locate the last sample below `at`, and the chunk
        // starts there. All sample loads are issued before any is used, then all chunk loads:
        // two memory round trips for up to kShortSlots lists. The loads are unconditional
        // (slots without a word load any in-range entry and are never read), so that the
        // compiler keeps them in flight together.
        auto establish = [&](int32_t at) {
            int32_t sv[kShortSlots][kSampleRounds];
            unsigned long long m = shorts;
#pragma unroll
            for (int c = 0; c < kShortSlots; ++c) {
                const int j = m ? __ffsll(static_cast<long long>(m)) - 1 : 0;
                m &= m - 1;
                const uint32_t first = ReadLane(info.x, j), cnt = ReadLane(info.y, j);
                const uint32_t i0 = (first + 15u) >> kSampleShift, i1 = (first + cnt + 15u) >> kSampleShift;
#pragma unroll
                for (int r = 0; r < kSampleRounds; ++r) {
                    const uint32_t i = i0 + r * kWave + lane;
                    const int32_t v = ix.samples[min(i, max(i1, 1u) - 1u)];
                    sv[c][r] = i < i1 ? v : INT32_MAX;
                }
            }
            int32_t dv[kShortSlots];
            m = shorts;
#pragma unroll
            for (int c = 0; c < kShortSlots; ++c) {
                const int j = m ? __ffsll(static_cast<long long>(m)) - 1 : 0;
                m &= m - 1;
                const uint32_t first = ReadLane(info.x, j), cnt = ReadLane(info.y, j);
                const uint32_t i0 = (first + 15u) >> kSampleShift;
                uint32_t ca = 0;
#pragma unroll
                for (int r = 0; r < kSampleRounds; ++r)
                    ca += sv[c][r] < at ? 1u : 0u;
                ca = WaveSum(ca);
                const uint32_t start = ca > 0 ? (i0 + ca - 1) << kSampleShift : first;
                const uint32_t idx = start + lane;
                const int32_t v = static_cast<int32_t>(ix.post[min(idx, max(first + cnt, 1u) - 1u)].x);
                dv[c] = idx < first + cnt ? v : INT32_MAX;
                cbaseV = static_cast<int>(lane) == c ? start : cbaseV;
            }
#pragma unroll
            for (int c = 0; c < kShortSlots; ++c)
                sChunk[c * kWave + lane] = dv[c];
        };
        // Posting range of short word j, cached in slot c, for docs [wlo, whi): exact from the
        // chunk (the chunk is reloaded when it lies wholly below wlo); a block that reaches past
        // the chunk's last entry takes the sample window instead. Lane j receives it.
        auto shortRange = [&](int j, int32_t wlo, int32_t whi, uint32_t &oLo, uint32_t &oLen) {
            const int c = __popcll(shorts & ((1ULL << j) - 1));
            if (c >= kShortSlots) {
                shortWindow(j, wlo, whi, oLo, oLen);
                return;
            }
            const uint32_t listEnd = ReadLane(info.x, j) + ReadLane(info.y, j);
            if (__popcll(__ballot(sChunk[c * kWave + lane] < wlo)) == kWave && ReadLane(cbaseV, c) + kWave < listEnd)
                establish(wlo); // the chunk lies wholly below wlo
            const int32_t d = sChunk[c * kWave + lane];
            const uint32_t cb = ReadLane(cbaseV, c);
            const uint32_t nl = static_cast<uint32_t>(__popcll(__ballot(d < wlo)));
            const uint32_t nh = static_cast<uint32_t>(__popcll(__ballot(d < whi)));
            if (nh == kWave && cb + kWave < listEnd) {
                shortWindow(j, wlo, whi, oLo, oLen);
                return;
            }
            if (static_cast<int>(lane) == j) {
                oLo = cb + nl;
                oLen = nh - nl;
            }
        };

        // Zeroes the accumulator row and adds the postings of segments [segLo, segLo+segLen)
        // (lane s: segment s, s < nseg, in query-word order) for docs in [base, end), in
        // rounds of U 64-posting columns. Returns the postings streamed.
        uint32_t inRange = 0;
        // Largest value this lane wrote to the accumulator row: impacts are non-negative, so every
        // partial sum is at most the doc's final score, and the wave maximum is the block's top score.
        float runMax = 0.0F;
        auto streamAdd = [&](uint32_t segLo, uint32_t segLen, int nseg) {
            const uint32_t myCols = static_cast<int>(lane) < nseg ? (segLen + kWave - 1) / kWave : 0u;
            uint32_t colEnd = myCols;
            colEnd = WaveInclusiveSum(colEnd);
            const uint32_t columns = ReadLane(colEnd, kWave - 1);
            const uint32_t colStart = colEnd - myCols;
            // Lanes past the last segment hold ~0: column c belongs to segment
            // popcount(ballot(colEnd <= c)).
            if (static_cast<int>(lane) >= nseg)
                colEnd = 0xFFFFFFFFu;
            if (columns == 0) {
#pragma unroll
                for (int r = 0; r < kVec; ++r)
                    acc4[r * kWave + lane] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
            }
            const uint32_t span = static_cast<uint32_t>(end - base);
            for (uint32_t r0 = 0; r0 < columns; r0 += U) {
                // Lanes past a column's end keep doc 0xFFFFFFFF, which the range test drops.
                uint2 e[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    e[u] = make_uint2(0xFFFFFFFFu, 0u);
                    const uint32_t c = r0 + u;
                    if (c < columns) {
                        const int sg = __popcll(__ballot(colEnd <= c));
                        const uint32_t off = (c - ReadLane(colStart, sg)) << 6;
                        if (lane + off < ReadLane(segLen, sg))
                            e[u] = ix.post[ReadLane(segLo, sg) + off + lane];
                    }
                }
                if (r0 == 0 && flushDue) {
                    // The previous tasks' appends, added to the query's histogram after this
                    // round's posting loads: waits for those loads do not wait for these atomics.
                    __builtin_amdgcn_sched_barrier(0);
                    flushHist(q);
                    flushDue = false;
                }
                if (r0 + U >= columns) { // the task's last round: nothing waits for these loads before
                                         // its selection
                    __builtin_amdgcn_sched_barrier(0);
                    tauLoaded = __hip_atomic_load(wk.tau + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (readHist) {
#pragma unroll
                        for (int i = 0; i < kHistPerLane; ++i)
                            hregs[i] = __hip_atomic_load(wk.hist + static_cast<int64_t>(q) * kHistBins +
                                                             lane * kHistPerLane + i,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (r0 == 0) {
#pragma unroll
                    for (int r = 0; r < kVec; ++r)
                        acc4[r * kWave + lane] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
                }
                // Read-add-write each column in order. When the whole round lies in one
                // segment, its docs are distinct, so all reads may go first (no column can
                // see another's write).
                const uint32_t cLast = min(r0 + U, columns) - 1;
                if (__popcll(__ballot(colEnd <= r0)) == __popcll(__ballot(colEnd <= cLast))) {
                    float v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const uint32_t rel = e[u].x - static_cast<uint32_t>(base);
                        v[u] = rel < span ? acc[rel] : 0.0F;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const uint32_t rel = e[u].x - static_cast<uint32_t>(base);
                        if (rel < span) {
                            const float nv = v[u] + __uint_as_float(e[u].y);
                            acc[rel] = nv;
                            runMax = fmaxf(runMax, nv);
                            ++inRange;
                        }
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const uint32_t rel = e[u].x - static_cast<uint32_t>(base);
                        if (rel < span) {
                            const float nv = acc[rel] + __uint_as_float(e[u].y);
                            acc[rel] = nv;
                            runMax = fmaxf(runMax, nv);
                            ++inRange;
                        }
                    }
                }
            }
            uint32_t streamed = 0;
#pragma unroll 1
            for (int sg = 0; sg < nseg; ++sg)
                streamed += ReadLane(segLen, sg);
            return streamed;
        };

        uint32_t lo = 0, len = 0;
        // 1. Plain task: the posting range of each word in the block (lane j: word j). Short
        // lists from the cached chunks, long lists from the cursor (this block starts where the
        // previous one ended; the next block's end is loaded one task ahead) or the directory.
        if (shorts && !chunksValid) {
            establish(base); // the chunks of the query's cached short lists, at its first task here
            chunksValid = true;
        }
        unsigned long long shortMask = __ballot(static_cast<int>(lane) < n && info.y > 0 && sl < 0);
        while (shortMask) {
            const int j = __ffsll(static_cast<long long>(shortMask)) - 1;
            shortMask &= shortMask - 1;
            shortRange(j, base, end, lo, len);
        }
        if (isLong) {
            if (cursorTask == t) {
                lo = hiPrev;
                len = info.x + dirNext - lo;
            } else {
                lo = info.x + dirRow[g0];
                len = info.x + dirRow[g0 + gpb] - lo;
            }
        }
        hiPrev = lo + len;
        cursorTask = 0xFFFFFFFFu;
        if (t + 1 < tEnd && blk + 1 < blocks) {
            cursorTask = t + 1;
            if (isLong)
                dirNext = dirRow[min(g0 + 2 * kGpb, ix.granules)];
        }
        if (WaveSum(len) == 0) {
            ++skipped; // no query word has postings here: every score is 0
            setCount(q, blk, 0u);
            continue;
        }
        // 2. Stream and add, word by word in query order.
        BM25MI_PHASE(1);
        scanned += streamAdd(lo, len, n);
        ++scored;
        useful += inRange;

        // 3. Selection. A block whose top score is below the floor appends nothing: the floor's
        // cheap part (tau, the start bound) is checked first, then the histogram bound, before
        // the accumulator row is scanned.
        BM25MI_PHASE(2);
        const uint32_t top = WaveMax(__float_as_uint(runMax));
        if (readHist)
            histBound = max(histBound, WaveHistBoundRegs(hregs, k));
        const uint32_t floor = max(max(tauLoaded, tau0q), histBound);
        floorSeen = max(floorSeen, floor);
        if (top < floor) {
            setCount(q, blk, 0u);
            continue;
        }
        const uint4 *bits4 = reinterpret_cast<const uint4 *>(acc);
        uint32_t mine = 0, mx = 0;
#pragma unroll 2
        for (int r = 0; r < kVec; ++r) {
            const uint4 v = bits4[r * kWave + lane];
            mine += (v.x >= floor ? 1u : 0u) + (v.y >= floor ? 1u : 0u) + (v.z >= floor ? 1u : 0u) +
                    (v.w >= floor ? 1u : 0u);
            mx = max(mx, max(max(v.x, v.y), max(v.z, v.w)));
        }
        uint32_t kept = WaveSum(mine);
        if (kept == 0) {
            setCount(q, blk, 0u);
            continue;
        }
        BM25MI_PHASE(3);
        // Cut (rare), then append the kept docs to the task's region of the pool as composite keys
        // (score bits << 32 | ~doc) and count them in this wave's histogram; the cut raises tau.
        LdsU32 *bitsL = (LdsU32 *)bits4; // acc is in LDS
        uint32_t tb = floor, te = 0, tauNew = 0;
        const uint32_t cap = static_cast<uint32_t>(BlockKeepCap(static_cast<int32_t>(k), B, keepCap));
        if (kept > cap) {
            const CutResult cr = SelectCut<B>(bitsL, floor, k, cap, mx);
            tb = cr.tb;
            te = cr.te;
            kept = cr.kept;
            mine = cr.mine;
            tauNew = cr.tauNew;
            histBound = max(histBound, tauNew); // the block holds k docs at or above it
            ++cuts;
        }
        // The wave's entries fit its regions (at most cap per task) and the pool: they end at most at
        // (last block) * cap + (its docs) <= ndocs, or (last block + 1) * cap <= blocks * cap.
        const uint32_t slot0 = static_cast<uint32_t>(segBlock) * cap + ownCount;
        setCount(q, blk, 0u);
        uint32_t before = mine;
        before = WaveInclusiveSum(before);
        before -= mine;
        // Lanes count into each other's bins, so the bins are written and read as volatile: the
        // compiler must not forward a lane's own zero to its read below, which would drop the
        // other lanes' counts (LDS operations of a wave are performed in order).
        volatile __attribute__((address_space(3))) uint32_t *bins =
            (volatile __attribute__((address_space(3))) uint32_t *)sHist;
#pragma unroll
        for (int r = 0; r < kHistPerLane; ++r)
            bins[lane * kHistPerLane + r] = 0;
        unsigned long long *pool = wk.pool + poolBase;
        const int64_t capq = poolCap;
        if (mine) {
            uint32_t slot = slot0 + before;
#pragma unroll 1
            for (int r = 0; r < kVec; ++r) {
                const U32x4 v = reinterpret_cast<LdsUint4 *>(bitsL)[r * kWave + lane];
                const uint32_t vals[4] = {v.x, v.y, v.z, v.w};
                const uint32_t e0 = (r * kWave + lane) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const uint32_t x = vals[c], rk = static_cast<uint32_t>(B - 1) - (e0 + c);
                    if (x > tb || (x == tb && rk >= te)) {
                        if (static_cast<int64_t>(slot) < capq)
                            pool[slot] = (static_cast<unsigned long long>(x) << 32) |
                                         static_cast<unsigned long long>(0xFFFFFFFFu - (static_cast<uint32_t>(base) + e0 + c));
                        else
                            ++overflow;
                        ++slot;
                        atomicAdd(sHist + HistBucket(x), 1u);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kHistPerLane; ++r) {
            const uint32_t c = bins[lane * kHistPerLane + r];
            ownHist[r] += c;
            pendHist[r] += c;
        }
        ownCount += kept;
        if (eagerFlush)
            flushHist(q);
        if (ownCount >= k && ownCount >= 2 * ownAt) {
            ownAt = ownCount;
            flushDue = true; // added at the query's next task with postings (or when the wave leaves it)
            const uint32_t ob = WaveHistBoundRegs(ownHist, k);
            if (ob > histBound && ob > tauLoaded)
                tauNew = max(tauNew, ob);
            histBound = max(histBound, ob);
        }
        if (lane == 0 && tauNew > tauLoaded)
            atomicMax(wk.tau + q, tauNew);
        emitted += kept;
        ++emitting;
    }
    if (qCur >= 0) {
        flushHist(qCur);
        finishSegment(qCur);
    }
    BM25MI_PHASE(0);
#undef BM25MI_PHASE
    // Counters go to this wave's own row (plain stores; added to the row the live kernel's wave
    // of the same index wrote when addRows); the final kernel sums the rows.
    useful = WaveSum64(useful);
    overflow = WaveSum64(overflow);
    if (lane == 0) {
        unsigned long long *row = wk.stats + static_cast<int64_t>(blockIdx.x) * kStats;
        const bool covered = static_cast<int32_t>(blockIdx.x) < addRows;
        if (covered && tBegin == tEnd)
            return; // no tasks: the live kernel's row already holds this Search's counters
        if (covered) {
            scanned += row[kStatPostings];
            scored += row[kStatScored];
            skipped += row[kStatSkipped];
            emitted += row[kStatCandidates];
            overflow += row[kStatOverflow];
            emitting += row[kStatEmitting];
            cuts += row[kStatExact];
            useful += row[kStatUseful];
        }
        row[kStatPostings] = scanned;
        row[kStatScored] = scored;
        row[kStatSkipped] = skipped;
        row[kStatCandidates] = emitted;
        row[kStatOverflow] = overflow;
        row[kStatEmitting] = emitting;
        row[kStatExact] = cuts;
        row[kStatUseful] = useful;
#ifdef BM25MI_PHASES
        for (int i = 0; i < 4; ++i)
            row[kStatPhase + i] = phase[i];
        row[kStatMax] = phase[0] + phase[1] + phase[2] + phase[3]; // busy time
        row[kStatMax + 1] = stamp;                                  // end
        row[kStatMax + 2] = ~waveStart;                             // start (max of ~ = min)
        row[kStatMax + 3] = waveStart;                              // latest start
#else
        for (int i = kStatPhase; i < kStats; ++i)
            row[i] = 0;
#endif
    }
}

// Threshold of query q of a learned batch (no reference counterpart), computed by one wave
// (every lane returns it): the MLP kFeatures -> 64 -> 64 -> 1 (ReLU, FP32) maps the query's
// standardised features to z, an estimate of (log(K-th score / upper) - yMean) / yScale, where
// upper is the FP32 sum of its terms' largest impacts; the threshold is
// clamp(safety * upper * exp(z * yScale + yMean), lower, upper), lower = tau0 (the K-th largest
// impact of any of its terms). A threshold below minRatio * upper, or K above liveMaxK, routes
// the query to the plain path (0). Lane j computes hidden unit j of both layers. Load() issues
// every load (the query's inputs and the lane's rows of both layers, one round trip) and
// Compute() waits for them, so that a caller can overlap them with its own loads.
struct ThresholdPredictor {
    float ub, lower, b1, b2, w3, b3;
    int32_t k;
    float4 f[kFeatures / 4], a[kFeatures / 4], w[kHidden / 4];
    float mu[kFeatures], sd[kFeatures]; // mean and scale follow b3 and are only 4-byte aligned

    __device__ __forceinline__ void Load(const ModelView &mv, const BatchView &bt, int q)
    {
        const int lane = static_cast<int>(__lane_id());
        const float *W1 = mv.model, *B1 = W1 + kHidden * kFeatures, *W2 = B1 + kHidden, *B2 = W2 + kHidden * kHidden;
        const float *W3 = B2 + kHidden, *B3 = W3 + kHidden, *mean = B3 + 1, *scale = mean + kFeatures;
        ub = bt.upper[q];
        lower = __uint_as_float(bt.tau0[q]);
        k = bt.k[q];
        const float4 *f4 = reinterpret_cast<const float4 *>(bt.feat + static_cast<int64_t>(q) * kFeatures);
        const float4 *w1 = reinterpret_cast<const float4 *>(W1) + lane; // device layout: see ModelView
        const float4 *w2 = reinterpret_cast<const float4 *>(W2) + lane;
#pragma unroll
        for (int v = 0; v < kFeatures / 4; ++v) {
            f[v] = f4[v];
            a[v] = w1[v * kHidden];
        }
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) {
            mu[i] = mean[i];
            sd[i] = scale[i];
        }
#pragma unroll
        for (int v = 0; v < kHidden / 4; ++v)
            w[v] = w2[v * kHidden];
        b1 = B1[lane];
        b2 = B2[lane];
        w3 = W3[lane];
        b3 = B3[0];
    }

    __device__ __forceinline__ float Compute(const ModelView &mv) const
    {
        float z1 = b1;
#pragma unroll
        for (int v = 0; v < kFeatures / 4; ++v) {
            z1 = fmaf(a[v].x, (f[v].x - mu[4 * v]) / sd[4 * v], z1);
            z1 = fmaf(a[v].y, (f[v].y - mu[4 * v + 1]) / sd[4 * v + 1], z1);
            z1 = fmaf(a[v].z, (f[v].z - mu[4 * v + 2]) / sd[4 * v + 2], z1);
            z1 = fmaf(a[v].w, (f[v].w - mu[4 * v + 3]) / sd[4 * v + 3], z1);
        }
        // Hidden unit i of the first layer, broadcast from lane i (a scalar read per unit).
        const uint32_t h1 = __float_as_uint(fmaxf(z1, 0.0F));
        const auto H = [&](int i) { return __uint_as_float(ReadLane(h1, i)); };
        // Four independent partial sums (a shorter dependent chain), then the wave sum.
        float z2a = b2, z2b = 0.0F, z2c = 0.0F, z2d = 0.0F;
#pragma unroll
        for (int v = 0; v < kHidden / 4; ++v) {
            z2a = fmaf(w[v].x, H(4 * v), z2a);
            z2b = fmaf(w[v].y, H(4 * v + 1), z2b);
            z2c = fmaf(w[v].z, H(4 * v + 2), z2c);
            z2d = fmaf(w[v].w, H(4 * v + 3), z2d);
        }
        const float z2 = (z2a + z2b) + (z2c + z2d);
        const float y = __ockl_wfred_add_f32(w3 * fmaxf(z2, 0.0F)) + b3;
        float thr = 0.0F;
        if (ub > 0.0F && k <= mv.liveMaxK) {
            thr = fminf(fmaxf(mv.safety * ub * expf(y * mv.yScale + mv.yMean), lower), ub);
            if (!(thr >= mv.minRatio * ub))
                thr = 0.0F;
        }
        return thr;
    }
};

__device__ __forceinline__ unsigned long long WaveOr64(unsigned long long v)
{
    return __ockl_wfred_or_u64(v);
}

constexpr int kPruneSpan = 64; // granules per planning item (one per lane)
constexpr int kPruneGroup = 4; // planning waves per workgroup

// Planning kernel of threshold queries (the learned path; no reference counterpart). Wave b of
// the first `waves` takes the items [items * b / waves, items * (b + 1) / waves) of the
// query-major item order (item q * ceil(granules / 64) + c: granules 64c .. 64c + 63 of query q),
// kPruneItems of one query at a time; lane g sums, over the query's words, an upper bound of
// each word's impact in granule g of each item: the granule maximum of a long list (all of the
// items' loads of up to eight long words in flight together), the term maximum of a short list
// when one of its postings lies in the granule (found through the posting samples, then the
// postings themselves). FP32 addition of non-negative values is monotone, so the FP32 score of
// every doc in the granule, summed in query order, is at most the same sum of these bounds; the
// sum here may be taken in another order, which the factor (1 - 2^-16) on the threshold covers
// for up to 64 words. A block with no granule reaching the threshold cannot hold a doc scoring at
// least it. The kernel writes every item's mask of live granules (0 for the items of plain
// queries) and, per wave, the live granules and dead blocks it found; the live kernel cuts its
// waves' shares from the live counts. For learned batches (header[3] = 1), whose routing the
// staging kernel decided, the extra wave `waves` writes the plain queries' cost prefix and the
// header counts. nqKnown: the batch's query count when the launch knows it (-1: read it from
// the packet, as a captured graph does), which saves a dependent load.
template <int B>
__global__ __launch_bounds__(kPruneGroup * kWave) void KPrune(IndexView ix, BatchView bt, WorkView wk, int32_t blocks,
                                                              int32_t waves, int32_t nqKnown)
{
    // kPruneGroup independent waves per workgroup (fewer workgroups to dispatch); wave w.
    const int32_t w = static_cast<int32_t>(blockIdx.x) * kPruneGroup + static_cast<int32_t>(threadIdx.x >> 6);
    if (w > waves)
        return;
    constexpr int kGpb = B / kGranule;
    constexpr int kBpc = kPruneSpan / kGpb; // blocks per item
    constexpr int C = kPruneItems;
    constexpr int kSampleRounds = (kLongMin >> kSampleShift) / kWave;
    static_assert(kPruneSpan % kGpb == 0 && kGpb <= kPruneSpan, "block size");
    static_assert(C * kBpc <= kWave, "one lane per block of a chunk");
    const uint32_t lane = __lane_id();
    const int32_t nq = nqKnown >= 0 ? nqKnown : bt.header[0];
    if (w == waves) {
        if (bt.header[3] == 0)
            return;
        // Learned batch: the wave plan of the plain kernel over the queries routed to the plain path.
        auto *cost = const_cast<int64_t *>(bt.costPrefix);
        auto *header = const_cast<int32_t *>(bt.header);
        int64_t acc = 0;
        int32_t live = 0;
        for (int base = 0; base < nq; base += kWave) {
            const int q = base + static_cast<int>(lane);
            const bool in = q < nq;
            const bool isLive = in && bt.thr[q] > 0.0F;
            int64_t c = in && !isLive ? bt.planCost[q] : 0;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int64_t y = __shfl_up(c, o);
                if (lane >= static_cast<uint32_t>(o))
                    c += y;
            }
            if (in)
                cost[q + 1] = acc + c;
            acc += __shfl(c, kWave - 1);
            live += __popcll(__ballot(isLive));
        }
        if (lane == 0) {
            cost[0] = 0;
            header[1] = live;
            header[2] = nq - live;
        }
        return;
    }
    const int64_t granules = ix.granules;
    const uint32_t cpq = static_cast<uint32_t>((granules + kPruneSpan - 1) / kPruneSpan);
    const uint32_t items = static_cast<uint32_t>(nq) * cpq;
    const uint32_t i0 = static_cast<uint32_t>(static_cast<uint64_t>(items) * static_cast<uint32_t>(w) / waves);
    const uint32_t i1 = static_cast<uint32_t>(static_cast<uint64_t>(items) * static_cast<uint32_t>(w + 1) / waves);
    int qCur = -1, n = 0;
    int64_t wq0 = 0; // the query's first word in the batch
    float thrLive = 0.0F;
    uint4 info = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
    uint32_t dead = 0, live = 0;
    for (uint32_t it = i0; it < i1;) {
        const int q = static_cast<int>(it / cpq);
        const uint32_t qEnd = min(i1, static_cast<uint32_t>(q + 1) * cpq); // this range's items of query q
        if (q != qCur) {
            qCur = q;
            // Independent loads, issued together.
            const float thr = bt.thr[q];
            const int64_t w0 = bt.qoff[q], w1 = bt.qoff[q + 1];
            thrLive = thr * 0.99998474F; // thr * (1 - 2^-16)
            n = 0;
            info = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
            if (thr > 0.0F) {
                wq0 = w0;
                n = static_cast<int>(w1 - w0);
                if (static_cast<int>(lane) < n)
                    info = bt.words[w0 + lane];
            }
        }
        if (!(thrLive > 0.0F)) {
            for (uint32_t j = it + lane; j < qEnd; j += kWave)
                wk.itemMask[j] = 0ULL; // a plain query
            it = qEnd;
            continue;
        }
        const uint32_t nIt = min(static_cast<uint32_t>(C), qEnd - it);
        const int64_t gBase = static_cast<int64_t>(it - static_cast<uint32_t>(q) * cpq) * kPruneSpan;
        bool valid[C];
        float ub[C];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            valid[u] = static_cast<uint32_t>(u) < nIt && gBase + u * kPruneSpan + lane < granules;
            ub[u] = 0.0F;
        }
        // Long words: the items' granule-maximum rows of eight words in flight at a time; the
        // first eight stay in flight through the short words' searches and are added after them.
        unsigned long long longs = __ballot(static_cast<int>(lane) < n && info.y > 0 && static_cast<int32_t>(info.z) >= 0);
        float vFirst[8][C];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            int64_t row = -1;
            if (longs) {
                const int j = __ffsll(static_cast<long long>(longs)) - 1;
                longs &= longs - 1;
                row = static_cast<int64_t>(static_cast<int32_t>(ReadLane(info.z, j))) * granules + gBase + lane;
            }
#pragma unroll
            for (int u = 0; u < C; ++u)
                vFirst[t][u] = row >= 0 && valid[u] ? ix.bmax[row + u * kPruneSpan] : 0.0F;
        }
        while (longs) {
            float v[8][C];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                int64_t row = -1;
                if (longs) {
                    const int j = __ffsll(static_cast<long long>(longs)) - 1;
                    longs &= longs - 1;
                    row = static_cast<int64_t>(static_cast<int32_t>(ReadLane(info.z, j))) * granules + gBase + lane;
                }
#pragma unroll
                for (int u = 0; u < C; ++u)
                    v[t][u] = row >= 0 && valid[u] ? ix.bmax[row + u * kPruneSpan] : 0.0F;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int u = 0; u < C; ++u)
                    ub[u] += v[t][u];
        }
        // Short words: the postings between the samples that bracket the chunk's docs.
        unsigned long long shorts = __ballot(static_cast<int>(lane) < n && info.y > 0 && static_cast<int32_t>(info.z) < 0);
        if (shorts) {
            const int32_t dlo = static_cast<int32_t>(gBase << kGranuleShift);
            const int32_t dhi = static_cast<int32_t>(
                min((gBase + static_cast<int64_t>(nIt) * kPruneSpan) << kGranuleShift, static_cast<int64_t>(ix.ndocs)));
            const uint32_t chunkBlocks = nIt * kBpc;
            while (shorts) {
                const int j = __ffsll(static_cast<long long>(shorts)) - 1;
                shorts &= shorts - 1;
                const uint32_t first = ReadLane(info.x, j), cnt = ReadLane(info.y, j);
                const uint32_t s0 = (first + 15u) >> kSampleShift, s1 = (first + cnt + 15u) >> kSampleShift;
                int32_t smp[kSampleRounds];
#pragma unroll
                for (int r = 0; r < kSampleRounds; ++r)
                    smp[r] = ix.samples[min(s0 + r * kWave + lane, s1 - 1u)];
                uint32_t ca = 0, cb = 0;
#pragma unroll
                for (int r = 0; r < kSampleRounds; ++r) {
                    const bool in = s0 + r * kWave + lane < s1;
                    ca += in && smp[r] < dlo ? 1u : 0u;
                    cb += in && smp[r] < dhi ? 1u : 0u;
                }
                ca = WaveSum(ca);
                cb = WaveSum(cb);
                const uint32_t ws = ca > 0 ? ((s0 + ca - 1) << kSampleShift) + 1 : first;
                const uint32_t we = cb < s1 - s0 ? (s0 + cb) << kSampleShift : first + cnt;
                unsigned long long m[C];
#pragma unroll
                for (int u = 0; u < C; ++u)
                    m[u] = 0;
                uint32_t below = 0; // lane b < chunkBlocks: postings of [ws, we) before block b of the chunk
                for (uint32_t p0 = ws; p0 < we; p0 += kWave) {
                    const bool v = p0 + lane < we;
                    const uint32_t d = v ? ix.post[p0 + lane].x : 0xFFFFFFFFu;
                    const uint32_t rel = (d >> kGranuleShift) - static_cast<uint32_t>(gBase);
#pragma unroll
                    for (int u = 0; u < C; ++u)
                        if (v && rel - static_cast<uint32_t>(u * kPruneSpan) < static_cast<uint32_t>(kPruneSpan))
                            m[u] |= 1ULL << (rel - static_cast<uint32_t>(u * kPruneSpan));
#pragma unroll
                    for (int b = 0; b < C * kBpc; ++b) {
                        const uint32_t c = static_cast<uint32_t>(
                            __popcll(__ballot(d < static_cast<uint32_t>(dlo) + static_cast<uint32_t>(b * B))));
                        below += static_cast<int>(lane) == b ? c : 0u;
                    }
                }
                const float mx = __uint_as_float(ReadLane(info.w, j));
#pragma unroll
                for (int u = 0; u < C; ++u)
                    if ((WaveOr64(m[u]) >> lane) & 1ULL)
                        ub[u] += mx;
                // Postings before ws lie below the chunk's first doc and those from we on at or
                // above its end, so ws + below is the word's first posting in block b: the live
                // kernel starts its search of the block's granules there.
                if (lane < chunkBlocks && gBase / kGpb + lane < blocks)
                    wk.cursor[(wq0 + j) * blocks + gBase / kGpb + lane] = ws + below;
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
#pragma unroll
            for (int t = 0; t < 8; ++t)
                ub[u] += vFirst[t][u];
            if (static_cast<uint32_t>(u) < nIt) {
                const unsigned long long liveG = __ballot(valid[u] && ub[u] >= thrLive);
                unsigned long long bm = 0; // lane b < kBpc: the live granules of block b of the item
                if (static_cast<int>(lane) < kBpc) {
                    if constexpr (kGpb == kPruneSpan)
                        bm = liveG;
                    else
                        bm = (liveG >> (lane * kGpb)) & ((1ULL << kGpb) - 1ULL);
                }
                const bool inRange = static_cast<int>(lane) < kBpc && gBase / kGpb + u * kBpc + lane < blocks;
                if (lane == 0)
                    wk.itemMask[it + u] = liveG;
                live += static_cast<uint32_t>(__popcll(liveG));
                dead += static_cast<uint32_t>(__popcll(__ballot(inRange && bm == 0)));
            }
        }
        it += nIt;
    }
    if (lane == 0) {
        wk.dead[w] = dead;
        wk.live[w] = live;
    }
}

#ifndef BM25MI_LIVE_SLOTS
#define BM25MI_LIVE_SLOTS 8
#endif
// Granules a live batch scores together (8 x 128 FP32 = 4 KB of LDS per wave; on the 10M-doc
// Q8 x 3-term K10 batches 8 measured 0.077 / 0.058 ms of device time per learned Search against
// 0.079 / 0.063 for 16 and 0.113 / 0.076 for 32).
constexpr int kLiveSlots = BM25MI_LIVE_SLOTS;
static_assert(kLiveSlots >= 2 && kLiveSlots <= 32 && (kLiveSlots & (kLiveSlots - 1)) == 0, "live batch size");

// Index (0..63) of the n-th set bit (n counted from 0) of m, which has more than n set bits.
__device__ __forceinline__ uint32_t NthSetBit(unsigned long long m, uint32_t n)
{
    uint32_t pos = 0;
#pragma unroll
    for (int width = 32; width > 0; width >>= 1) {
        const uint32_t low = static_cast<uint32_t>(__popcll(m & ((1ULL << width) - 1ULL)));
        if (n >= low) {
            n -= low;
            m >>= width;
            pos += static_cast<uint32_t>(width);
        }
    }
    return pos;
}

// Live kernel of threshold queries (the learned path; no reference counterpart). Every wave
// scores its share of the live granules (a contiguous run in item order; the waves' shares differ
// by at most one granule), kLiveSlots at a time, from any query. It locates up to
// 64 of its granules at once from a window of 64 item masks (one load per lane), then takes them
// into slots. A batch resolves all its (granule, word) pairs together: one gather of the query
// words, one load of two directory entries per long word (or a window of the short list's
// postings from the planning kernel's cursor, with a two-level 16-ary search of its samples as
// the fallback), then it streams the pairs' postings as segments in (granule, query word) order
// into its granules' FP32 accumulators, so every doc's score is the same sequential FP32 sum in
// query-word order as on the plain path. Docs scoring at least the query's threshold (and its
// start bound) are appended to the query's pool; a pool that overflows makes the final kernel
// rerun the query on the plain path. Batching turns the per-granule chain of dependent loads
// into one chain per kLiveSlots granules. Counter rows go to row rowBase + wave. pruneWaves: the
// planning kernel's item waves, a multiple of 256 up to kMaxPruneWaves.
template <int B>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4))) void KLive(IndexView ix, BatchView bt, WorkView wk, int32_t blocks, int32_t rowBase, int32_t pruneWaves)
{
    constexpr int kGpb = B / kGranule;
    constexpr int U = kLiveColumns;
    constexpr int kLevel = 16; // samples per level of the short-list search (16 x 16 >= 4096 / 16)
    static_assert(kLevel * kLevel >= (kLongMin >> kSampleShift), "short lists must fit two levels");
    static_assert(64 % kGpb == 0, "blocks tile the 64 granules of an item");
    __shared__ __align__(16) float acc[kLiveSlots * kGranule];
    float4 *acc4 = reinterpret_cast<float4 *>(acc);
    const uint32_t lane = __lane_id();
    const uint32_t itemsPerQuery = static_cast<uint32_t>((ix.granules + kWave - 1) / kWave);
    const uint32_t items = static_cast<uint32_t>(bt.header[0]) * itemsPerQuery;
    // This wave's share of the live granules: ranks [T * w / W, T * (w + 1) / W) of the T live
    // granules in item order (W waves), located from the planning waves' live counts: lane l
    // sums those of planning waves [l * per, l * per + per) (per = pruneWaves / 64), the lane
    // holding rank T * w / W then finds its planning wave b among its counts, and the window of
    // item masks starts at b's first item.
    uint32_t winItem = 0; // first item of the window of item masks (lane l: item winItem + l)
    uint32_t skip = 0;    // live granules of the window before the next one to take
    uint32_t left = 0;    // granules still to take
    {
        const uint32_t per = static_cast<uint32_t>(pruneWaves) / kWave; // a multiple of 4, at most 64
        const uint4 *own = reinterpret_cast<const uint4 *>(wk.live + lane * per);
        uint4 c4[kMaxPruneWaves / (4 * kWave)];
#pragma unroll
        for (int j = 0; j < kMaxPruneWaves / (4 * kWave); ++j)
            c4[j] = static_cast<uint32_t>(4 * j) < per ? own[j] : make_uint4(0u, 0u, 0u, 0u);
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kMaxPruneWaves / (4 * kWave); ++j)
            sum += c4[j].x + c4[j].y + c4[j].z + c4[j].w;
        uint32_t incl = sum;
        incl = WaveInclusiveSum(incl);
        const uint32_t total = ReadLane(incl, kWave - 1);
        const uint32_t r0 = static_cast<uint32_t>(static_cast<uint64_t>(total) * blockIdx.x / gridDim.x);
        const uint32_t r1 = static_cast<uint32_t>(static_cast<uint64_t>(total) * (blockIdx.x + 1) / gridDim.x);
        left = r1 - r0;
        if (left > 0) {
            const int l = __popcll(__ballot(incl <= r0)); // the lane whose planning waves hold rank r0
            // Walk lane l's counts (read from its registers) to the planning wave b holding r0.
            uint32_t before = ReadLane(incl - sum, l), b = 0;
            bool found = false;
#pragma unroll
            for (int j = 0; j < kMaxPruneWaves / kWave; ++j) {
                const uint4 &v4 = c4[j / 4];
                const uint32_t v = ReadLane(j % 4 == 0 ? v4.x : j % 4 == 1 ? v4.y : j % 4 == 2 ? v4.z : v4.w, l);
                if (!found && static_cast<uint32_t>(j) < per) {
                    if (before + v > r0) {
                        found = true;
                        b = static_cast<uint32_t>(l) * per + static_cast<uint32_t>(j);
                    } else {
                        before += v;
                    }
                }
            }
            skip = r0 - before;
            winItem = static_cast<uint32_t>(static_cast<uint64_t>(items) * b / static_cast<uint32_t>(pruneWaves));
        }
    }
    // Located granules not yet in a slot: lane i < listN holds granule i of the list as
    // item << 6 | bit (items < 2^26: the session bounds granules per Search by 2^32), its
    // query's word range as first word << 7 | words (first word <= 2^24, words <= 64), K and
    // floor (score bits: the threshold, at least the start bound and tau; tau only grows, so a
    // stale read is still a lower bound).
    uint32_t listIB = 0, listWN = 0, listK = 1, listFloor = 1, listN = 0, listAt = 0;
    unsigned long long scanned = 0, scored = 0, emitted = 0, emitting = 0, useful = 0;

    // Per segment of a batch in position-major order (lane r: segment r): first posting, count,
    // slot | word position << 8.
    __shared__ uint32_t sSeg[3][kWave];
    __shared__ uint32_t sHist[kHistBins]; // one slot's appends per histogram bucket
#ifdef BM25MI_PHASES
    // Shader-clock ticks per phase (live masks and slots, words and ranges, streaming, append),
    // measurement builds only.
    unsigned long long phase[4] = {0, 0, 0, 0}, stamp = wall_clock64();
    const unsigned long long waveStart = stamp;
    int cur = 0;
#define BM25MI_LPHASE(next)                                                                                            \
    do {                                                                                                               \
        const unsigned long long now_ = wall_clock64();                                                                \
        phase[cur] += now_ - stamp;                                                                                    \
        stamp = now_;                                                                                                  \
        cur = (next);                                                                                                  \
    } while (0)
#else
#define BM25MI_LPHASE(next)                                                                                            \
    do {                                                                                                               \
    } while (0)
#endif
    int ns = 0; // slots held (lane s < ns: slot s's query, granule, word range, K and floor), some carried over
    // Word offsets fit 32 bits (a batch holds at most 2^24 words).
    uint32_t slotQ = 0, slotG = 0, slotW0 = 0, slotNw = 0, slotK = 1, slotFloor = 1;
    for (;;) {
        BM25MI_LPHASE(0);
        // 1. Up to kLiveSlots granules of the share.
        while (ns < kLiveSlots) {
            if (listAt == listN) {
                if (left == 0)
                    break;
                // Locate the next min(left, 64) granules: load the window of item masks (with
                // each item's query state, in the same round trip), moving it on while it holds
                // fewer than skip + 1 live granules; then lane i finds the granule of rank
                // skip + i in it.
                unsigned long long winMask = 0;
                uint32_t winIncl = 0, winTotal = 0; // inclusive live-granule counts over the lanes
                uint32_t winWN = 0, winK = 1, winFloor = 1;
                for (;;) {
                    if (winItem >= items) {
                        left = 0; // past the last item (the plan counts only granules of the masks)
                        break;
                    }
                    const uint32_t it = winItem + lane;
                    const bool inWin = it < items;
                    const uint32_t qw = inWin ? it / itemsPerQuery : 0u;
                    winMask = inWin ? wk.itemMask[it] : 0ULL;
                    const uint32_t w0 = static_cast<uint32_t>(bt.qoff[qw]), w1 = static_cast<uint32_t>(bt.qoff[qw + 1]);
                    winK = static_cast<uint32_t>(bt.k[qw]);
                    const uint32_t tauW = __hip_atomic_load(wk.tau + qw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    winFloor = max(max(max(__float_as_uint(bt.thr[qw]), bt.tau0[qw]), tauW), 1u);
                    winWN = w0 << 7 | (w1 - w0);
                    winIncl = static_cast<uint32_t>(__popcll(winMask));
                    winIncl = WaveInclusiveSum(winIncl);
                    winTotal = ReadLane(winIncl, kWave - 1);
                    if (skip < winTotal)
                        break;
                    skip -= winTotal;
                    winItem += kWave;
                }
                if (left == 0)
                    break;
                const uint32_t take = min(left, min(static_cast<uint32_t>(kWave), winTotal - skip));
                const uint32_t r = skip + min(lane, take - 1u);
                uint32_t idx = 0; // first window lane whose inclusive count exceeds r
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(winIncl), static_cast<int>(idx) + step - 1));
                    if (v <= r)
                        idx += static_cast<uint32_t>(step);
                }
                const unsigned long long m = __shfl(winMask, static_cast<int>(idx));
                const uint32_t incl = static_cast<uint32_t>(__shfl(static_cast<int>(winIncl), static_cast<int>(idx)));
                const uint32_t bit = NthSetBit(m, r - (incl - static_cast<uint32_t>(__popcll(m))));
                const uint32_t item = winItem + idx;
                listIB = item << 6 | bit;
                listWN = static_cast<uint32_t>(__shfl(static_cast<int>(winWN), static_cast<int>(idx)));
                listK = static_cast<uint32_t>(__shfl(static_cast<int>(winK), static_cast<int>(idx)));
                listFloor = static_cast<uint32_t>(__shfl(static_cast<int>(winFloor), static_cast<int>(idx)));
                // A live block counts once, at its first live granule.
                const uint32_t b0 = bit & ~static_cast<uint32_t>(kGpb - 1);
                scored += static_cast<unsigned long long>(
                    __popcll(__ballot(lane < take && ((m >> b0) & ((1ULL << (bit - b0)) - 1ULL)) == 0)));
                listN = take;
                listAt = 0;
                skip += take;
                left -= take;
            }
            const int i = static_cast<int>(listAt);
            const uint32_t ib = ReadLane(listIB, i), wn = ReadLane(listWN, i);
            const uint32_t kl = ReadLane(listK, i), fl = ReadLane(listFloor, i);
            const uint32_t qi = (ib >> 6) / itemsPerQuery;
            if (static_cast<int>(lane) == ns) {
                slotQ = qi;
                slotG = ((ib >> 6) - qi * itemsPerQuery) * kWave + (ib & 63u);
                slotW0 = wn >> 7;
                slotNw = wn & 127u;
                slotK = kl;
                slotFloor = fl;
            }
            ++ns;
            ++listAt;
        }
        if (ns == 0)
            break;

        BM25MI_LPHASE(1);
        // 2. The slots' query words (K and floor came with the granules).
        const bool isSlot = static_cast<int>(lane) < ns;
        const int64_t w0 = isSlot ? static_cast<int64_t>(slotW0) : 0;
        const uint32_t nw = isSlot ? slotNw : 0u;
        const uint32_t floorS = isSlot ? slotFloor : 1u, kS = isSlot ? slotK : 1u;
        uint32_t pairEnd = nw; // inclusive prefix of the slots' word counts
        pairEnd = WaveInclusiveSum(pairEnd);
        // The leading m slots whose (granule, word) pairs fit the 64 lanes (m >= 1: a query has
        // at most 64 words); the others carry over to the next batch.
        const int m = __popcll(__ballot(isSlot && pairEnd <= static_cast<uint32_t>(kWave)));
        const uint32_t pairs = ReadLane(pairEnd, m - 1);
        const uint32_t pairStart = pairEnd - nw;
        static_assert((kLiveSlots * kGranule) % (4 * kWave) == 0, "live accumulator size");
#pragma unroll
        for (int r = 0; r < kLiveSlots * kGranule / (4 * kWave); ++r)
            acc4[r * kWave + lane] = make_float4(0.0F, 0.0F, 0.0F, 0.0F);

        // 3. Lane p < pairs: pair p = (slot s, word j), slot-major. Its postings in the slot's
        // granule: two directory entries of a long list; a short list's range from its samples.
        {
            const uint32_t idx = lane;
            const bool has = idx < pairs;
            int s = 0;
#pragma unroll
            for (int r = 1; r < kLiveSlots; ++r)
                s += r < m && ReadLane(pairStart, r) <= idx ? 1 : 0;
            const uint32_t j = idx - static_cast<uint32_t>(__shfl(static_cast<int>(pairStart), s));
            const int64_t wBase = __shfl(w0, s);
            const uint32_t g = static_cast<uint32_t>(__shfl(static_cast<int>(slotG), s));
            uint4 wd = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
            uint32_t cur = 0; // a short word's first posting in the granule's block (planning kernel)
            if (has) {
                wd = bt.words[wBase + j];
                cur = wk.cursor[(wBase + j) * blocks + g / kGpb];
            }
            uint32_t sLo = 0, sLen = 0;
            const bool isLongW = has && wd.y > 0 && static_cast<int32_t>(wd.z) >= 0;
            bool isShortW = has && wd.y > 0 && static_cast<int32_t>(wd.z) < 0;
            if (isLongW) {
                const uint32_t *row = ix.dir + static_cast<int64_t>(static_cast<int32_t>(wd.z)) * (ix.granules + 1) + g;
                const uint32_t d0 = row[0], d1 = row[1];
                sLo = wd.x + d0;
                sLen = d1 - d0;
            }
            const unsigned long long shortPairs = __ballot(isShortW);
            if (shortPairs) {
                // The kWin postings from the cursor, loaded kWin lanes per pair (coalesced), kPer
                // pairs per round: round r, lane l holds posting l % kWin of pair r * kPer + l / kWin.
                // They start at the block's first doc; when they reach the granule's end or the
                // list's end, the granule's postings are exactly those among them in [dlo, dhi).
                constexpr int kWin = 8, kPer = kWave / kWin;
                static_assert(kWin * kPer == kWave, "kPer pairs of kWin lanes per round");
                const uint32_t gDlo = g << kGranuleShift;
                const uint32_t gDhi = min(gDlo + static_cast<uint32_t>(kGranule), static_cast<uint32_t>(ix.ndocs));
                const uint32_t wEnd = wd.x + wd.y;
                constexpr int kFly = 4; // rounds in flight (32 pairs per memory round trip)
                const int rounds = static_cast<int>((pairs + kPer - 1) / kPer);
                uint32_t na = 0, nb = 0;
                bool full = false;
                for (int r0 = 0; r0 < rounds; r0 += kFly) {
                    uint32_t dw[kFly];
#pragma unroll
                    for (int u = 0; u < kFly; ++u) {
                        const int pr = (r0 + u) * kPer + static_cast<int>(lane) / kWin;
                        const uint32_t idx = static_cast<uint32_t>(__shfl(static_cast<int>(cur), pr & (kWave - 1))) + lane % kWin;
                        const uint32_t pe = static_cast<uint32_t>(__shfl(static_cast<int>(wEnd), pr & (kWave - 1)));
                        dw[u] = 0xFFFFFFFFu;
                        if (pr < kWave && ((shortPairs >> pr) & 1ULL) && idx < pe)
                            dw[u] = ix.post[idx].x;
                    }
#pragma unroll
                    for (int u = 0; u < kFly; ++u) {
                        const int r = r0 + u;
                        const int pr = r * kPer + static_cast<int>(lane) / kWin;
                        const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(gDlo), pr & (kWave - 1)));
                        const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(gDhi), pr & (kWave - 1)));
                        const unsigned long long ba = __ballot(dw[u] < lo), bb = __ballot(dw[u] < hi);
                        const unsigned long long bl = __ballot(lane % kWin == kWin - 1 && dw[u] < hi);
                        if (static_cast<int>(lane) / kPer == r) {
                            const int sh = static_cast<int>(lane % kPer) * kWin;
                            na = static_cast<uint32_t>(__popcll((ba >> sh) & ((1ULL << kWin) - 1ULL)));
                            nb = static_cast<uint32_t>(__popcll((bb >> sh) & ((1ULL << kWin) - 1ULL)));
                            full = (bl >> (sh + kWin - 1)) & 1ULL;
                        }
                    }
                }
                if (isShortW && !full) {
                    sLo = cur + na;
                    sLen = nb - na;
                    isShortW = false; // resolved
                }
            }
            if (__ballot(isShortW)) {
                // A window that ends below the granule's end (dense lists, small corpora): the
                // short list's postings of docs [dlo, dhi): count its samples (doc of every
                // 16th posting, global index) below dlo and dhi with a two-level 16-ary search,
                // then the postings below each bound among the at most 15 between its bracketing
                // samples.
                const int32_t dlo = static_cast<int32_t>(g << kGranuleShift);
                const int32_t dhi = min(dlo + kGranule, ix.ndocs);
                const uint32_t i0 = (wd.x + 15u) >> kSampleShift, i1 = (wd.x + wd.y + 15u) >> kSampleShift;
                const uint32_t S = isShortW ? i1 - i0 : 0u;
                uint32_t ca = 0, cb = 0;
#pragma unroll
                for (int mm = 0; mm < kLevel; ++mm) {
                    const uint32_t at = static_cast<uint32_t>(mm) * kLevel;
                    const int32_t v = at < S ? ix.samples[i0 + at] : INT32_MAX;
                    ca += v < dlo ? 1u : 0u;
                    cb += v < dhi ? 1u : 0u;
                }
                const uint32_t ba = ca ? (ca - 1) * kLevel : 0u, bb = cb ? (cb - 1) * kLevel : 0u;
                uint32_t ca2 = 0, cb2 = 0;
#pragma unroll
                for (int mm = 0; mm < kLevel; ++mm) {
                    const uint32_t ia = ba + mm, ib = bb + mm;
                    const int32_t va = ca && ia < S ? ix.samples[i0 + ia] : INT32_MAX;
                    const int32_t vb = cb && ib < S ? ix.samples[i0 + ib] : INT32_MAX;
                    ca2 += va < dlo ? 1u : 0u;
                    cb2 += vb < dhi ? 1u : 0u;
                }
                ca = ca ? ba + ca2 : 0u; // samples of the list with doc < dlo
                cb = cb ? bb + cb2 : 0u; // ... < dhi
                // Postings before as are below dlo, postings from ae on are not (likewise bs, be
                // for dhi); at most 15 lie between.
                const uint32_t listEnd = wd.x + wd.y;
                const uint32_t as = ca > 0 ? ((i0 + ca - 1) << kSampleShift) + 1 : wd.x;
                const uint32_t ae = min(ca < S ? (i0 + ca) << kSampleShift : listEnd, listEnd);
                const uint32_t bs = cb > 0 ? ((i0 + cb - 1) << kSampleShift) + 1 : wd.x;
                const uint32_t be = min(cb < S ? (i0 + cb) << kSampleShift : listEnd, listEnd);
                uint32_t na = 0, nb = 0;
#pragma unroll
                for (int mm = 0; mm < (1 << kSampleShift) - 1; ++mm) {
                    const uint32_t ia = as + mm, ib = bs + mm;
                    const int32_t da = isShortW && ia < ae ? static_cast<int32_t>(ix.post[ia].x) : INT32_MAX;
                    const int32_t db = isShortW && ib < be ? static_cast<int32_t>(ix.post[ib].x) : INT32_MAX;
                    na += da < dlo ? 1u : 0u;
                    nb += db < dhi ? 1u : 0u;
                }
                if (isShortW) {
                    sLo = as + na;
                    const uint32_t hiIdx = bs + nb;
                    sLen = hiIdx > sLo ? hiIdx - sLo : 0u;
                }
            }
            // Position-major rank: pairs ordered by (word position j, slot s).
            uint32_t below = 0, before = 0;
#pragma unroll
            for (int r = 0; r < kLiveSlots; ++r) {
                const uint32_t nr = r < m ? ReadLane(nw, r) : 0u;
                below += min(nr, j);
                before += r < s && nr > j ? 1u : 0u;
            }
            if (has) {
                sSeg[0][below + before] = sLo;
                sSeg[1][below + before] = sLen;
                sSeg[2][below + before] = static_cast<uint32_t>(s) | (j << 8);
            }
        }
        // Lane r < pairs: segment r in position-major order. Lane p < groups: word position p's
        // first segment, segments, postings; its columns start at gColStart.
        const bool hasSeg = lane < pairs;
        const uint32_t rLo = hasSeg ? sSeg[0][lane] : 0u, rLen = hasSeg ? sSeg[1][lane] : 0u;
        const uint32_t rTag = hasSeg ? sSeg[2][lane] : 0u;
        const uint32_t groups = WaveMax(static_cast<int>(lane) < m ? nw : 0u);
        uint32_t gFirst = 0, gCount = 0;
#pragma unroll
        for (int r = 0; r < kLiveSlots; ++r) {
            const uint32_t nr = r < m ? ReadLane(nw, r) : 0u;
            gFirst += min(nr, lane);
            gCount += nr > lane ? 1u : 0u;
        }
        uint32_t gTotal = 0;
#pragma unroll
        for (int t = 0; t < kLiveSlots; ++t) {
            const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(rLen), static_cast<int>(min(gFirst + t, 63u))));
            gTotal += static_cast<uint32_t>(t) < gCount ? v : 0u;
        }
        const uint32_t gCols = lane < groups ? (gTotal + kWave - 1) / kWave : 0u;
        uint32_t gColEnd = gCols;
        gColEnd = WaveInclusiveSum(gColEnd);
        const uint32_t columns = ReadLane(gColEnd, kWave - 1);
        const uint32_t gColStart = gColEnd - gCols;
        if (lane >= groups)
            gColEnd = 0xFFFFFFFFu;
        // Segment r's first posting position within its word position's group.
        const uint32_t rFirst = static_cast<uint32_t>(__shfl(static_cast<int>(gFirst), static_cast<int>(rTag >> 8)));
        uint32_t rFlat = 0;
#pragma unroll
        for (int t = 0; t < kLiveSlots - 1; ++t) {
            const uint32_t v = static_cast<uint32_t>(__shfl(static_cast<int>(rLen), static_cast<int>(min(rFirst + t, 63u))));
            rFlat += rFirst + t < lane ? v : 0u;
        }
        uint32_t totalLen = WaveSum(rLen);
        scanned += totalLen;

        BM25MI_LPHASE(2);
        // 4. Stream the columns, U per round. A column lies within one word position, where the
        // segments are of different granules and hold distinct docs; positions are added in
        // order, so every doc's adds follow its query's word order.
        uint32_t in = 0;
        for (uint32_t r0 = 0; r0 < columns; r0 += U) {
            uint2 e[U];
            uint32_t cb[U], co[U];
            int grp[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                e[u] = make_uint2(0xFFFFFFFFu, 0u);
                cb[u] = 0;
                co[u] = 0;
                grp[u] = -1;
                const uint32_t c = r0 + u;
                if (c < columns) {
                    const int jg = __popcll(__ballot(gColEnd <= c));
                    grp[u] = jg;
                    const uint32_t f = ((c - ReadLane(gColStart, jg)) << 6) + lane;
                    const uint32_t first = ReadLane(gFirst, jg), cnt = ReadLane(gCount, jg);
                    uint32_t r = first;
#pragma unroll
                    for (int t = 1; t < kLiveSlots; ++t)
                        if (static_cast<uint32_t>(t) < cnt && ReadLane(rFlat, static_cast<int>(min(first + t, 63u))) <= f)
                            r = first + t;
                    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(rLo), static_cast<int>(r)));
                    const uint32_t fl = static_cast<uint32_t>(__shfl(static_cast<int>(rFlat), static_cast<int>(r)));
                    const uint32_t sl = static_cast<uint32_t>(__shfl(static_cast<int>(rTag), static_cast<int>(r))) & 0xFFu;
                    cb[u] = static_cast<uint32_t>(__shfl(static_cast<int>(slotG), static_cast<int>(sl))) << kGranuleShift;
                    co[u] = sl << kGranuleShift;
                    if (f < ReadLane(gTotal, jg))
                        e[u] = ix.post[lo + f - fl];
                }
            }
            const uint32_t cLast = min(r0 + U, columns) - 1 - r0;
            if (grp[0] == __builtin_amdgcn_readfirstlane(grp[cLast < U ? cLast : 0])) {
                // One word position: distinct docs, all reads first.
                float v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t rel = e[u].x - cb[u];
                    v[u] = rel < static_cast<uint32_t>(kGranule) ? acc[co[u] + rel] : 0.0F;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t rel = e[u].x - cb[u];
                    if (rel < static_cast<uint32_t>(kGranule)) {
                        acc[co[u] + rel] = v[u] + __uint_as_float(e[u].y);
                        ++in;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t rel = e[u].x - cb[u];
                    if (rel < static_cast<uint32_t>(kGranule)) {
                        acc[co[u] + rel] += __uint_as_float(e[u].y);
                        ++in;
                    }
                }
            }
        }
        useful += in;

        BM25MI_LPHASE(3);
        // 5. Per slot: append the docs scoring at least its floor to its query's pool, at most
        // min(K, kGranule) of them: past the granule's K-th key a doc has K docs of its query
        // above it, so the cut keeps exactly the granule's top K and raises the query's bound tau
        // (a threshold query's pool holds min(K, kGranule) entries per granule). Lane l holds the
        // slot's entries l * kVpl .. l * kVpl + kVpl - 1.
        constexpr int kVpl = kGranule / kWave;
        static_assert(kVpl == 2 || kVpl == 4, "values per lane of a live accumulator");
        const auto *accBits = reinterpret_cast<const uint32_t *>(acc);
        for (int r = 0; r < m; ++r) {
            const uint32_t fl = ReadLane(floorS, r);
            uint32_t vals[kVpl];
            if constexpr (kVpl == 2) {
                const uint2 v2 = reinterpret_cast<const uint2 *>(accBits + r * kGranule)[lane];
                vals[0] = v2.x;
                vals[1] = v2.y;
            } else {
                const uint4 v4 = reinterpret_cast<const uint4 *>(accBits + r * kGranule)[lane];
                vals[0] = v4.x;
                vals[1] = v4.y;
                vals[2] = v4.z;
                vals[3] = v4.w;
            }
            uint32_t mine = 0, mx = 0;
#pragma unroll
            for (int c = 0; c < kVpl; ++c) {
                mine += vals[c] >= fl ? 1u : 0u;
                mx = max(mx, vals[c]);
            }
            uint32_t kept = WaveSum(mine);
            if (kept == 0)
                continue;
            const uint32_t q = ReadLane(slotQ, r);
            const uint32_t kq = ReadLane(kS, r);
            const uint32_t capG = min(kq, static_cast<uint32_t>(kGranule));
            uint32_t tb = fl, te = 0;
            if (kept > capG) {
                const CutResult cr = SelectCut<kGranule>((LdsU32 *)(accBits + r * kGranule), fl, kq, capG, mx);
                tb = cr.tb;
                te = cr.te;
                kept = cr.kept;
                mine = cr.mine;
                if (lane == 0 && cr.tauNew)
                    atomicMax(wk.tau + q, cr.tauNew);
            }
            const uint32_t doc0 = (ReadLane(slotG, r) << kGranuleShift) + lane * kVpl;
            uint32_t slot0 = 0;
            if (lane == 0)
                slot0 = atomicAdd(wk.count + q, kept);
            uint32_t before = mine;
            before = WaveInclusiveSum(before);
            uint32_t slot = ReadLane(slot0, 0) + before - mine;
            unsigned long long *pool = wk.pool + bt.poolOff[q];
            const int64_t capq = bt.poolOff[q + 1] - bt.poolOff[q];
#pragma unroll
            for (int i = 0; i < kHistPerLane; ++i)
                sHist[lane * kHistPerLane + i] = 0;
            // Lanes count into each other's bins (see the scoring kernel): no forwarding of a
            // lane's own zero to its read below.
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
            for (int c = 0; c < kVpl; ++c)
                if (vals[c] > tb ||
                    (vals[c] == tb && static_cast<uint32_t>(kGranule - 1) - (lane * kVpl + c) >= te)) {
                    if (static_cast<int64_t>(slot) < capq)
                        pool[slot] = (static_cast<unsigned long long>(vals[c]) << 32) |
                                     static_cast<unsigned long long>(0xFFFFFFFFu - (doc0 + c));
                    ++slot;
                    atomicAdd(sHist + HistBucket(vals[c]), 1u);
                }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // The query's score histogram, from which the final kernel bounds its K-th score.
            uint32_t *histw = wk.hist + static_cast<int64_t>(q) * kHistBins;
#pragma unroll
            for (int i = 0; i < kHistPerLane; ++i) {
                const uint32_t bin = lane * kHistPerLane + i;
                const uint32_t cnt = sHist[bin];
                if (cnt)
                    atomicAdd(histw + bin, cnt);
            }
            emitted += kept;
            ++emitting;
        }
        // Slots [m, ns) carry over to the next batch.
        slotQ = static_cast<uint32_t>(__shfl(static_cast<int>(slotQ), static_cast<int>(min(lane + m, 63u))));
        slotG = static_cast<uint32_t>(__shfl(static_cast<int>(slotG), static_cast<int>(min(lane + m, 63u))));
        slotW0 = static_cast<uint32_t>(__shfl(static_cast<int>(slotW0), static_cast<int>(min(lane + m, 63u))));
        slotNw = static_cast<uint32_t>(__shfl(static_cast<int>(slotNw), static_cast<int>(min(lane + m, 63u))));
        slotK = static_cast<uint32_t>(__shfl(static_cast<int>(slotK), static_cast<int>(min(lane + m, 63u))));
        slotFloor = static_cast<uint32_t>(__shfl(static_cast<int>(slotFloor), static_cast<int>(min(lane + m, 63u))));
        ns -= m;
    }
    // This wave's counter row: rowBase + blockIdx.x (rowBase = gridDim.x when the plain kernel's
    // waves write rows [0, gridDim.x) of the same Search).
    useful = WaveSum64(useful);
    if (lane == 0) {
        const int64_t at = static_cast<int64_t>(rowBase) + blockIdx.x;
        unsigned long long *row = wk.stats + at * kStats;
        for (int i = 0; i < kStats; ++i)
            row[i] = 0;
        row[kStatPostings] = scanned;
        row[kStatScored] = scored;
        // The planning kernel's dead-block counts (its grid differs from this one).
        unsigned long long deadBlocks = 0;
        for (int32_t b = static_cast<int32_t>(blockIdx.x); b < pruneWaves; b += static_cast<int32_t>(gridDim.x))
            deadBlocks += wk.dead[b];
        row[kStatSkipped] = deadBlocks;
        row[kStatCandidates] = emitted;
        row[kStatEmitting] = emitting;
        row[kStatUseful] = useful;
#ifdef BM25MI_PHASES
        BM25MI_LPHASE(0);
        for (int i = 0; i < 4; ++i)
            row[kStatPhase + i] = phase[i];
        row[kStatMax] = phase[0] + phase[1] + phase[2] + phase[3];
        row[kStatMax + 1] = stamp;
        row[kStatMax + 2] = ~waveStart;
        row[kStatMax + 3] = waveStart;
#endif
    }
#undef BM25MI_LPHASE
}

constexpr int kStageThreads = 256;
constexpr int kStageBatch = 8; // 16-byte loads in flight per thread

// Staging of a Search's request packet (no reference counterpart): the workgroups copy the
// packet's ranges in use from pinned, device-mapped, coherent host memory to the packet's device
// copy, in stream order before the kernels that read it. Unit u of the ranges' concatenation goes
// to thread u % (threads of the grid); each thread issues kStageBatch loads before its stores, so
// a packet of up to threads * kStageBatch units costs one host-memory round trip. For learned
// batches (predict = the queries to predict, at least the batch's; the host packet's thresholds
// are not staged) wave w of workgroup b also predicts the threshold of query 4b + w
// (ThresholdPredictor, reading the query's inputs from the host packet, its loads in flight
// with the copy's) and writes it into the device packet when the query is in the batch.
__global__ __launch_bounds__(kStageThreads) void KStage(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                        StageRanges r, BatchView host, float *thrOut, ModelView mv,
                                                        int32_t predict)
{
    const int q = static_cast<int>(blockIdx.x * (kStageThreads / kWave) + (threadIdx.x >> 6));
    const bool predicting = q < predict;
    ThresholdPredictor tp;
    int32_t nq = 0;
    if (predicting) {
        nq = host.header[0];
        tp.Load(mv, host, q);
    }
    uint32_t total = 0;
    for (int k = 0; k < r.n; ++k)
        total += r.len[k];
    const uint32_t threads = gridDim.x * kStageThreads, me = blockIdx.x * kStageThreads + threadIdx.x;
    for (uint32_t u0 = 0; u0 < total; u0 += threads * kStageBatch) {
        uint4 v[kStageBatch];
        uint32_t at[kStageBatch];
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j) {
            const uint32_t u = u0 + j * threads + me;
            at[j] = 0xFFFFFFFFu;
            if (u < total) {
                uint32_t rest = u;
                int k = 0;
                while (rest >= r.len[k]) {
                    rest -= r.len[k];
                    ++k;
                }
                at[j] = r.off[k] + rest;
                v[j] = src[at[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j)
            if (at[j] != 0xFFFFFFFFu)
                dst[at[j]] = v[j];
    }
    if (!predicting)
        return;
    const float thr = tp.Compute(mv);
    if ((threadIdx.x & 63) == 0 && q < nq)
        thrOut[q] = thr;
}

void LaunchStage(const uint4 *src, uint4 *dst, const StageRanges &ranges, const BatchView &host, float *thrOut,
                 const ModelView &mv, int32_t predictQueries, hipStream_t stream)
{
    const int grid = predictQueries > 0 ? (predictQueries + kStageThreads / kWave - 1) / (kStageThreads / kWave) : 1;
    hipLaunchKernelGGL(KStage, dim3(grid), dim3(kStageThreads), 0, stream, src, dst, ranges, host, thrOut, mv,
                       predictQueries);
}

// Initial state of the per-query counters; run once at session creation. After that
// the scoring and final kernels restore it for the next Search.
__global__ void KReset(WorkView wk, int32_t maxQ)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < maxQ) {
        wk.count[i] = 0;
        wk.tau[i] = 1;
    }

    if (i < static_cast<int64_t>(maxQ) * kHistBins)
        wk.hist[i] = 0;
}

__device__ __forceinline__ unsigned long long Shfl64(unsigned long long v, uint32_t src)
{
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(src)));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(src)));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long Shfl64Xor(unsigned long long v, uint32_t mask)
{
    return Shfl64(v, __lane_id() ^ mask);
}

// One key per lane sorted across the wave, descending (lane 0 the largest): a bitonic sort
// through lane exchanges. Every lane active.
__device__ __forceinline__ unsigned long long WaveSortDesc(unsigned long long v)
{
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t size = 2; size <= kWave; size <<= 1)
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            const unsigned long long o = Shfl64Xor(v, stride);
            const bool lower = (lane & stride) == 0, desc = (lane & size) == 0;
            v = (lower == desc) == (o > v) ? o : v; // the larger where lower == desc, else the smaller
        }
    return v;
}

// The 64 largest of two descending wave-sorted lists, descending: max(a[l], b[63 - l]) is a
// bitonic sequence holding them, then a bitonic merge.
__device__ __forceinline__ unsigned long long WaveMergeTop(unsigned long long a, unsigned long long b)
{
    const uint32_t lane = __lane_id();
    const unsigned long long r = Shfl64(b, kWave - 1 - lane);
    unsigned long long v = a > r ? a : r;
#pragma unroll
    for (uint32_t stride = kWave / 2; stride > 0; stride >>= 1) {
        const unsigned long long o = Shfl64Xor(v, stride);
        v = ((lane & stride) == 0) == (o > v) ? o : v;
    }
    return v;
}

// The k-th largest of p unique keys in global memory (p > k), by a 64-step bisection
// with block-wide counts. All threads call; every thread gets the result.
__device__ unsigned long long KthFromGlobal(const unsigned long long *keys, uint32_t p, uint32_t k, uint32_t *red)
{
    unsigned long long kappa = 0;
    for (int bit = 63; bit >= 0; --bit) {
        const unsigned long long cand = kappa | (1ULL << bit);
        uint32_t c = 0;
        for (uint32_t i = threadIdx.x; i < p; i += kFinalThreads)
            c += keys[i] >= cand ? 1u : 0u;
        if (BlockSum<kFinalThreads>(c, red) >= k)
            kappa = cand;
    }
    return kappa;
}

__device__ __forceinline__ uint32_t ShflU32(uint32_t v, uint32_t src)
{
    return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(src)));
}

// Visits the candidates of a plain query, gathered from the per-task regions of its pool: block
// b's region holds counts[b] keys from entry b * cap. Wave w of the workgroup takes the groups of
// 64 blocks w, w + 16, ..., kGatherGroups at a time: their counts are loaded together, then each
// lane finds the block of its entry by a binary search over the group's count prefix (six lane
// reads) and every group's first two rounds of 64 entries are loaded before any is visited
// (further rounds of a group with more than 128 entries follow). visit(key, valid) is called by
// every lane of the wave, also for lanes without an entry (valid false).
constexpr int kGatherGroups = 5;
// The final kernel's gather of a plain query: blocks per window (8 per thread; their counts and
// the threads' prefix in LDS) and entries per thread per pass.
constexpr int kGatherWin = 8 * kFinalThreads;
constexpr int kGatherPerThread = 8;
constexpr uint32_t kNoEntry = 0xFFFFFFFFu; // rounds of 64 entries per group the final kernel loads at once
template <class Visit>
__device__ __forceinline__ void GatherBlocks(const unsigned long long *pool, const uint32_t *counts, int32_t blocks,
                                             uint32_t cap, Visit &&visit)
{
    constexpr int kWaves = kFinalThreads / kWave;
    const int wave = static_cast<int>(threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const int groups = (blocks + kWave - 1) / kWave;
    // The key of entry i of a group (incl: the group's inclusive count prefix, lane l holding
    // block g * 64 + l), loaded when i < total. The block is found by a binary search over the
    // prefix with the whole wave active (i clamped into the group), so that every lane read
    // reads an active lane.
    auto entry = [&](uint32_t incl, uint32_t c, int g, uint32_t total, uint32_t i) -> unsigned long long {
        const uint32_t ic = min(i, total - 1u);
        uint32_t pos = 0; // lanes whose inclusive count prefix is <= ic
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1)
            pos += ShflU32(incl, pos + step - 1) <= ic ? step : 0u;
        const uint32_t start = ShflU32(incl - c, pos);
        const int64_t b = static_cast<int64_t>(g) * kWave + pos;
        return i < total ? pool[b * cap + (ic - start)] : 0ULL;
    };
    for (int g0 = wave; g0 < groups; g0 += kWaves * kGatherGroups) {
        uint32_t c[kGatherGroups], incl[kGatherGroups], total[kGatherGroups];
#pragma unroll
        for (int j = 0; j < kGatherGroups; ++j) {
            const int b = (g0 + j * kWaves) * kWave + static_cast<int>(lane);
            c[j] = b < blocks ? counts[b] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kGatherGroups; ++j) {
            incl[j] = WaveInclusiveSum(c[j]);
            total[j] = ReadLane(incl[j], kWave - 1);
        }
        unsigned long long key[kGatherGroups][2];
#pragma unroll
        for (int j = 0; j < kGatherGroups; ++j)
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                key[j][r] = 0ULL;
                if (static_cast<uint32_t>(r * kWave) < total[j]) // uniform
                    key[j][r] = entry(incl[j], c[j], g0 + j * kWaves, total[j], r * kWave + lane);
            }
#pragma unroll
        for (int j = 0; j < kGatherGroups; ++j)
#pragma unroll
            for (int r = 0; r < 2; ++r)
                visit(key[j][r], r * kWave + lane < total[j]);
#pragma unroll
        for (int j = 0; j < kGatherGroups; ++j)
            for (uint32_t r0 = 2 * kWave; r0 < total[j]; r0 += kWave) {
                const unsigned long long v = entry(incl[j], c[j], g0 + j * kWaves, total[j], r0 + lane);
                visit(v, r0 + lane < total[j]);
            }
    }
}

// Per query: the exact top-K of the candidate pool. Entries scoring below the query's
// final lower bound tau cannot be in the top-K and are dropped; the rest are sorted in
// LDS (score desc, doc asc) when they fit, else reduced by a 64-step bisection first.
// Fewer than K positive docs are padded with the lowest-ID zero-score docs. A plain query's
// candidates are gathered from its tasks' pool regions (blocks, blockDocs and keepCap give
// their size, BlockKeepCap); a threshold query's pool is filled from entry 0.
__global__ __launch_bounds__(kFinalThreads) void KFinal(BatchView bt, WorkView wk, unsigned long long *hostStats,
                                                         int32_t scoreGrid, int32_t blocks, int32_t blockDocs,
                                                         int32_t keepCap)
{
    __shared__ unsigned long long sKeys[kSortCap];
    __shared__ unsigned long long sIds[1024];
    __shared__ uint32_t sRed[kFinalThreads / 64];
    __shared__ uint32_t sN;
    // Plain query, per window of kGatherWin blocks: the prefix of each thread's 8 block counts,
    // and the counts (16-bit, 8 per thread).
    __shared__ uint32_t sGrp[kFinalThreads];
    __shared__ __align__(16) uint16_t sCnt[kGatherWin];
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    // The query's state, loaded with the query count and before the counter rows so that all
    // these loads share one memory round trip (wave 0 also loads the score histogram). The
    // grid holds at most the session's query capacity, so the loads stay within the packet and
    // the per-query arrays also for workgroups beyond the batch's queries, which ignore them.
    const int32_t nq = bt.header[0];
    uint32_t hregs[kHistPerLane];
    uint32_t k = static_cast<uint32_t>(bt.k[q]);
    const int64_t poolLo = bt.poolOff[q], poolHi = bt.poolOff[q + 1];
    const uint32_t countQ = wk.count[q], tauQ = wk.tau[q], tau0Q = bt.tau0[q];
    const float thrQ = bt.thr[q];
#pragma unroll
    for (int i = 0; i < kHistPerLane; ++i)
        hregs[i] = tid < kWave ? wk.hist[static_cast<int64_t>(q) * kHistBins + tid * kHistPerLane + i] : 0u;
    const bool active = q < nq;
    // A plain query's block counts of the gather's first window (thread t: blocks 8t .. 8t+7 of
    // the window), loaded early. The count row is padded to a multiple of 8 blocks, so that the
    // two 16-byte loads stay within it; padding entries count 0 (blocks past the query's).
    const uint32_t *counts = wk.blockCount + static_cast<int64_t>(q) * BlockCountStride(blocks);
    const bool plainActive = active && !(thrQ > 0.0F);
    auto loadCounts = [&](int w0, uint4 (&c8)[2]) {
        const int b = w0 + 8 * tid;
        const uint4 *p = reinterpret_cast<const uint4 *>(counts + b);
        c8[0] = c8[1] = make_uint4(0u, 0u, 0u, 0u);
        if (plainActive && b < blocks) {
            c8[0] = p[0];
            c8[1] = p[1];
            // Entries past the last block (row padding) are not written by the scoring kernel.
            const int n = blocks - b; // >= 1
            c8[0].y = n > 1 ? c8[0].y : 0u;
            c8[0].z = n > 2 ? c8[0].z : 0u;
            c8[0].w = n > 3 ? c8[0].w : 0u;
            c8[1].x = n > 4 ? c8[1].x : 0u;
            c8[1].y = n > 5 ? c8[1].y : 0u;
            c8[1].z = n > 6 ? c8[1].z : 0u;
            c8[1].w = n > 7 ? c8[1].w : 0u;
        }
    };
    uint4 cg0[2];
    bool prefetched = false;
    auto prefetchCounts = [&]() {
        prefetched = true;
        loadCounts(0, cg0);
    };
    {
        // The scoring kernel has finished: this workgroup sums its share of the per-wave counter
        // rows (counters 12..15 take the maximum) into its row of hostStats, which the host adds
        // up. The rows are read as contiguous words (thread t: words t, t + 1024, ... of the
        // share, all of counter t % 16), eight per thread in flight; a plain query's first block
        // counts are loaded in the same round trip.
        static_assert(kFinalThreads % kStats == 0 && kStats == 16, "a thread sums one counter");
        const int stat = tid & (kStats - 1);
        const int per = (scoreGrid + gridDim.x - 1) / gridDim.x;
        const int r0 = q * per, r1 = min(scoreGrid, r0 + per);
        const int64_t w0 = static_cast<int64_t>(r0) * kStats, w1 = static_cast<int64_t>(max(r0, r1)) * kStats;
        unsigned long long acc = 0;
        for (int64_t wb = w0 + tid; wb < w1 || wb == w0 + tid; wb += 8 * kFinalThreads) {
            unsigned long long v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t w = wb + u * kFinalThreads;
                v[u] = w < w1 ? wk.stats[w] : 0ULL;
            }
            if (wb == w0 + tid)
                prefetchCounts();
#pragma unroll
            for (int u = 0; u < 8; ++u)
                acc = stat >= kStatMax ? (v[u] > acc ? v[u] : acc) : acc + v[u];
        }
        // Lanes l, l^16, l^32, l^48 hold counter l % 16; then the 16 waves through LDS.
#pragma unroll
        for (uint32_t m = 16; m < kWave; m <<= 1) {
            const unsigned long long o = Shfl64Xor(acc, m);
            acc = stat >= kStatMax ? (o > acc ? o : acc) : acc + o;
        }
        if ((tid & 63) < kStats)
            sIds[(tid >> 6) * kStats + stat] = acc;
        __syncthreads();
        if (tid < kStats) {
            unsigned long long t = 0;
            for (int w = 0; w < kFinalThreads / kWave; ++w) {
                const unsigned long long o = sIds[w * kStats + tid];
                t = stat >= kStatMax ? (o > t ? o : t) : t + o;
            }
            hostStats[static_cast<int64_t>(q) * kStats + tid] = t;
        }
        __syncthreads(); // sIds is used again below
    }
    if (!active)
        return;
    const int64_t cap = poolHi - poolLo;
    const uint32_t p = static_cast<uint32_t>(min(static_cast<int64_t>(countQ), cap));
    const bool overflowed = static_cast<int64_t>(countQ) > cap; // only threshold queries (rerun)
    const uint32_t blockCap = static_cast<uint32_t>(BlockKeepCap(static_cast<int32_t>(k), blockDocs, keepCap));
    uint32_t *hist = wk.hist + static_cast<int64_t>(q) * kHistBins;
    if (tid < 64) {
        const uint32_t hb = WaveHistBoundRegs(hregs, k);
        if (tid == 0)
            sRed[0] = max(max(max(tauQ, hb), tau0Q), 1u);
    }
    __syncthreads();
    const uint32_t tau = sRed[0];
    const unsigned long long *pool = wk.pool + poolLo;
    const bool pruned = thrQ > 0.0F;
    if (tid == 0)
        sN = 0;
    __syncthreads();
    // Restore the per-query state for the next Search.
    if (tid == 0) {
        wk.count[q] = 0;
        wk.tau[q] = 1;
    }

    for (int i = tid; i < kHistBins; i += kFinalThreads)
        hist[i] = 0;
    // Keep entries with score bits >= tau.
    if (pruned) {
        for (uint32_t i0 = 0; i0 < p; i0 += kFinalThreads) {
            const uint32_t i = i0 + tid;
            const unsigned long long key = i < p ? pool[i] : 0ULL;
            const uint32_t at = WaveAppend(i < p && static_cast<uint32_t>(key >> 32) >= tau, &sN);
            if (at < static_cast<uint32_t>(kSortCap))
                sKeys[at] = key;
        }
    } else {
        // Plain query, in windows of kGatherWin blocks: thread t holds the counts of the window's
        // blocks 8t .. 8t+7 (two 16-byte loads, made with the counter rows' loads for the first
        // window); one workgroup-wide prefix over the threads' totals goes to LDS with the counts
        // (16-bit, 8 per thread); then thread t takes the window's entries t, t + 1024, ...
        // (kGatherPerThread per pass, all loads in flight together), finds each entry's thread
        // by a binary search over the prefix and its block among that thread's 8 counts, and the
        // survivors (score bits >= tau) are placed with one wave-wide prefix.
        if (!prefetched)
            prefetchCounts();
        for (int w0 = 0; w0 < blocks; w0 += kGatherWin) {
            uint4 c8[2];
            if (w0 == 0) {
                c8[0] = cg0[0];
                c8[1] = cg0[1];
            } else {
                loadCounts(w0, c8);
            }
            const uint32_t mineCount = c8[0].x + c8[0].y + c8[0].z + c8[0].w + c8[1].x + c8[1].y + c8[1].z + c8[1].w;
            uint32_t total = 0;
            const uint32_t ex = BlockExclusive<kFinalThreads>(mineCount, sRed, &total);
            sGrp[tid] = ex;
            reinterpret_cast<uint4 *>(sCnt)[tid] =
                make_uint4(c8[0].x | c8[0].y << 16, c8[0].z | c8[0].w << 16, c8[1].x | c8[1].y << 16,
                           c8[1].z | c8[1].w << 16);
            __syncthreads();
            for (uint32_t e0 = 0; e0 < total; e0 += kFinalThreads * kGatherPerThread) {
                // The last thread whose prefix is <= i holds entry i (threads without entries
                // share their successor's prefix), then its block among that thread's 8. The
                // searches of a thread's entries advance together, step by step (index clamped
                // into the window's entries), so that their LDS reads are in flight together.
                uint32_t g[kGatherPerThread], ic[kGatherPerThread];
#pragma unroll
                for (int m = 0; m < kGatherPerThread; ++m) {
                    g[m] = 0;
                    ic[m] = min(e0 + m * kFinalThreads + tid, total - 1u);
                }
#pragma unroll
                for (uint32_t step = kFinalThreads / 2; step > 0; step >>= 1)
#pragma unroll
                    for (int m = 0; m < kGatherPerThread; ++m)
                        g[m] += sGrp[g[m] + step] <= ic[m] ? step : 0u;
                uint32_t at[kGatherPerThread];
#pragma unroll
                for (int m = 0; m < kGatherPerThread; ++m) {
                    uint32_t r = ic[m] - sGrp[g[m]];
                    const uint4 cc = reinterpret_cast<const uint4 *>(sCnt)[g[m]];
                    const uint32_t cnt[8] = {cc.x & 0xFFFFu, cc.x >> 16, cc.y & 0xFFFFu, cc.y >> 16,
                                             cc.z & 0xFFFFu, cc.z >> 16, cc.w & 0xFFFFu, cc.w >> 16};
                    uint32_t blk = 0;
                    bool found = false;
#pragma unroll
                    for (int jb = 0; jb < 8; ++jb) {
                        const bool here = !found && r < cnt[jb];
                        blk = here ? static_cast<uint32_t>(jb) : blk;
                        r = !found && !here ? r - cnt[jb] : r;
                        found = found || here;
                    }
                    at[m] = e0 + m * kFinalThreads + tid < total
                                ? (static_cast<uint32_t>(w0) + 8 * g[m] + blk) * blockCap + r
                                : kNoEntry;
                }
                unsigned long long key[kGatherPerThread];
#pragma unroll
                for (int m = 0; m < kGatherPerThread; ++m)
                    key[m] = at[m] != kNoEntry ? pool[at[m]] : 0ULL;
                uint32_t mine = 0;
#pragma unroll
                for (int m = 0; m < kGatherPerThread; ++m)
                    mine += at[m] != kNoEntry && static_cast<uint32_t>(key[m] >> 32) >= tau ? 1u : 0u;
                const uint32_t inclMine = WaveInclusiveSum(mine);
                uint32_t slot = 0;
                if ((tid & 63) == kWave - 1 && inclMine)
                    slot = atomicAdd(&sN, inclMine);
                slot = ReadLane(slot, kWave - 1) + inclMine - mine;
#pragma unroll
                for (int m = 0; m < kGatherPerThread; ++m)
                    if (at[m] != kNoEntry && static_cast<uint32_t>(key[m] >> 32) >= tau) {
                        if (slot < static_cast<uint32_t>(kSortCap))
                            sKeys[slot] = key[m];
                        ++slot;
                    }
            }
            __syncthreads(); // sGrp and sCnt are rewritten by the next window
        }
    }
    __syncthreads();
    const uint32_t c = sN;
    if (pruned && (c < k || overflowed)) {
        if (tid == 0)
            wk.status[q] = 1; // fewer than K survivors, or a full pool: the host reruns this query
        return;                // on the plain path
    }
    const uint32_t m = c < k ? c : k;
    if (c <= static_cast<uint32_t>(kWave)) {
        // Up to 64 survivors (the usual case at small K): one wave ranks them without barriers.
        // Keys are unique, so a key's position is the number of keys above it, counted against
        // every lane's key read as two scalars (its reads come before its writes in the wave's
        // program order).
        if (tid < kWave) {
            const unsigned long long v = tid < static_cast<int>(c) ? sKeys[tid] : 0ULL;
            const uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
            uint32_t rank = 0;
            for (int j = 0; j < static_cast<int>(c); ++j) {
                const unsigned long long o =
                    (static_cast<unsigned long long>(ReadLane(hi, j)) << 32) | ReadLane(lo, j);
                rank += o > v ? 1u : 0u;
            }
            if (tid < static_cast<int>(c))
                sKeys[rank] = v;
        }
        __syncthreads();
    } else if (c <= static_cast<uint32_t>(kSortCap) && k <= static_cast<uint32_t>(kWave)) {
        // More than 64 survivors, K <= 64: only the top K are needed. Each wave keeps the 64
        // largest keys of its share (every 1024th chunk of 64) sorted in its registers, merging
        // each chunk with a register sort (no workgroup barriers), writes its top K, and wave 0
        // merges the 16 lists the same way.
        const int wave = tid >> 6;
        unsigned long long best = 0;
        for (uint32_t i0 = static_cast<uint32_t>(wave) * kWave; i0 < c; i0 += kFinalThreads) {
            const uint32_t i = i0 + (tid & 63);
            best = WaveMergeTop(best, WaveSortDesc(i < c ? sKeys[i] : 0ULL));
        }
        __syncthreads(); // every wave has read its share of sKeys
        if ((tid & 63) < static_cast<int>(k))
            sIds[wave * k + (tid & 63)] = best;
        __syncthreads();
        // The 16 lists (16K keys, zeros past a wave's share): thread i ranks key i by the keys above
        // it (every thread reads the same key at a time: LDS broadcasts); the top K go to sKeys.
        const uint32_t n = (kFinalThreads / kWave) * k;
        if (static_cast<uint32_t>(tid) < n) {
            const unsigned long long v = sIds[tid];
            uint32_t rank = 0;
            for (uint32_t j = 0; j < n; ++j)
                rank += sIds[j] > v ? 1u : 0u;
            if (rank < k && v != 0ULL)
                sKeys[rank] = v;
        }
        __syncthreads();
    } else if (c <= static_cast<uint32_t>(kSortCap)) {
        int n2 = 1;
        while (n2 < static_cast<int>(c))
            n2 <<= 1;
        for (int i = static_cast<int>(c) + tid; i < n2; i += kFinalThreads)
            sKeys[i] = 0ULL;
        __syncthreads();
        BitonicDesc<kFinalThreads>(sKeys, n2);
    } else {
        // Too many to sort: the K-th key over the whole pool, then the K keys at or above it.
        unsigned long long kappa = 0;
        if (pruned) {
            kappa = KthFromGlobal(pool, p, k, sRed);
        } else {
            // The same 64-step bisection over the gathered regions (rare: more than kSortCap
            // candidates at or above the final bound, i.e. large K with massive score ties).
            for (int bit = 63; bit >= 0; --bit) {
                const unsigned long long cand = kappa | (1ULL << bit);
                uint32_t n = 0;
                GatherBlocks(pool, counts, blocks, blockCap,
                             [&](unsigned long long key, bool valid) { n += valid && key >= cand ? 1u : 0u; });
                if (BlockSum<kFinalThreads>(n, sRed) >= k)
                    kappa = cand;
            }
        }
        if (tid == 0)
            sN = 0;
        __syncthreads();
        if (pruned) {
            for (uint32_t i = tid; i < p; i += kFinalThreads)
                if (pool[i] >= kappa) {
                    const uint32_t at = atomicAdd(&sN, 1u);
                    if (at < k)
                        sIds[at] = pool[i];
                }
        } else {
            GatherBlocks(pool, counts, blocks, blockCap, [&](unsigned long long key, bool valid) {
                if (valid && key >= kappa) {
                    const uint32_t at = atomicAdd(&sN, 1u);
                    if (at < k)
                        sIds[at] = key;
                }
            });
        }
        __syncthreads();
        int n2 = 1;
        while (n2 < static_cast<int>(k))
            n2 <<= 1;
        for (int i = tid; i < n2; i += kFinalThreads)
            sKeys[i] = i < static_cast<int>(k) ? sIds[i] : 0ULL;
        __syncthreads();
        BitonicDesc<kFinalThreads>(sKeys, n2);
    }
    const int64_t row = bt.outOff[q];
    for (uint32_t r = tid; r < m; r += kFinalThreads) {
        const unsigned long long key = sKeys[r];
        wk.scores[row + r] = __uint_as_float(static_cast<uint32_t>(key >> 32));
        wk.ids[row + r] = static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(key & 0xFFFFFFFFu));
    }
    if (tid == 0)
        wk.status[q] = 0;
    if (m < k) {
        // Pad with the lowest-ID docs that scored zero (every positive doc is in the pool):
        // sort the positive IDs ascending (descending on their complements), then walk.
        int n2 = 1;
        while (n2 < static_cast<int>(m))
            n2 <<= 1;
        for (int r = tid; r < n2; r += kFinalThreads)
            sIds[r] = r < static_cast<int>(m)
                          ? ~static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<uint32_t>(sKeys[r] & 0xFFFFFFFFu))
                          : 0ULL;
        __syncthreads();
        BitonicDesc<kFinalThreads>(sIds, n2);
        if (tid == 0) {
            uint32_t j = 0;
            int32_t d = 0;
            for (uint32_t r = m; r < k; ++r) {
                while (j < m && static_cast<int32_t>(~sIds[j]) == d) {
                    ++j;
                    ++d;
                }
                wk.scores[row + r] = 0.0F;
                wk.ids[row + r] = d;
                ++d;
            }
        }
    }
}

void LaunchInit(const WorkView &wk, int32_t maxQ, hipStream_t stream)
{
    const int threads = 256;
    const int grid = static_cast<int>((static_cast<int64_t>(maxQ) * kHistBins +
                                       threads - 1) / threads);
    hipLaunchKernelGGL(KReset, dim3(grid), dim3(threads), 0, stream, wk, maxQ);
}

bool LaunchScore(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t grid, hipStream_t stream, hipEvent_t start, hipEvent_t stop, int32_t addRows, int32_t keepCap,
                 int32_t nqKnown)
{
    // With events, hipExtLaunchKernelGGL stamps them at the kernel's own start and end (the
    // interval the profiler reports), not around the launch.
    switch (blockDocs) {
    case 2048:
        hipExtLaunchKernelGGL(KScore<2048>, dim3(grid), dim3(kWave), 0, stream, start, stop, 0, ix, bt, wk, blocks, addRows,
                              keepCap, nqKnown);
        return true;
    case 4096:
        hipExtLaunchKernelGGL(KScore<4096>, dim3(grid), dim3(kWave), 0, stream, start, stop, 0, ix, bt, wk, blocks, addRows,
                              keepCap, nqKnown);
        return true;
    case 8192:
        hipExtLaunchKernelGGL(KScore<8192>, dim3(grid), dim3(kWave), 0, stream, start, stop, 0, ix, bt, wk, blocks, addRows,
                              keepCap, nqKnown);
        return true;
    default:
        return false;
    }
}

bool LaunchPrune(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                 int32_t waves, int32_t nqKnown, hipStream_t stream)
{
    // waves item waves and the learned batches' header wave, kPruneGroup waves per workgroup.
    const dim3 groups((waves + 1 + kPruneGroup - 1) / kPruneGroup), threads(kPruneGroup * kWave);
    switch (blockDocs) {
    case 2048:
        hipLaunchKernelGGL(KPrune<2048>, groups, threads, 0, stream, ix, bt, wk, blocks, waves, nqKnown);
        return true;
    case 4096:
        hipLaunchKernelGGL(KPrune<4096>, groups, threads, 0, stream, ix, bt, wk, blocks, waves, nqKnown);
        return true;
    case 8192:
        hipLaunchKernelGGL(KPrune<8192>, groups, threads, 0, stream, ix, bt, wk, blocks, waves, nqKnown);
        return true;
    default:
        return false;
    }
}

bool LaunchLive(int blockDocs, const IndexView &ix, const BatchView &bt, const WorkView &wk, int32_t blocks,
                int32_t grid, int32_t rowBase, int32_t pruneWaves, hipStream_t stream)
{
    switch (blockDocs) {
    case 2048:
        hipLaunchKernelGGL(KLive<2048>, dim3(grid), dim3(kWave), 0, stream, ix, bt, wk, blocks, rowBase, pruneWaves);
        return true;
    case 4096:
        hipLaunchKernelGGL(KLive<4096>, dim3(grid), dim3(kWave), 0, stream, ix, bt, wk, blocks, rowBase, pruneWaves);
        return true;
    case 8192:
        hipLaunchKernelGGL(KLive<8192>, dim3(grid), dim3(kWave), 0, stream, ix, bt, wk, blocks, rowBase, pruneWaves);
        return true;
    default:
        return false;
    }
}

int32_t PruneWaves(int64_t items)
{
    // About kPruneItems items per wave, in whole 256-wave steps.
    const int64_t w = (items + kPruneItems - 1) / kPruneItems;
    return static_cast<int32_t>(std::min<int64_t>(kMaxPruneWaves, std::max<int64_t>(256, (w + 255) / 256 * 256)));
}

// Resident planning-kernel workgroups (kPruneGroup waves each) per CU, or 0.
int PruneBlocksPerCu(int blockDocs)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (blockDocs) {
    case 2048:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KPrune<2048>, kPruneGroup * kWave, 0);
        break;
    case 4096:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KPrune<4096>, kPruneGroup * kWave, 0);
        break;
    case 8192:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KPrune<8192>, kPruneGroup * kWave, 0);
        break;
    default:
        break;
    }
    return e == hipSuccess ? n : 0;
}

// Resident live-kernel waves per CU for a block size, or 0.
int LiveBlocksPerCu(int blockDocs)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (blockDocs) {
    case 2048:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KLive<2048>, kWave, 0);
        break;
    case 4096:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KLive<4096>, kWave, 0);
        break;
    case 8192:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KLive<8192>, kWave, 0);
        break;
    default:
        break;
    }
    return e == hipSuccess ? n : 0;
}

// Resident scoring waves per CU for a block size (registers and LDS), or 0.
int ScoreBlocksPerCu(int blockDocs)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (blockDocs) {
    case 2048:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<2048>, kWave, 0);
        break;
    case 4096:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<4096>, kWave, 0);
        break;
    case 8192:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, KScore<8192>, kWave, 0);
        break;
    default:
        break;
    }
    return e == hipSuccess ? n : 0;
}

void LaunchFinal(const BatchView &bt, const WorkView &wk, int32_t workgroups, unsigned long long *hostStats,
                 int32_t scoreGrid, int32_t blocks, int32_t blockDocs, int32_t keepCap, hipStream_t stream)
{
    hipLaunchKernelGGL(KFinal, dim3(workgroups), dim3(kFinalThreads), 0, stream, bt, wk, hostStats, scoreGrid, blocks,
                       blockDocs, keepCap);
}

} // namespace bm25mi
